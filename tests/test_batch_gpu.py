"""Batched solve of many independent meshes (the library's svg_batch_*):
each mesh's result equals a solve of that mesh alone."""
import ctypes as C

import numpy as np
import pytest

from svfsiplus_amd import _lib, batch as B, workload as W
from svfsiplus_amd.newton import DeviceLhs, FluidNewton, fsils_solve, ls_create

pytestmark = pytest.mark.gpu


def _meshes(count=3, tets_max=1.2e4, seed=5):
    return B.make_batch(W.bifurcation_batch(count, seed=seed, tets_max=tets_max))


def _assembled_union(torch, meshes, phys):
    bn = B.BatchNewton(meshes, phys)
    fn = bn.fn
    lib = _lib.load()
    from svfsiplus_amd.newton import _stream, b_assem_neu_bc, construct_fluid
    _lib.check(lib.svg_pici(bn.pipe.nNo, phys.am, phys.af, fn.Ao.data_ptr(), fn.An.data_ptr(), fn.Yo.data_ptr(),
                            fn.Yn.data_ptr(), fn.Ag.data_ptr(), fn.Yg.data_ptr(), _stream()))
    fn.R.zero_()
    fn.Val.zero_()
    construct_fluid(fn.msh, phys, fn.Ag, fn.Yg, fn.R, fn.Val)
    b_assem_neu_bc(fn.outlet, phys, fn.Yg, fn.R, fn.Val)
    torch.cuda.synchronize()
    return bn


@pytest.mark.parametrize("precision", [_lib.PRECISION_FP64, _lib.PRECISION_MIXED])
def test_batch_solve_equals_individual_solves(torch_cuda, precision):
    torch = torch_cuda
    meshes = _meshes()
    phys = W.Physics()
    bn = _assembled_union(torch, meshes, phys)
    R0, V0 = bn.fn.R.clone(), bn.fn.Val.clone()
    # per-mesh systems: the union's rows and blocks of each mesh
    boff = np.concatenate([[0], np.cumsum([m.nnz for m in meshes])])
    singles = []
    for q, m in enumerate(meshes):
        a, b = int(bn.noff[q]), int(bn.noff[q + 1])
        Rq = R0[a:b].clone()
        Vq = V0[boff[q]:boff[q + 1]].clone()
        ls = ls_create(_lib.LS_GMRES, relTol=1e-6, absTol=1e-30, maxItr=1000, dimKry=50)
        fsils_solve(DeviceLhs(m), ls, 4, Rq, Vq, precision=precision)
        singles.append((Rq, ls.RI.itr, ls.RI.iNorm))
    Rb, Vb = R0.clone(), V0.clone()
    _lib.check(_lib.load().svg_batch_fsils_solve(bn.h, C.byref(bn.proto), Rb.data_ptr(), Vb.data_ptr(), precision,
                                                 bn.out, None))
    torch.cuda.synchronize()
    for q, (Rq, itr, iN) in enumerate(singles):
        a, b = int(bn.noff[q]), int(bn.noff[q + 1])
        diff = float((Rb[a:b] - Rq).abs().max())
        print(f"mesh {q}: nNo {b - a} itr batch {bn.out[q].itr} single {itr} iNorm {bn.out[q].iNorm:.6e}/{iN:.6e} "
              f"max|dx| {diff:.2e}")
        assert bn.out[q].itr == itr
        assert bn.out[q].iNorm == iN
        assert diff == 0.0


@pytest.mark.parametrize("device_setup", [False, True])
def test_batch_newton_time_step(torch_cuda, device_setup):
    """Three Newton iterations of every mesh at once against the same iterations
    run on each mesh alone; also with the union's pattern, colouring and slot
    map built on the device."""
    meshes = _meshes(4, seed=9)
    phys = W.Physics()
    bn = B.BatchNewton(meshes, phys, device_setup=device_setup)
    recs = bn.time_step(3)
    for q, m in enumerate(meshes):
        fn = FluidNewton(m, phys, ls_create(_lib.LS_GMRES, relTol=1e-6, absTol=1e-30, maxItr=1000, dimKry=50))
        h = fn.time_step()
        for it in range(3):
            rb, rs = recs[it][q], h[it]
            print(f"mesh {q} newton {it + 1}: batch itr {rb['itr']} iNorm {rb['iNorm']:.6e} | "
                  f"single itr {rs['ls_itr']} iNorm {rs['iNorm']:.6e}")
            # the assembly adds element contributions in element order, so the
            # device-built set-up (another colouring) assembles the same sums
            assert abs(rb["itr"] - rs["ls_itr"]) <= max(2, 0.02 * rs["ls_itr"])
            assert abs(rb["iNorm"] - rs["iNorm"]) <= 1e-8 * rs["iNorm"]


@pytest.mark.parametrize("precision", [_lib.PRECISION_FP64, _lib.PRECISION_MIXED])
def test_batch_newton_graph_equals_eager(torch_cuda, precision):
    """The graph-captured Newton iteration (svg_batch_newton_iteration) against
    the same launches enqueued one by one: identical records and state, bit for
    bit, on every replay (two time steps, eager work and a reset in between)."""
    torch = torch_cuda
    meshes = _meshes(3, seed=11)
    phys = W.Physics()
    bn = B.BatchNewton(meshes, phys, precision=precision)
    ref = bn.time_step(3, graph=False)
    state = [t.clone() for t in (bn.fn.An, bn.fn.Yn, bn.fn.R)]
    for rep in range(2):
        bn.reset()
        torch.cuda.synchronize()
        got = bn.time_step(3, graph=True)
        torch.cuda.synchronize()
        for it in range(3):
            for q in range(len(meshes)):
                assert got[it][q] == ref[it][q], (rep, it, q, got[it][q], ref[it][q])
        for a, b in zip((bn.fn.An, bn.fn.Yn, bn.fn.R), state):
            assert torch.equal(a, b), rep
        # eager work between the replays
        bn.fn.R.mul_(0.5)
        print(f"replay {rep}: itr {[r['itr'] for r in got[0]]} identical to the eager iteration")


@pytest.mark.parametrize("precision", [_lib.PRECISION_FP64, _lib.PRECISION_MIXED])
def test_batch_learned_equals_individual_learned(torch_cuda, precision):
    """The learned preconditioner in the batched solve: each mesh's flexible
    GMRES (per-mesh fallback) is bit for bit the single-system learned solve of
    that mesh (iterations, norms, solution, fallback point): the trained network
    on every 5th step, and a harmful one (large random decoder) on every step,
    which makes the meshes fall back to the diagonal preconditioner."""
    import os
    from oracle import lpc as O
    from svfsiplus_amd.lpc import LearnedPC, init_params
    from svfsiplus_amd.newton import solve_opts
    torch = torch_cuda
    trained = LearnedPC.load(os.path.join(os.path.dirname(_lib.LIB_PATH), "data", "lpc_weights.npz"))
    p = init_params(5)
    p[O.OFF_WD:] = (np.random.default_rng(105).standard_normal(4 * O.F) * 50.0).astype(np.float32)
    harmful = LearnedPC(p)
    meshes = _meshes(4, seed=13)
    phys = W.Physics()
    bn = _assembled_union(torch, meshes, phys)
    R0, V0 = bn.fn.R.clone(), bn.fn.Val.clone()
    boff = np.concatenate([[0], np.cumsum([m.nnz for m in meshes])])
    window = 20
    for case, net, period in (("trained", trained, 5), ("harmful", harmful, 1)):
        singles = []
        for q, m in enumerate(meshes):
            a, b = int(bn.noff[q]), int(bn.noff[q + 1])
            Rq = R0[a:b].clone()
            Vq = V0[boff[q]:boff[q + 1]].clone()
            ls = ls_create(_lib.LS_GMRES, relTol=1e-6, absTol=1e-30, maxItr=1000, dimKry=50)
            info = {}
            fsils_solve(DeviceLhs(m), ls, 4, Rq, Vq, precision=precision, lpc=net, fallback_window=window,
                        net_period=period, info=info)
            singles.append((Rq, ls.RI.itr, ls.RI.iNorm, ls.RI.fNorm, ls.RI.suc, info["fallback_itr"]))
        Rb, Vb = R0.clone(), V0.clone()
        opts = solve_opts(precision, net, window, period)
        _lib.check(_lib.load().svg_batch_fsils_solve_ex(bn.h, C.byref(bn.proto), Rb.data_ptr(), Vb.data_ptr(),
                                                        C.byref(opts), bn.out, None))
        torch.cuda.synchronize()
        fb = np.empty(len(meshes), dtype=np.int32)
        _lib.check(_lib.load().svg_batch_fallback_itr(bn.h, fb.ctypes.data))
        for q, (Rq, itr, iN, fN, suc, fbi) in enumerate(singles):
            a, b = int(bn.noff[q]), int(bn.noff[q + 1])
            diff = float((Rb[a:b] - Rq).abs().max())
            print(f"{case} mesh {q}: itr batch {bn.out[q].itr} single {itr} fallback {fb[q]}/{fbi} "
                  f"fNorm {bn.out[q].fNorm:.6e}/{fN:.6e} max|dx| {diff:.2e}")
            assert (bn.out[q].itr, bn.out[q].iNorm, bn.out[q].fNorm, bn.out[q].suc) == (itr, iN, fN, suc)
            assert fb[q] == fbi
            assert diff == 0.0
            assert bn.out[q].suc == 1
        if case == "harmful":
            assert (fb >= window).all(), fb


def test_batch_graph_recaptured_after_network_reallocation(torch_cuda):
    """The batched restart-cycle graph holds the network's activation buffers.
    A training step on a larger system reallocates them; the next batched time
    step recaptures the graph and gives, bit for bit, what a batch created
    afresh with the same (trained) network gives."""
    import os
    from svfsiplus_amd import lpc as L
    torch = torch_cuda
    net = L.LearnedPC.load(os.path.join(os.path.dirname(_lib.LIB_PATH), "data", "lpc_weights.npz"))
    meshes = _meshes(3, seed=17)
    phys = W.Physics()
    bn = B.BatchNewton(meshes, phys, precision=_lib.PRECISION_MIXED, lpc=net, net_period=5)
    first = bn.time_step(1)
    # training on a system larger than any mesh of the batch: new activation buffers
    big = L.ScaledSystem(W.make_pipe(10), phys)
    Rt = (big.R / torch.linalg.vector_norm(big.R)).reshape(1, big.n, 4).repeat(2, 1, 1).contiguous()
    loss = C.c_double()
    _lib.check(_lib.load().svg_lpc_train_step(net.h, big.lhs.h, big.Val.data_ptr(), 2, Rt.data_ptr(),
                                              C.c_float(1e-4), C.byref(loss), None))
    torch.cuda.synchronize()
    bn.reset()
    again = bn.time_step(1)
    fresh = B.BatchNewton(meshes, phys, precision=_lib.PRECISION_MIXED, lpc=net, net_period=5)
    ref = fresh.time_step(1)
    print("before training", [r["itr"] for r in first[0]], "after (reused batch)", [r["itr"] for r in again[0]],
          "after (fresh batch)", [r["itr"] for r in ref[0]])
    assert again == ref
    assert torch.equal(bn.fn.Yn, fresh.fn.Yn) and torch.equal(bn.fn.An, fresh.fn.An)
    assert all(r["suc"] == 1 for r in again[0])
    fresh.free()
    bn.free()


@pytest.mark.parametrize("precision", [_lib.PRECISION_FP64, _lib.PRECISION_MIXED])
def test_batch_twolevel_equals_individual_twolevel(torch_cuda, precision):
    """The two-level preconditioner in the batched solve (one dense coarse block
    per mesh): each mesh's flexible GMRES is bit for bit the single-system
    two-level solve of that mesh (iteration count, norms, solution), and the
    aggregates of the union are each mesh's own aggregates, offset."""
    torch = torch_cuda
    meshes = _meshes(3, seed=23)
    phys = W.Physics()
    agg = 16
    bn = _assembled_union(torch, meshes, phys)
    R0, V0 = bn.fn.R.clone(), bn.fn.Val.clone()
    boff = np.concatenate([[0], np.cumsum([m.nnz for m in meshes])])
    opts = _lib.SolveOpts()
    opts.precision, opts.method, opts.coarse, opts.fallback_window = precision, _lib.METHOD_LEARNED, agg, 20
    Rb, Vb = R0.clone(), V0.clone()
    _lib.check(_lib.load().svg_batch_fsils_solve_ex(bn.h, C.byref(bn.proto), Rb.data_ptr(), Vb.data_ptr(),
                                                    C.byref(opts), bn.out, None))
    torch.cuda.synchronize()
    na_u = C.c_int()
    agg_u = np.empty(bn.pipe.nNo, dtype=np.int32)
    _lib.check(_lib.load().svg_coarse_info(bn.fn.lhs.h, C.byref(na_u), agg_u.ctypes.data))
    a_off = 0
    for q, m in enumerate(meshes):
        a, b = int(bn.noff[q]), int(bn.noff[q + 1])
        Rq = R0[a:b].clone()
        Vq = V0[boff[q]:boff[q + 1]].clone()
        L = DeviceLhs(m)
        ls = ls_create(_lib.LS_GMRES, relTol=1e-6, absTol=1e-30, maxItr=1000, dimKry=50)
        fsils_solve(L, ls, 4, Rq, Vq, precision=precision, coarse=agg, fallback_window=20)
        na = C.c_int()
        agg_q = np.empty(m.nNo, dtype=np.int32)
        _lib.check(_lib.load().svg_coarse_info(L.h, C.byref(na), agg_q.ctypes.data))
        assert np.array_equal(agg_u[a:b], agg_q + a_off)
        a_off += na.value
        diff = float((Rb[a:b] - Rq).abs().max())
        print(f"mesh {q}: nNo {m.nNo} aggregates {na.value} itr batch {bn.out[q].itr} single {ls.RI.itr} "
              f"max|dx| {diff:.2e}")
        assert bn.out[q].suc == ls.RI.suc == 1
        assert bn.out[q].itr == ls.RI.itr
        assert bn.out[q].iNorm == ls.RI.iNorm and bn.out[q].fNorm == ls.RI.fNorm
        assert diff == 0.0
        L.free()
    assert a_off == na_u.value
    bn.free()


def test_batch_twolevel_newton_graph_equals_eager(torch_cuda):
    """The graph-captured Newton iteration with the two-level preconditioner
    (its coarse blocks assembled and inverted inside the captured graph)
    against the eager launches: identical records and state on every replay."""
    torch = torch_cuda
    meshes = _meshes(3, seed=29)
    phys = W.Physics()
    bn = B.BatchNewton(meshes, phys, precision=_lib.PRECISION_MIXED, coarse=16)
    ref = bn.time_step(3, graph=False)
    state = [t.clone() for t in (bn.fn.An, bn.fn.Yn, bn.fn.R)]
    for rep in range(2):
        bn.reset()
        torch.cuda.synchronize()
        got = bn.time_step(3, graph=True)
        torch.cuda.synchronize()
        for it in range(3):
            for q in range(len(meshes)):
                assert got[it][q] == ref[it][q], (rep, it, q, got[it][q], ref[it][q])
                assert got[it][q]["suc"] == 1
        for a, b in zip((bn.fn.An, bn.fn.Yn, bn.fn.R), state):
            assert torch.equal(a, b), rep
        bn.fn.R.mul_(0.5)
        print(f"replay {rep}: itr {[r['itr'] for r in got[0]]} identical to the eager iteration")
    bn.free()


@pytest.mark.parametrize("mItr", [3, 1000])
def test_batch_cycles_per_replay(torch_cuda, mItr):
    """Several restart cycles per graph replay: every mesh's count, norms and
    solution equal its single-mesh solve bit for bit, also when the restart
    budget mItr cuts the solves short (the budget is enforced per mesh on the
    device, in the middle of a replay), and the host synchronises once per
    replay, fewer times than the cycles the slowest mesh runs."""
    torch = torch_cuda
    meshes = _meshes(4, seed=13)
    phys = W.Physics()
    bn = _assembled_union(torch, meshes, phys)
    R0, V0 = bn.fn.R.clone(), bn.fn.Val.clone()
    boff = np.concatenate([[0], np.cumsum([m.nnz for m in meshes])])
    tol, sD = 1e-9, 20
    singles = []
    for q, m in enumerate(meshes):
        a, b = int(bn.noff[q]), int(bn.noff[q + 1])
        Rq, Vq = R0[a:b].clone(), V0[boff[q]:boff[q + 1]].clone()
        ls = ls_create(_lib.LS_GMRES, relTol=tol, absTol=1e-30, maxItr=mItr, dimKry=sD)
        fsils_solve(DeviceLhs(m), ls, 4, Rq, Vq)
        singles.append((Rq, ls.RI.itr, ls.RI.iNorm, ls.RI.fNorm, ls.RI.suc))
    h = C.c_void_p()
    _lib.check(_lib.load().svg_batch_create(bn.fn.lhs.h, bn.nsys, bn.noff.ctypes.data, sD, C.byref(h)))
    proto = _lib.SubLs()
    proto.mItr, proto.sD, proto.absTol, proto.relTol = mItr, sD, 1e-30, tol
    out = (_lib.SubLs * bn.nsys)()
    Rb, Vb = R0.clone(), V0.clone()
    r0 = C.c_longlong()
    _lib.check(_lib.load().svg_batch_replays(h, C.byref(r0)))
    _lib.check(_lib.load().svg_batch_fsils_solve(h, C.byref(proto), Rb.data_ptr(), Vb.data_ptr(),
                                                 _lib.PRECISION_FP64, out, None))
    torch.cuda.synchronize()
    r1 = C.c_longlong()
    _lib.check(_lib.load().svg_batch_replays(h, C.byref(r1)))
    cycles = []
    for q, (Rq, itr, iN, fN, suc) in enumerate(singles):
        a, b = int(bn.noff[q]), int(bn.noff[q + 1])
        print(f"mesh {q}: itr batch {out[q].itr} single {itr} fNorm {out[q].fNorm:.6e}/{fN:.6e} suc {out[q].suc}/{suc}")
        assert (out[q].itr, out[q].iNorm, out[q].fNorm, out[q].suc) == (itr, iN, fN, suc)
        assert torch.equal(Rb[a:b], Rq)
        cycles.append(-(-itr // (sD + 1)))
    replays = r1.value - r0.value
    print(f"mItr {mItr}: {replays} replays, the slowest mesh ran about {max(cycles)} cycles")
    assert replays < max(cycles) or max(cycles) <= 1
    _lib.load().svg_batch_free(h)
