"""Learned preconditioner on the GPU: the library's BF16-MFMA forward and
backward kernels against the CPU oracle (oracle/lpc.py), training, the
learned-preconditioned flexible GMRES (FP64 and mixed precision), its fallback
to the diagonal preconditioner, and what a trained network gains over it."""
import ctypes as C
import os

import numpy as np
import pytest

from oracle import lpc as O
from svfsiplus_amd import _lib, lpc as L, workload as W
from svfsiplus_amd.newton import _stream

pytestmark = pytest.mark.gpu


def _sys(torch_cuda, n=4, seed=None, mu=0.04, peak=30.0, sten=0.0):
    kw = {} if seed is None else {"seed": seed}
    if sten > 0.0:
        kw.update(sten_depth=sten, sten_z0=1.25, sten_len=0.5)
    return L.ScaledSystem(W.make_pipe(n, **kw), W.Physics(mu=mu, peak=peak))


def _params(seed=1, wd=0.3):
    p = L.init_params(seed)
    rng = np.random.default_rng(seed + 100)
    p[O.OFF_WD:] = (rng.standard_normal(4 * O.F) * wd).astype(np.float32)
    for l in range(O.NL):
        o = O.OFF_L + l * O.LSZ + 2 * O.F * O.F
        p[o:o + O.F] = (rng.standard_normal(O.F) * 0.05).astype(np.float32)
    return p


def _host(sys):
    return (sys.Val.cpu().numpy(), sys.pipe.rowPtr, sys.pipe.colPtr, sys.pipe.diagPtr())


def _vectors(torch, sys, B, seed):
    g = torch.Generator(device="cuda").manual_seed(seed)
    V = torch.randn((B, sys.n, 4), dtype=torch.float64, device="cuda", generator=g)
    V = V * (sys.R.abs() > 0).to(torch.float64)  # zero on the Dirichlet dofs, as the solver's vectors
    return (V / torch.linalg.vector_norm(V.reshape(B, -1), dim=1)[:, None, None]).contiguous()


def test_lpc_forward_matches_oracle(torch_cuda):
    torch = torch_cuda
    sys = _sys(torch)
    net = L.LearnedPC(_params())
    V = _vectors(torch, sys, 3, 0)
    Z = torch.empty_like(V)
    _lib.check(_lib.load().svg_lpc_apply(net.h, sys.lhs.h, 3, V.data_ptr(), Z.data_ptr(), _stream()))
    Val, rp, cp, dp = _host(sys)
    zo, _ = O.forward(net.params(), Val, rp, cp, dp, V.cpu().numpy())
    sc = np.sqrt(sys.n)
    yg = (Z.cpu().numpy() - V.cpu().numpy()) * sc
    yo = (zo - V.cpu().numpy()) * sc
    rel = np.linalg.norm(yg - yo) / np.linalg.norm(yo)
    print(f"lpc forward: |y| {np.linalg.norm(yo):.3e} rel err gpu vs oracle {rel:.2e}")
    assert np.linalg.norm(yo) > 0.1
    # the kernels round to BF16 where the oracle does; what remains is the FP32
    # accumulation order (observed ~8e-7)
    assert rel < 1e-5


def test_lpc_gradient_matches_oracle(torch_cuda):
    torch = torch_cuda
    sys = _sys(torch)
    net = L.LearnedPC(_params(3))
    R = _vectors(torch, sys, 2, 1)
    loss = C.c_double()
    g = np.empty(L.num_params(), dtype=np.float32)
    _lib.check(_lib.load().svg_lpc_loss_grad(net.h, sys.lhs.h, sys.Val.data_ptr(), 2, R.data_ptr(), C.byref(loss),
                                             g.ctypes.data, _stream()))
    Val, rp, cp, dp = _host(sys)
    lo, go, _ = O.loss_grad(net.params(), Val, rp, cp, dp, R.cpu().numpy())
    print(f"lpc loss gpu {loss.value:.6e} oracle {lo:.6e}")
    assert abs(loss.value - lo) <= 1e-5 * lo
    blocks = [("W0", O.OFF_W0, O.OFF_L)]
    for l in range(O.NL):
        o = O.OFF_L + l * O.LSZ
        blocks += [(f"W{l + 1}", o, o + 2 * O.F * O.F), (f"b{l + 1}", o + 2 * O.F * O.F, o + O.LSZ)]
    blocks.append(("Wd", O.OFF_WD, O.NPARAM))
    for name, a, b in blocks:
        rel = np.linalg.norm(g[a:b] - go[a:b]) / max(np.linalg.norm(go[a:b]), 1e-30)
        print(f"  grad {name}: |g| {np.linalg.norm(go[a:b]):.3e} rel err {rel:.2e}")
        assert rel < 1e-5, name  # observed <= 3e-7


def test_lpc_training_reduces_loss(torch_cuda):
    torch = torch_cuda
    sys = _sys(torch)
    R, E = L.solver_triples(sys, k_max=4)
    net = L.LearnedPC(seed=0)
    data = [(sys, R, E)]
    before = L.evaluate(net, data)
    hist = L.train(net, data, steps=150, lr=2e-3, batch=4)
    after = L.evaluate(net, data)
    print(f"lpc training: loss {before:.4e} -> {after:.4e} (first/last step {hist[0]:.4e}/{hist[-1]:.4e})")
    assert after < 0.8 * before


def _solve(sys, lpc=None, precision=_lib.PRECISION_FP64, window=0, tol=1e-6, period=1):
    from svfsiplus_amd.newton import fsils_solve, ls_create
    x = sys.R.clone()
    ls = ls_create(_lib.LS_GMRES, relTol=tol, absTol=1e-30, maxItr=1000, dimKry=50)
    info = {}
    fsils_solve(sys.lhs, ls, 4, x, sys.Val, prec=_lib.PREC_NONE, precision=precision, lpc=lpc,
                fallback_window=window, net_period=period, info=info)
    import torch
    rel = float(torch.linalg.vector_norm(sys.residual(x)) / torch.linalg.vector_norm(sys.R))
    return x, ls, rel, info.get("fallback_itr", -1)


@pytest.mark.parametrize("precision", [_lib.PRECISION_FP64, _lib.PRECISION_MIXED])
def test_fgmres_identity_network_equals_gmres(torch_cuda, precision):
    """With Wd = 0 the flexible solve is the plain GMRES (z_i = u_i)."""
    torch = torch_cuda
    sys = _sys(torch, n=6)
    p = L.init_params(0)  # Wd = 0: M = identity on the scaled system
    net = L.LearnedPC(p)
    _, ls0, rel0, _ = _solve(sys, precision=precision)
    _, ls1, rel1, fb = _solve(sys, lpc=net, precision=precision)
    print(f"identity network: plain itr {ls0.RI.itr} rel {rel0:.2e} | flexible itr {ls1.RI.itr} rel {rel1:.2e} fb {fb}")
    assert ls1.RI.suc == 1 and rel1 < 1e-6
    assert abs(ls1.RI.itr - ls0.RI.itr) <= max(3, ls0.RI.itr // 20)


def test_fgmres_fallback(torch_cuda):
    torch = torch_cuda
    sys = _sys(torch, n=6)
    p = _params(5, wd=50.0)  # a harmful network: large random decoder
    net = L.LearnedPC(p)
    _, ls, rel, fb = _solve(sys, lpc=net, window=20)
    print(f"harmful network: itr {ls.RI.itr} fallback at {fb} rel {rel:.2e}")
    assert fb >= 20
    assert ls.RI.suc == 1 and rel < 1e-6


def test_trained_network_gain(torch_cuda):
    """A network trained on solver triples of small seeded pipes cuts the
    flexible GMRES iteration count on a held-out pipe below the plain
    diagonal-preconditioned GMRES count at the same 1e-6 FP64 residual."""
    torch = torch_cuda
    data = []
    for d in L.draws(8, seed=7, sizes=(4, 6)):
        sys = L.ScaledSystem(d.pipe(), d.physics())
        R, E = L.solver_triples(sys, k_max=6)
        data.append((sys, R, E))
    net = L.LearnedPC(seed=0)
    L.train(net, data, steps=600, lr=2e-3, batch=6)
    held = _sys(torch, n=6, seed=99, mu=0.05, peak=25.0, sten=0.3)
    _, ls0, rel0, _ = _solve(held)
    for period in (1, 5):
        _, ls1, rel1, fb = _solve(held, lpc=net, window=20, period=period)
        print(f"trained network on held-out pipe: plain itr {ls0.RI.itr} | learned (network every {period}. step) "
              f"itr {ls1.RI.itr} (fallback {fb}) rel {rel0:.2e} / {rel1:.2e}")
        assert ls1.RI.suc == 1 and rel1 < 1e-6
        assert ls1.RI.itr < ls0.RI.itr


def test_learned_after_larger_then_smaller_krylov_dimension(torch_cuda):
    """One lhs solved with the plain GMRES at Krylov dimension 100, then the
    learned flexible GMRES at 30 and at 60: the workspace kept from the first
    solve serves the later ones (the flexible basis is sized for its capacity).
    Each meets the FP64 true residual, and the last equals, bit for bit, the
    same learned solve on a freshly created lhs."""
    import torch
    from svfsiplus_amd.newton import fsils_solve, ls_create
    sys = _sys(torch_cuda, n=6)
    net = L.LearnedPC.load(os.path.join(os.path.dirname(_lib.LIB_PATH), "data", "lpc_weights.npz"))

    def solve(s, sD, lpc):
        x = s.R.clone()
        ls = ls_create(_lib.LS_GMRES, relTol=1e-6, absTol=1e-30, maxItr=1000, dimKry=sD)
        fsils_solve(s.lhs, ls, 4, x, s.Val, prec=_lib.PREC_NONE, precision=_lib.PRECISION_MIXED, lpc=lpc,
                    fallback_window=20, net_period=5)
        rel = float(torch.linalg.vector_norm(s.residual(x)) / torch.linalg.vector_norm(s.R))
        return x, ls, rel

    out = {}
    for sD, lpc in ((100, None), (30, net), (60, net)):
        x, ls, rel = solve(sys, sD, lpc)
        print(f"sD {sD} {'learned' if lpc else 'plain'}: itr {ls.RI.itr} true rel residual {rel:.2e}")
        assert ls.RI.suc == 1 and rel <= 1e-6
        out[sD] = (x, ls.RI.itr)
    fresh = _sys(torch_cuda, n=6)
    x, ls, rel = solve(fresh, 60, net)
    assert ls.RI.itr == out[60][1]
    assert torch.equal(x, out[60][0])


def test_lpc_apply_bitwise_regression(torch_cuda):
    """svg_lpc_apply on seeded inputs against the stored output of the same
    kernels (tests/golden/lpc_apply_n4.npz, written by
    tests/golden/make_lpc_golden.py): bit for bit, so a change to the network
    kernels meant to keep their arithmetic is checked to keep it."""
    import sys as _sys
    golden = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    _sys.path.insert(0, golden)
    import make_lpc_golden as G
    d = np.load(os.path.join(golden, "lpc_apply_n4.npz"))
    p, s, V = G.inputs()
    assert np.array_equal(p, d["params"]) and np.array_equal(V, d["V"])
    Z = G.apply(p, s, V)
    print(f"lpc apply: {(Z != d['Z']).sum()} of {Z.size} entries differ from the stored output")
    assert np.array_equal(Z, d["Z"])
