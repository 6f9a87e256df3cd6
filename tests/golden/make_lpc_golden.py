"""Writes tests/golden/lpc_apply_n4.npz: the learned preconditioner's output
svg_lpc_apply(net, lhs, 2, V) on the scaled first Newton system of the n = 4
pipe, for seeded parameters and vectors, as the HIP kernels compute it. A
regression fixture of the device kernels themselves (run on a GPU):
tests/test_lpc_gpu.py::test_lpc_apply_bitwise_regression recomputes it and
compares bit for bit, so that a later change to the network kernels that is
meant to leave their arithmetic unchanged is checked to do so.

    python tests/golden/make_lpc_golden.py [out.npz]
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))


def inputs():
    """The seeded inputs: network parameters, the n = 4 pipe and two vectors
    (zero on the Dirichlet dofs, unit norm)."""
    from oracle import lpc as O
    from svfsiplus_amd import lpc as L, workload as W
    p = L.init_params(7)
    rng = np.random.default_rng(107)
    p[O.OFF_WD:] = (rng.standard_normal(4 * O.F) * 0.3).astype(np.float32)
    for l in range(O.NL):
        o = O.OFF_L + l * O.LSZ + 2 * O.F * O.F
        p[o:o + O.F] = (rng.standard_normal(O.F) * 0.05).astype(np.float32)
    sys_ = L.ScaledSystem(W.make_pipe(4), W.Physics())
    mask = (sys_.R.abs() > 0).cpu().numpy()
    V = np.random.default_rng(11).standard_normal((2, sys_.n, 4)) * mask
    V /= np.linalg.norm(V.reshape(2, -1), axis=1)[:, None, None]
    return p, sys_, V


def apply(p, sys_, V):
    import torch
    from svfsiplus_amd import _lib, lpc as L
    from svfsiplus_amd.newton import _stream
    net = L.LearnedPC(p)
    dV = torch.tensor(V, dtype=torch.float64, device="cuda").contiguous()
    dZ = torch.empty_like(dV)
    _lib.check(_lib.load().svg_lpc_apply(net.h, sys_.lhs.h, V.shape[0], dV.data_ptr(), dZ.data_ptr(), _stream()))
    torch.cuda.synchronize()
    return dZ.cpu().numpy()


if __name__ == "__main__":
    out = sys.argv[1] if len(sys.argv) > 1 else os.path.join(HERE, "lpc_apply_n4.npz")
    p, s, V = inputs()
    Z = apply(p, s, V)
    np.savez_compressed(out, V=V, Z=Z, params=p)
    print(f"wrote {out}: Z {Z.shape}, |Z - V| {np.linalg.norm(Z - V):.6e}")
