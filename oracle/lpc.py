"""CPU oracle of the learned preconditioner network — TEST INFRASTRUCTURE ONLY.

A numpy restatement of the message-passing network of the HIP library
(forward pass, the training loss and its hand-derived gradient), used by the
tests to check the library's BF16-MFMA forward and backward kernels. Only
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import it.

The network has no counterpart in the reference (it stands beside the
diagonal preconditioner of precond.cpp:122-257). With bf16=True the oracle
rounds to BF16 (round-to-nearest-even, via FP32) at exactly the points where
the kernels store BF16: encoder input, weights, the matrix copy, every
activation h, every message m, the relu mask source p, and the gradient tiles
fed to the MFMA weight-gradient products; sums are in float64 (the kernels sum
in FP32), so the two agree to BF16 resolution. With bf16=False the oracle is
plain float64 and its gradient is pinned by central finite differences.

Parameter vector (float32): W0 [64][32] | 3 x (W_l [64][128], b_l [64]) | Wd [4][64].

    x   = v sqrt(n);  u = [x (4), D_i (16), 1, 0 x 11];  h0 = W0 u
    m_i = sum_{j in row i} A_ij (x) h_j      m[c][d] = sum_e A_ij[d][e] h_j[c][e]
    p   = W_l [h; m] + b_l;  h <- h + relu(p)         (3 layers)
    z   = v + (Wd h) / sqrt(n)
    L   = (1/B) sum_b |r_b - A z_b|^2 with v_b = r_b.
"""
from __future__ import annotations

import numpy as np

F, NIN, NL = 64, 32, 3
OFF_W0 = 0
OFF_L = F * NIN
LSZ = F * 2 * F + F
OFF_WD = OFF_L + NL * LSZ
NPARAM = OFF_WD + 4 * F


def bf16(x):
    """Round to BF16 (nearest even) through FP32, returned as float64."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _f32(x):
    return np.asarray(x, dtype=np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    W0 = p[OFF_W0:OFF_W0 + F * NIN].reshape(F, NIN)
    layers = []
    for l in range(NL):
        o = OFF_L + l * LSZ
        layers.append((p[o:o + 2 * F * F].reshape(F, 2 * F), p[o + 2 * F * F:o + LSZ]))
    Wd = p[OFF_WD:OFF_WD + 4 * F].reshape(4, F)
    return W0, layers, Wd


def _rows(rowPtr):
    rowPtr = np.asarray(rowPtr, dtype=np.int64)
    return np.repeat(np.arange(len(rowPtr) - 1), np.diff(rowPtr))


def message(A, rows, cols, n, h):
    """m[b,i,c,d] = sum_{k: rows[k]=i} sum_e A[k,d,e] h[b,cols[k],c,e]; h (B,n,64)."""
    B = h.shape[0]
    hb = h.reshape(B, n, 16, 4)[:, cols]                      # (B,nnz,16,4)
    contrib = np.einsum("kde,bkce->bkcd", A, hb)               # (B,nnz,16,4)
    m = np.zeros((B, n, 16, 4))
    for b in range(B):
        np.add.at(m[b], rows, contrib[b])
    return m.reshape(B, n, F)


def message_T(A, rows, cols, n, g):
    """Adjoint of message: out[b,j,c,e] = sum_{k: cols[k]=j} sum_d A[k,d,e] g[b,rows[k],c,d]."""
    B = g.shape[0]
    gb = g.reshape(B, n, 16, 4)[:, rows]
    contrib = np.einsum("kde,bkcd->bkce", A, gb)
    out = np.zeros((B, n, 16, 4))
    for b in range(B):
        np.add.at(out[b], cols, contrib[b])
    return out.reshape(B, n, F)


def bsr_matvec(Val, rows, cols, n, x):
    """y = A x for the (nnz,16) row-major 4x4 blocks; x (n,4)."""
    A = Val.reshape(-1, 4, 4)
    contrib = np.einsum("kde,ke->kd", A, x[cols])
    y = np.zeros((n, 4))
    np.add.at(y, rows, contrib)
    return y


def bsr_rmatvec(Val, rows, cols, n, x):
    """y = A^T x."""
    A = Val.reshape(-1, 4, 4)
    contrib = np.einsum("kde,kd->ke", A, x[rows])
    y = np.zeros((n, 4))
    np.add.at(y, cols, contrib)
    return y


def forward(params, Val, rowPtr, colPtr, diagPtr, V, bf16_mode=True):
    """z = M(V) for V (B,n,4) on the scaled matrix Val (nnz,16); returns z and the cache."""
    rnd = bf16 if bf16_mode else (lambda x: np.asarray(x, dtype=np.float64))
    V = np.asarray(V, dtype=np.float64)
    B, n, _ = V.shape
    sc = np.sqrt(n)
    rows = _rows(rowPtr)
    cols = np.asarray(colPtr, dtype=np.int64)
    W0, layers, Wd = unpack(params)
    W0b = rnd(W0)
    Wdb = rnd(Wd)
    A = rnd(_f32(Val)).reshape(-1, 4, 4) if bf16_mode else np.asarray(Val, dtype=np.float64).reshape(-1, 4, 4)
    D = rnd(_f32(np.asarray(Val)[np.asarray(diagPtr)])) if bf16_mode else np.asarray(Val)[np.asarray(diagPtr)]
    x = rnd(_f32(V * sc)) if bf16_mode else V * sc
    u = np.zeros((B, n, NIN))
    u[:, :, 0:4] = x
    u[:, :, 4:20] = D[None]
    u[:, :, 20] = 1.0
    h = rnd(u @ W0b.T)
    cache = dict(u=u, H=[h], M=[], P=[], A=A, rows=rows, cols=cols, W0b=W0b, Wdb=Wdb, n=n, B=B, sc=sc,
                 Wb=[rnd(W) for W, _ in layers], bias=[b for _, b in layers])
    for l, (W, b) in enumerate(layers):
        m = rnd(message(A, rows, cols, n, h))
        X = np.concatenate([h, m], axis=2)
        p = X @ cache["Wb"][l].T + b[None, None, :]
        pr = rnd(p)
        h = rnd(h + np.maximum(p, 0.0))
        cache["M"].append(m)
        cache["P"].append(pr)
        cache["H"].append(h)
    y = h @ Wdb.T
    z = V + y / sc
    return z, cache


def loss_grad(params, Val, rowPtr, colPtr, diagPtr, R, bf16_mode=True):
    """L = (1/B) sum_b |R_b - A M(R_b)|^2 (A = the FP64 scaled matrix) and dL/dparams."""
    rnd = bf16 if bf16_mode else (lambda x: np.asarray(x, dtype=np.float64))
    R = np.asarray(R, dtype=np.float64)
    z, c = forward(params, Val, rowPtr, colPtr, diagPtr, R, bf16_mode)
    B, n, sc = c["B"], c["n"], c["sc"]
    rows, cols = c["rows"], c["cols"]
    Val = np.asarray(Val, dtype=np.float64)
    res = np.stack([R[b] - bsr_matvec(Val, rows, cols, n, z[b]) for b in range(B)])
    loss = float(np.sum(res * res) / B)
    gz = np.stack([-2.0 / B * bsr_rmatvec(Val, rows, cols, n, res[b]) for b in range(B)])
    grad = np.zeros(NPARAM)
    gy = gz / sc                                             # (B,n,4)
    h3 = c["H"][NL]
    grad[OFF_WD:OFF_WD + 4 * F] = np.einsum("bnd,bnf->df", gy, h3).ravel()
    G = gy @ c["Wdb"]                                        # (B,n,64)
    for l in range(NL - 1, -1, -1):
        gp = rnd(G * (c["P"][l] > 0))
        X = np.concatenate([c["H"][l], c["M"][l]], axis=2)
        o = OFF_L + l * LSZ
        grad[o:o + 2 * F * F] = np.einsum("bno,bni->oi", gp, X).ravel()
        grad[o + 2 * F * F:o + LSZ] = gp.sum(axis=(0, 1))
        gX = gp @ c["Wb"][l]                                 # (B,n,128)
        G = G + gX[:, :, :F] + message_T(c["A"], rows, cols, n, gX[:, :, F:])
    gh0 = rnd(G)
    grad[OFF_W0:OFF_W0 + F * NIN] = np.einsum("bno,bni->oi", gh0, c["u"]).ravel()
    return loss, grad, z
