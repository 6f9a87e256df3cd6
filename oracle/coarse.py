"""CPU oracle of the two-level preconditioner (test infrastructure only).

The preconditioner has no counterpart in the reference (it stands beside
precond_diag, Code/Source/liner_solver/precond.cpp:122-257): on the diagonally
scaled system A (Val(16, nnz) after precond_diag, block j row-major as in
spar_mul.cpp:228-237)

    M(v) = P zc + B (v - A P zc),   zc = Ac^-1 P^T v,   Ac = P^T A P

with P the piecewise-constant prolongation of the pressure (dof 3) over node
aggregates and B the inverses of the 4x4 diagonal blocks. This module restates
it in numpy/scipy in FP64 (the library applies it in FP32):

* ``aggregate``  the library's breadth-first aggregation of the node graph
  (seeds in node order; each aggregate grows over unaggregated neighbours, in
  CSR order, to `target` nodes; aggregates under target/4 nodes join the
  neighbouring aggregate of at least target/4 nodes they share most edges with,
  ties to the lowest id; labels renumbered by first node);
* ``TwoLevel``   Ac, its inverse, B and the three products the library exposes
  (apply, the coarse-corrected residual, the block-Jacobi smoothing).

Parity of this restatement is pinned by properties, not by retyping: the
coarse-corrected residual has zero aggregate pressure sums (P^T r2 = 0, the
Galerkin condition), M reproduces the coarse space (M(A P c) = P c), B inverts
the diagonal blocks, and every aggregate is a connected set of nodes.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp


def aggregate(rowPtr, colPtr, target: int):
    """Node -> aggregate labels (int64) and the number of aggregates."""
    rp = np.asarray(rowPtr)
    cp = np.asarray(colPtr)
    nNo = rp.shape[0] - 1
    agg = np.full(nNo, -1, dtype=np.int64)
    size = []
    na = 0
    for s in range(nNo):
        if agg[s] >= 0:
            continue
        q = [s]
        agg[s] = na
        cnt = 1
        h = 0
        while h < len(q) and cnt < target:
            v = q[h]
            h += 1
            for k in range(rp[v], rp[v + 1]):
                if cnt >= target:
                    break
                w = cp[k]
                if agg[w] < 0:
                    agg[w] = na
                    q.append(w)
                    cnt += 1
        size.append(cnt)
        na += 1
    size = np.asarray(size)
    small = max(1, target // 4)
    dest = np.arange(na)
    members = [[] for _ in range(na)]
    for a in range(nNo):
        members[agg[a]].append(a)
    for a in range(na):
        if size[a] >= small:
            continue
        edges = {}
        for v in members[a]:
            for k in range(rp[v], rp[v + 1]):
                b = int(agg[cp[k]])
                if b != a and size[b] >= small:
                    edges[b] = edges.get(b, 0) + 1
        if edges:
            best = max(edges.values())
            dest[a] = min(b for b, c in edges.items() if c == best)
    label = np.full(na, -1, dtype=np.int64)
    out = np.empty(nNo, dtype=np.int64)
    n2 = 0
    for a in range(nNo):
        d = dest[agg[a]]
        if label[d] < 0:
            label[d] = n2
            n2 += 1
        out[a] = label[d]
    return out, n2


def block_matrix(rowPtr, colPtr, Val):
    """The scaled 4x4-block matrix as scipy CSR (rows 4a + i, columns 4b + j)."""
    rp = np.asarray(rowPtr)
    cp = np.asarray(colPtr)
    nNo = rp.shape[0] - 1
    V = np.asarray(Val, dtype=np.float64).reshape(-1, 4, 4)
    return sp.bsr_matrix((V, cp, rp), shape=(4 * nNo, 4 * nNo)).tocsr()


class TwoLevel:
    """M(v) = P Ac^-1 P^T v + B (v - A P Ac^-1 P^T v) on the scaled system."""

    def __init__(self, rowPtr, colPtr, Val, agg, nAgg):
        rp = np.asarray(rowPtr)
        nNo = rp.shape[0] - 1
        self.nNo, self.nAgg = nNo, nAgg
        self.agg = np.asarray(agg, dtype=np.int64)
        self.A = block_matrix(rowPtr, colPtr, Val)
        rows = 4 * np.arange(nNo) + 3
        self.P = sp.csr_matrix((np.ones(nNo), (rows, self.agg)), shape=(4 * nNo, nAgg))
        self.Ac = np.asarray((self.P.T @ self.A @ self.P).todense())
        self.Aci = np.linalg.inv(self.Ac)
        # diagonal blocks; a zero diagonal entry (Dirichlet row and column zeroed
        # by the scaling) becomes 1; a singular block gives the identity
        D = np.zeros((nNo, 4, 4))
        idx = 4 * np.arange(nNo)
        A = self.A.tocsr()
        for i in range(4):
            for j in range(4):
                D[:, i, j] = np.asarray(A[idx + i, idx + j]).ravel()
        for i in range(4):
            z = D[:, i, i] == 0.0
            D[z, i, i] = 1.0
        self.D = D
        Bi = np.empty_like(D)
        for a in range(nNo):
            try:
                Bi[a] = np.linalg.inv(D[a])
            except np.linalg.LinAlgError:
                Bi[a] = np.eye(4)
        self.Binv = Bi

    def _zc(self, v):
        return self.Aci @ (self.P.T @ v)

    def residual(self, v):
        """r2 = v - A P zc (flat, 4 nNo)."""
        return v - self.A @ (self.P @ self._zc(v))

    def smooth(self, v):
        """B r2."""
        r2 = self.residual(v).reshape(-1, 4)
        return np.einsum("nij,nj->ni", self.Binv, r2).ravel()

    def apply(self, v):
        """M(v) = P zc + B r2."""
        return self.P @ self._zc(v) + self.smooth(v)


def fgmres(A, b, M, m=50, tol=1e-6, maxit=5000):
    """Right-preconditioned flexible GMRES(m) (modified Gram-Schmidt, Givens
    rotations), x0 = 0; stops when the Givens estimate of |b - A x| is below
    tol |b|. Returns (x, iterations, true relative residual)."""
    import scipy.linalg as sla
    n = b.shape[0]
    x = np.zeros(n)
    bn = np.linalg.norm(b)
    itr = 0
    while itr < maxit:
        r = b - A @ x
        beta = np.linalg.norm(r)
        if beta <= tol * bn:
            break
        Vb = np.zeros((m + 1, n))
        Z = np.zeros((m, n))
        H = np.zeros((m + 1, m))
        Vb[0] = r / beta
        g = np.zeros(m + 1)
        g[0] = beta
        cs, sn = np.zeros(m), np.zeros(m)
        k = 0
        for k in range(m):
            Z[k] = M(Vb[k])
            w = A @ Z[k]
            for j in range(k + 1):
                H[j, k] = w @ Vb[j]
                w = w - H[j, k] * Vb[j]
            H[k + 1, k] = np.linalg.norm(w)
            Vb[k + 1] = w / H[k + 1, k]
            for j in range(k):
                t = cs[j] * H[j, k] + sn[j] * H[j + 1, k]
                H[j + 1, k] = -sn[j] * H[j, k] + cs[j] * H[j + 1, k]
                H[j, k] = t
            den = np.hypot(H[k, k], H[k + 1, k])
            cs[k], sn[k] = H[k, k] / den, H[k + 1, k] / den
            H[k, k] = den
            H[k + 1, k] = 0.0
            g[k + 1] = -sn[k] * g[k]
            g[k] = cs[k] * g[k]
            itr += 1
            if abs(g[k + 1]) <= tol * bn or itr >= maxit:
                break
        y = sla.solve_triangular(H[:k + 1, :k + 1], g[:k + 1])
        x = x + y @ Z[:k + 1]
    return x, itr, np.linalg.norm(b - A @ x) / bn
