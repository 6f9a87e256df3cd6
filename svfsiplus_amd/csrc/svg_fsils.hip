// FSILS solve on the device for one rank (lhs.map is the identity and every
// halo sum / allreduce of the reference is the identity):
//   fsils_solve          liner_solver/solve.cpp:50-193
//   precond_diag         liner_solver/precond.cpp:122-257 (+ pre_mul, pos_mul)
//   gmres_v              liner_solver/gmres.cpp:450-629
//   fsils_spar_mul_vv    liner_solver/spar_mul.cpp:191-258
//   fsils_nc_dot_v/norm  liner_solver/dot.cpp:134-171, norm.cpp:67-112
//   omp_sum_v/omp_mul_v  liner_solver/omp_la.cpp:39-144
//
// Design:
//   * SpMV: one lane per (row, block-row i); a lane walks its row's blocks in
//     CSR order and accumulates K(4i+c) U(c) left to right exactly as the
//     reference loop does (no FMA contraction), so the product is bitwise
//     equal to the reference's. Four lanes of a row read the 128-byte block
//     as one coalesced line.
//   * Reductions: two-pass and deterministic (fixed element partition, fixed
//     butterfly inside a wavefront, fixed-order sums across blocks); the
//     result does not depend on scheduling. The order differs from the
//     reference's serial left-to-right sum, so dots agree to rounding.
//   * GMRES: classical Gram-Schmidt with all dots against the unorthogonalised
//     vector and the Pythagorean norm, exactly as gmres.cpp:550-566. The
//     Hessenberg column, Givens rotations and the stopping test run on the
//     device in a single-workgroup tail, so a whole restart cycle is enqueued
//     without host round trips; kernels after convergence exit at once.
#include "svg_device.h"

#include <chrono>
#include <cfloat>
#include <cmath>
#include <cstring>

namespace svg {

constexpr int RED_BS = 256;   // threads per reduction block (4 wavefronts)
constexpr int RED_EPT = 4;    // elements per thread
constexpr int RED_SPAN = RED_BS * RED_EPT;
constexpr int LOOKAHEAD = 3;  // Arnoldi steps enqueued ahead of the host's convergence check

__device__ __forceinline__ double wave_sum(double v)
{
  // Fixed butterfly over the 64 lanes of a wavefront.
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// ------------------------------------------------------------------ SpMV --
// Blocks are dispatched round-robin over the 8 XCDs; remap so that each XCD
// walks one contiguous band of rows and its L2 keeps the gathered U entries.
__device__ __forceinline__ int xcd_block(int b, int nb)
{
  const int per = (nb + 7) >> 3;
  const int xcd = b & 7, idx = b >> 3;
  const int r = xcd * per + idx;
  return r < nb ? r : -1;
}

template <int D, int MODE>
__global__ __launch_bounds__(256) void k_spmv(int nNo, const int* __restrict__ rowPtr, const int* __restrict__ colPtr,
                                              const double* __restrict__ K, const double* __restrict__ U,
                                              double* __restrict__ KU, const double* __restrict__ B,
                                              const GmState* __restrict__ st)
{
#pragma clang fp contract(off)
  if (st && st->done) return;
  int bx = blockIdx.x;
  if (gridDim.x >= 64) {
    // gridDim is a multiple of 8 for the remapped launches
    bx = xcd_block(blockIdx.x, gridDim.x);
    if (bx < 0) return;
  }
  const int gid = bx * blockDim.x + threadIdx.x;
  const int row = gid / D;
  const int i = gid - row * D;
  if (row >= nNo) return;
  const int s = rowPtr[row], e = rowPtr[row + 1];
  double acc = 0.0;
  if constexpr (D == 4) {
    const double2* Kb = reinterpret_cast<const double2*>(K) + 2 * i;
    const double2* Ub = reinterpret_cast<const double2*>(U);
    int j = s;
    // four blocks per trip: all loads issued first, the adds keep the CSR order
    for (; j + 4 <= e; j += 4) {
      int c[4];
      double2 k0[4], k1[4], u0[4], u1[4];
#pragma unroll
      for (int q = 0; q < 4; q++) c[q] = __builtin_nontemporal_load(colPtr + j + q);
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const double* kp = reinterpret_cast<const double*>(Kb + 8 * (size_t)(j + q));
        k0[q].x = __builtin_nontemporal_load(kp + 0);
        k0[q].y = __builtin_nontemporal_load(kp + 1);
        k1[q].x = __builtin_nontemporal_load(kp + 2);
        k1[q].y = __builtin_nontemporal_load(kp + 3);
        u0[q] = Ub[2 * (size_t)c[q]];
        u1[q] = Ub[2 * (size_t)c[q] + 1];
      }
#pragma unroll
      for (int q = 0; q < 4; q++) {
        double t = acc + k0[q].x * u0[q].x;
        t = t + k0[q].y * u0[q].y;
        t = t + k1[q].x * u1[q].x;
        acc = t + k1[q].y * u1[q].y;
      }
    }
    for (; j < e; j++) {
      const int col = colPtr[j];
      const double2 k0 = Kb[8 * (size_t)j], k1 = Kb[8 * (size_t)j + 1];
      const double2 u0 = Ub[2 * (size_t)col], u1 = Ub[2 * (size_t)col + 1];
      double t = acc + k0.x * u0.x;
      t = t + k0.y * u0.y;
      t = t + k1.x * u1.x;
      acc = t + k1.y * u1.y;
    }
  } else {
    for (int j = s; j < e; j++) {
      const int col = colPtr[j];
      const double* k = K + (size_t)D * D * j + D * i;
      const double* u = U + (size_t)D * col;
      double t = acc + k[0] * u[0];
#pragma unroll
      for (int c = 1; c < D; c++) t = t + k[c] * u[c];
      acc = t;
    }
  }
  if (MODE == 0)
    KU[gid] = acc;
  else
    KU[gid] = B[gid] - acc;
}

template <int D, int MODE>
static void spmv(const svg_lhs* L, const double* K, const double* U, double* KU, const double* B, const GmState* st,
                 hipStream_t s)
{
  size_t nt = (size_t)D * L->nNo;
  unsigned nb = blocks_for(nt, 256);
  if (nb >= 64) nb = (nb + 7) & ~7u;
  launch("spmv_bcsr", k_spmv<D, MODE>, dim3(nb), dim3(256), 0, s, L->nNo, (const int*)L->d_rowPtr,
         (const int*)L->d_colPtr, K, U, KU, B, st);
}

// ------------------------------------------------------------ reductions --
enum FinMode { FIN_DOT = 0, FIN_NORM_R = 1, FIN_RESTART = 2, FIN_ARNOLDI = 3 };

struct FinArgs {
  int mode;
  int i;   // Arnoldi column
  int sD;
  double absTol, relTol;
};

constexpr int MAXM = 512;  // sums per reduction (Krylov dimension + 2)

// Dots of w against m vectors base + j*ld (j < m), reduced deterministically:
// every block writes its partial sums (transposed, partial[j*nblk + b]); the
// last block to finish (atomic ticket) sums the partials of every column in a
// fixed order and runs the tail selected by A.mode:
//   FIN_DOT      out[0] = sum
//   FIN_NORM_R   iNorm = fNorm = sqrt(sum), eps = max(absTol, relTol iNorm)   (gmres.cpp:482-487)
//   FIN_RESTART  err(0) = sqrt(sum), itr += 1                                 (gmres.cpp:501-531)
//   FIN_ARNOLDI  Hessenberg column i, Pythagorean h(i+1,i), Givens, err, stop (gmres.cpp:550-592)
__global__ __launch_bounds__(RED_BS) void k_mdot(int n, const double* __restrict__ w, const double* __restrict__ base,
                                                 size_t ld, int m, double* __restrict__ partial,
                                                 unsigned* __restrict__ ticket, FinArgs A, GmState* st,
                                                 double* __restrict__ h, double* __restrict__ cs,
                                                 double* __restrict__ sn, double* __restrict__ err,
                                                 double* __restrict__ coef, double* __restrict__ out,
                                                 volatile int* host_done)
{
#pragma clang fp contract(off)
  extern __shared__ double lds[];  // [4 waves][m]
  __shared__ double sums[MAXM];
  __shared__ double lh[MAXM], lcs[MAXM], lsn[MAXM];
  __shared__ int amLast;
  if (A.mode == FIN_ARNOLDI && st->done) return;
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const int nblk = gridDim.x;
  const size_t k0 = (size_t)blockIdx.x * RED_SPAN + t;
  double wr[RED_EPT];
#pragma unroll
  for (int r = 0; r < RED_EPT; r++) {
    size_t k = k0 + (size_t)r * RED_BS;
    wr[r] = k < (size_t)n ? w[k] : 0.0;
  }
  for (int j = 0; j < m; j++) {
    const double* uj = base + (size_t)j * ld;
    double acc = 0.0;
#pragma unroll
    for (int r = 0; r < RED_EPT; r++) {
      size_t k = k0 + (size_t)r * RED_BS;
      double v = k < (size_t)n ? uj[k] : 0.0;
      acc = acc + v * wr[r];
    }
    acc = wave_sum(acc);
    if (lane == 0) lds[wv * m + j] = acc;
  }
  __syncthreads();
  for (int j = t; j < m; j += RED_BS)
    partial[(size_t)j * nblk + blockIdx.x] = ((lds[j] + lds[m + j]) + lds[2 * m + j]) + lds[3 * m + j];
  __threadfence();
  __syncthreads();
  if (t == 0) amLast = (atomicAdd(ticket, 1u) == (unsigned)(nblk - 1));
  __syncthreads();
  if (!amLast) return;
  __threadfence();
  // ---- last block: fixed-order sums of the partials of every column
  for (int j = wv; j < m; j += RED_BS / 64) {
    const double* pj = partial + (size_t)j * nblk;
    double acc = 0.0;
    for (int b = lane; b < nblk; b += 64) acc = acc + pj[b];
    acc = wave_sum(acc);
    if (lane == 0) sums[j] = acc;
  }
  if (t == 0) *ticket = 0u;
  if (A.mode == FIN_ARNOLDI) {
    for (int j = t; j < A.i; j += RED_BS) {
      lcs[j] = cs[j];
      lsn[j] = sn[j];
    }
  }
  __syncthreads();
  if (A.mode == FIN_DOT) {
    if (t == 0) out[0] = sums[0];
    return;
  }
  if (A.mode == FIN_NORM_R) {
    if (t == 0) {
      double nrm = sqrt(sums[0]);
      st->iNorm = nrm;
      st->fNorm = nrm;
      st->eps = fmax(A.absTol, A.relTol * nrm);
      st->itr = 0;
      st->suc = 0;
      st->done = 0;
      st->divzero = 0;
      st->last_i = 0;
    }
    return;
  }
  if (A.mode == FIN_RESTART) {
    if (t == 0) {
      double e0 = sqrt(sums[0]);
      err[0] = e0;
      st->itr = st->itr + 1;
      st->done = 0;
      st->scale = e0;
      if (e0 == 0.0) {
        st->divzero = 1;
        st->done = 1;
      }
    }
    return;
  }
  // FIN_ARNOLDI (serial part on LDS by one lane, then parallel stores)
  const int i = A.i, ldh = A.sD + 1;
  __shared__ double lerr[2];
  __shared__ int conv;
  if (t == 0) {
    for (int j = 0; j <= i + 1; j++) lh[j] = sums[j];
    double hh = lh[i + 1];
    for (int j = 0; j <= i; j++) hh = hh - lh[j] * lh[j];
    hh = sqrt(fabs(hh));
    lh[i + 1] = hh;
    st->scale = 1.0 / hh;
    for (int j = 0; j <= i - 1; j++) {
      double tmp = lcs[j] * lh[j] + lsn[j] * lh[j + 1];
      lh[j + 1] = -lsn[j] * lh[j] + lcs[j] * lh[j + 1];
      lh[j] = tmp;
    }
    double tmp = sqrt(lh[i] * lh[i] + lh[i + 1] * lh[i + 1]);
    double c = lh[i] / tmp, sv = lh[i + 1] / tmp;
    cs[i] = c;
    sn[i] = sv;
    lh[i] = tmp;
    lh[i + 1] = 0.0;
    double ei = err[i];
    lerr[1] = -sv * ei;
    lerr[0] = c * ei;
    err[i + 1] = lerr[1];
    err[i] = lerr[0];
    st->itr = st->itr + 1;
    st->last_i = i;
    conv = fabs(lerr[1]) < st->eps;
    if (conv) {
      st->suc = 1;
      st->done = 1;
      if (host_done) {
        *host_done = 1;
        __threadfence_system();
      }
    }
  }
  __syncthreads();
  double* hc = h + (size_t)ldh * i;
  for (int j = t; j <= i + 1; j += RED_BS) hc[j] = lh[j];
  for (int j = t; j <= i; j += RED_BS) coef[j] = sums[j];
}

// u_{i+1} = (u_{i+1} - sum_j coef_j u_j) * scale, sequentially in j (omp_sum_v, omp_mul_v)
__global__ __launch_bounds__(256) void k_gs_update(int n, double* __restrict__ u, size_t ld, int i,
                                                   const double* __restrict__ coef, const GmState* __restrict__ st)
{
#pragma clang fp contract(off)
  if (st->done) return;
  const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (size_t)n) return;
  double* w = u + (size_t)(i + 1) * ld;
  double v = w[k];
  for (int j = 0; j <= i; j++) v = v + (-coef[j]) * u[(size_t)j * ld + k];
  w[k] = st->scale * v;
}

// u0 = u0 / err0 (Array division, Array.h:793-805)
__global__ __launch_bounds__(256) void k_div(int n, double* __restrict__ u, const GmState* __restrict__ st)
{
  const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (size_t)n) return;
  u[k] = u[k] / st->scale;
}

// Back substitution, gmres.cpp:595-608; the triangle is staged in LDS when it fits.
constexpr int BS_LDS = 6144;  // doubles of the staged Hessenberg (48 KB)
__global__ __launch_bounds__(256) void k_backsolve(int sD, const double* __restrict__ h, const double* __restrict__ err,
                                                   double* __restrict__ y, GmState* st)
{
#pragma clang fp contract(off)
  __shared__ double lh[BS_LDS];
  __shared__ double ly[MAXM];
  int last_i = st->last_i;
  if (last_i >= sD) last_i = sD - 1;
  const int ldh = sD + 1;
  const int nc = last_i + 1;
  const bool staged = (size_t)nc * nc <= (size_t)BS_LDS;
  if (staged) {
    for (int q = threadIdx.x; q < nc * nc; q += blockDim.x) {
      int r = q % nc, c = q / nc;
      lh[q] = h[r + (size_t)ldh * c];
    }
  }
  for (int q = threadIdx.x; q < nc; q += blockDim.x) ly[q] = err[q];
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int j = last_i; j >= 0; j--) {
      double yj = ly[j];
      for (int k = j + 1; k <= last_i; k++) yj = yj - (staged ? lh[j + nc * k] : h[j + (size_t)ldh * k]) * ly[k];
      ly[j] = yj / (staged ? lh[j + nc * j] : h[j + (size_t)ldh * j]);
    }
    st->fNorm = fabs(err[last_i + 1]);
  }
  __syncthreads();
  for (int q = threadIdx.x; q < nc; q += blockDim.x) y[q] = ly[q];
}

// X += sum_j y_j u_j, sequentially in j (gmres.cpp:610-612)
__global__ __launch_bounds__(256) void k_xupdate(int n, double* __restrict__ X, const double* __restrict__ u, size_t ld,
                                                 const double* __restrict__ y, const GmState* __restrict__ st)
{
#pragma clang fp contract(off)
  const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (size_t)n) return;
  int last_i = st->last_i;
  double x = X[k];
  for (int j = 0; j <= last_i; j++) x = x + y[j] * u[(size_t)j * ld + k];
  X[k] = x;
}

// ------------------------------------------------------ preconditioning --
// W = diag(Val) -> 1 where zero -> 1/sqrt|W| -> * Dirichlet mask; R *= W.
template <int D>
__global__ __launch_bounds__(256) void k_pc_w(int nNo, const int* __restrict__ diagPtr, const double* __restrict__ Val,
                                              const double* __restrict__ mask, double* __restrict__ W,
                                              double* __restrict__ R)
{
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= D * nNo) return;
  const int a = gid / D, i = gid - a * D;
  double w = Val[(size_t)D * D * diagPtr[a] + i * D + i];
  if (w == 0.0) w = 1.0;
  w = 1.0 / sqrt(fabs(w));
  w = w * mask[gid];
  W[gid] = w;
  if (R) R[gid] = w * R[gid];
}

// Val(r*D+c, j) = (Val * W(r,row)) * W(c,col): pre_mul then pos_mul.
template <int D>
__global__ __launch_bounds__(256) void k_pc_val(int nNo, const int* __restrict__ rowPtr, const int* __restrict__ colPtr,
                                                const double* __restrict__ W, double* __restrict__ Val)
{
  constexpr int E = D * D;
  constexpr int LPR = E <= 16 ? 16 : 64;  // lanes per row
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  const int row = gid / LPR, l = gid - row * LPR;
  if (row >= nNo || l >= E) return;
  const int r = l / D, c = l - r * D;
  const double wr = W[(size_t)D * row + r];
  for (int j = rowPtr[row]; j < rowPtr[row + 1]; j++) {
    const int col = colPtr[j];
    double* v = Val + (size_t)E * j + l;
    *v = (*v * wr) * W[(size_t)D * col + c];
  }
}

__global__ __launch_bounds__(256) void k_mul(int n, const double* __restrict__ a, const double* __restrict__ b,
                                             double* __restrict__ out)
{
  const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (size_t)n) return;
  out[k] = a[k] * b[k];
}

// --------------------------------------------------------- workspace ----
static int ensure_work(svg_lhs* L, int dof, int sD)
{
  auto& k = L->kw;
  int n = dof * L->nNo;
  if (k.u && k.n == n && k.sD >= sD) return SVG_OK;
  dev_free(k.u);
  dev_free(k.X);
  dev_free(k.W);
  dev_free(k.h);
  dev_free(k.cs);
  dev_free(k.sn);
  dev_free(k.err);
  dev_free(k.coef);
  dev_free(k.y);
  dev_free(k.partial);
  dev_free(k.st);
  dev_free(k.scal);
  dev_free(k.ticket);
  if (k.h_st) (void)hipHostFree(k.h_st);
  k.h_st = nullptr;
  k.n = n;
  k.sD = sD;
  k.ld = ((size_t)n + 63) & ~(size_t)63;
  k.nblk = (int)((n + RED_SPAN - 1) / RED_SPAN);
  int rc;
  if ((rc = dev_alloc(&k.u, k.ld * (sD + 1))) || (rc = dev_alloc(&k.X, (size_t)n)) || (rc = dev_alloc(&k.W, (size_t)n)) ||
      (rc = dev_alloc(&k.h, (size_t)(sD + 1) * sD)) || (rc = dev_alloc(&k.cs, sD)) || (rc = dev_alloc(&k.sn, sD)) ||
      (rc = dev_alloc(&k.err, sD + 1)) || (rc = dev_alloc(&k.coef, sD + 2)) || (rc = dev_alloc(&k.y, sD)) ||
      (rc = dev_alloc(&k.partial, (size_t)k.nblk * (sD + 2))) || (rc = dev_alloc(&k.st, 1)) ||
      (rc = dev_alloc(&k.scal, 8)) || (rc = dev_alloc(&k.ticket, 1)))
    return rc;
  SVG_HIP(hipHostMalloc((void**)&k.h_st, sizeof(GmState), hipHostMallocDefault));
  if (!k.h_done) {
    SVG_HIP(hipHostMalloc((void**)&k.h_done, sizeof(int), hipHostMallocMapped | hipHostMallocCoherent));
    SVG_HIP(hipHostGetDevicePointer((void**)&k.d_done, k.h_done, 0));
    for (auto& e : k.ev) SVG_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  }
  SVG_HIP(hipMemset(k.h, 0, sizeof(double) * (sD + 1) * sD));
  SVG_HIP(hipMemset(k.st, 0, sizeof(GmState)));
  SVG_HIP(hipMemset(k.ticket, 0, sizeof(unsigned)));
  return SVG_OK;
}

// Scratch for the standalone dot/norm entry points (not tied to an lhs).
struct DotScratch {
  double* partial = nullptr;
  int cap = 0;
  double* out = nullptr;
  unsigned* ticket = nullptr;
};
static DotScratch g_dot;

static int dot_device(int n, const double* w, const double* base, size_t ld, int m, double* partial,
                      unsigned* ticket, double* out, GmState* st, FinArgs fa, double* h, double* cs, double* sn,
                      double* err, double* coef, hipStream_t s, int* host_done = nullptr)
{
  int nblk = (n + RED_SPAN - 1) / RED_SPAN;
  if (nblk < 1) nblk = 1;
  if (m > MAXM) return set_error(SVG_ERR_ARG, "reduction: too many sums");
  launch("mdot", k_mdot, dim3(nblk), dim3(RED_BS), (size_t)4 * m * sizeof(double), s, n, w, base, ld, m, partial,
         ticket, fa, st, h, cs, sn, err, coef, out, (volatile int*)host_done);
  return SVG_OK;
}

// gmres_v on the device; R holds the (preconditioned) right-hand side on
// entry and the solution on exit, as in the reference.
static int gmres_v_dev(svg_lhs* L, svg_subls* ls, int dof, const double* Val, double* R, hipStream_t s)
{
  if (dof != 4) return set_error(SVG_ERR_UNSUPPORTED, "gmres_v: only dof = 4 is built");
  if (ls->sD < 1 || ls->sD > 500) return set_error(SVG_ERR_ARG, "gmres_v: Krylov dimension must be in [1, 500]");
  int rc = ensure_work(L, dof, ls->sD);
  if (rc) return rc;
  auto& k = L->kw;
  const int n = k.n, sD = ls->sD;
  auto t0 = std::chrono::steady_clock::now();
  FinArgs fa{};
  fa.sD = sD;
  fa.absTol = ls->absTol;
  fa.relTol = ls->relTol;
  fa.mode = FIN_NORM_R;
  dot_device(n, R, R, 0, 1, k.partial, k.ticket, k.scal, k.st, fa, k.h, k.cs, k.sn, k.err, k.coef, s);
  SVG_HIP(hipMemcpyAsync(k.h_st, k.st, sizeof(GmState), hipMemcpyDeviceToHost, s));
  SVG_HIP(hipStreamSynchronize(s));
  ls->suc = 0;
  ls->iNorm = k.h_st->iNorm;
  ls->fNorm = k.h_st->fNorm;
  ls->itr = 0;
  if (ls->iNorm <= ls->absTol) {
    ls->callD = DBL_EPSILON;
    ls->dB = 0.0;
    return SVG_OK;  // R is left as the right-hand side (gmres.cpp:489-493)
  }
  svg_fill(k.X, (size_t)n, 0.0, s);
  const unsigned nbv = blocks_for(n, 256);
  double dB = ls->fNorm;
  for (int l = 0; l < ls->mItr; l++) {
    dB = ls->fNorm;
    // u0 = R - A X; err0 = |u0|; u0 /= err0
    spmv<4, 1>(L, Val, k.X, k.u, R, nullptr, s);
    fa.mode = FIN_RESTART;
    dot_device(n, k.u, k.u, 0, 1, k.partial, k.ticket, k.scal, k.st, fa, k.h, k.cs, k.sn, k.err, k.coef, s);
    launch("div", k_div, dim3(nbv), dim3(256), 0, s, n, k.u, (const GmState*)k.st);
    *k.h_done = 0;
    for (int i = 0; i < sD; i++) {
      // Keep at most LOOKAHEAD steps in flight: wait for step i-LOOKAHEAD, then
      // stop enqueuing once the device has flagged convergence.
      if (i >= LOOKAHEAD) {
        SVG_HIP(hipEventSynchronize(k.ev[(i - LOOKAHEAD) % 8]));
        if (*(volatile int*)k.h_done) break;
      }
      spmv<4, 0>(L, Val, k.u + k.ld * i, k.u + k.ld * (i + 1), nullptr, k.st, s);
      fa.mode = FIN_ARNOLDI;
      fa.i = i;
      dot_device(n, k.u + k.ld * (i + 1), k.u, k.ld, i + 2, k.partial, k.ticket, k.scal, k.st, fa, k.h, k.cs, k.sn,
                 k.err, k.coef, s, k.d_done);
      launch("gs_update", k_gs_update, dim3(nbv), dim3(256), 0, s, n, k.u, k.ld, i, (const double*)k.coef,
             (const GmState*)k.st);
      SVG_HIP(hipEventRecord(k.ev[i % 8], s));
    }
    launch("backsolve", k_backsolve, dim3(1), dim3(256), 0, s, sD, (const double*)k.h, (const double*)k.err, k.y, k.st);
    launch("xupdate", k_xupdate, dim3(nbv), dim3(256), 0, s, n, k.X, (const double*)k.u, k.ld, (const double*)k.y,
           (const GmState*)k.st);
    SVG_HIP(hipGetLastError());
    SVG_HIP(hipMemcpyAsync(k.h_st, k.st, sizeof(GmState), hipMemcpyDeviceToHost, s));
    SVG_HIP(hipStreamSynchronize(s));
    if (k.h_st->divzero) return set_error(SVG_ERR_ARG, "Array Divide by zero.");
    ls->itr = k.h_st->itr;
    ls->fNorm = k.h_st->fNorm;
    if (k.h_st->suc) {
      ls->suc = 1;
      break;
    }
  }
  SVG_HIP(hipMemcpyAsync(R, k.X, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
  auto t1 = std::chrono::steady_clock::now();
  ls->callD = std::chrono::duration<double>(t1 - t0).count();
  ls->dB = 10.0 * log(ls->fNorm / dB);
  return SVG_OK;
}

// Rebuilds the per-dof Dirichlet multiplier when the included faces change.
static int build_mask(svg_lhs* L, int dof)
{
  std::vector<int> key(L->nFaces);
  for (int f = 0; f < L->nFaces; f++) key[f] = L->face[f].incFlag ? 1 : 0;
  if (L->maskBuilt && key == L->maskKey) return SVG_OK;
  std::vector<double> m((size_t)dof * L->nNo, 1.0);
  for (int f = 0; f < L->nFaces; f++) {
    const auto& F = L->face[f];
    if (!F.created || !F.incFlag || F.bGrp != SVG_BC_DIR) continue;
    int nd = F.dof < dof ? F.dof : dof;
    for (int a = 0; a < F.nNo; a++)
      for (int i = 0; i < nd; i++) m[(size_t)dof * F.glob[a] + i] = m[(size_t)dof * F.glob[a] + i] * F.val[(size_t)F.dof * a + i];
  }
  SVG_HIP(hipMemcpy(L->d_mask, m.data(), sizeof(double) * m.size(), hipMemcpyHostToDevice));
  L->maskKey = key;
  L->maskBuilt = true;
  return SVG_OK;
}

static int precond_diag_dev(svg_lhs* L, int dof, double* Val, double* R, double* W, hipStream_t s)
{
  int rc = build_mask(L, dof);
  if (rc) return rc;
  const int n = dof * L->nNo;
  if (dof != 4) return set_error(SVG_ERR_UNSUPPORTED, "precond_diag: only dof = 4 is built");
  launch("pc_w", k_pc_w<4>, dim3(blocks_for(n, 256)), dim3(256), 0, s, L->nNo, (const int*)L->d_diagPtr,
         (const double*)Val, (const double*)L->d_mask, W, R);
  launch("pc_val", k_pc_val<4>, dim3(blocks_for((size_t)16 * L->nNo, 256)), dim3(256), 0, s, L->nNo,
         (const int*)L->d_rowPtr, (const int*)L->d_colPtr, (const double*)W, Val);
  SVG_HIP(hipGetLastError());
  return SVG_OK;
}

}  // namespace svg

using namespace svg;

extern "C" {

int svg_ls_create(svg_ls* ls, int LS_type)
{
  if (!ls) return set_error(SVG_ERR_ARG, "svg_ls_create: null ls");
  std::memset(ls, 0, sizeof *ls);
  ls->LS_type = LS_type;
  switch (LS_type) {
    case SVG_LS_NS:
      ls->RI.relTol = 0.4;
      ls->GM.relTol = 1.E-2;
      ls->CG.relTol = 0.2;
      ls->RI.mItr = 10;
      ls->GM.mItr = 2;
      ls->CG.mItr = 500;
      ls->GM.sD = 100;
      ls->RI.sD = 100;
      break;
    case SVG_LS_GMRES:
      ls->RI.relTol = 0.1;
      ls->RI.mItr = 4;
      ls->RI.sD = 250;
      break;
    case SVG_LS_CG:
      ls->RI.relTol = 1.E-2;
      ls->RI.mItr = 1000;
      break;
    case SVG_LS_BICGS:
      ls->RI.relTol = 1.E-2;
      ls->RI.mItr = 500;
      break;
    default:
      break;
  }
  ls->RI.absTol = 1.E-10;
  ls->GM.absTol = 1.E-10;
  ls->CG.absTol = 1.E-10;
  return SVG_OK;
}

int svg_spar_mul_vv(const svg_lhs* L, int dof, const double* K, const double* U, double* KU, void* stream)
{
  if (!L || !K || !U || !KU) return set_error(SVG_ERR_ARG, "svg_spar_mul_vv: null argument");
  hipStream_t s = (hipStream_t)stream;
  switch (dof) {
    case 1: spmv<1, 0>(L, K, U, KU, nullptr, nullptr, s); break;
    case 2: spmv<2, 0>(L, K, U, KU, nullptr, nullptr, s); break;
    case 3: spmv<3, 0>(L, K, U, KU, nullptr, nullptr, s); break;
    case 4: spmv<4, 0>(L, K, U, KU, nullptr, nullptr, s); break;
    default: return set_error(SVG_ERR_UNSUPPORTED, "svg_spar_mul_vv: dof must be 1..4");
  }
  SVG_HIP(hipGetLastError());
  return SVG_OK;
}

int svg_precond_diag(svg_lhs* L, int dof, double* Val, double* R, double* W, void* stream)
{
  if (!L || !Val || !W) return set_error(SVG_ERR_ARG, "svg_precond_diag: null argument");
  for (auto& f : L->face) f.incFlag = f.created;
  return precond_diag_dev(L, dof, Val, R, W, (hipStream_t)stream);
}

static int dot_api(int n, const double* U, const double* V, double* out, hipStream_t s)
{
  int nblk = (n + RED_SPAN - 1) / RED_SPAN;
  if (nblk < 1) nblk = 1;
  if (g_dot.cap < nblk) {
    dev_free(g_dot.partial);
    int rc = dev_alloc(&g_dot.partial, (size_t)nblk);
    if (rc) return rc;
    g_dot.cap = nblk;
  }
  if (!g_dot.out) {
    int rc = dev_alloc(&g_dot.out, 1);
    if (!rc) rc = dev_alloc(&g_dot.ticket, 1);
    if (rc) return rc;
    SVG_HIP(hipMemset(g_dot.ticket, 0, sizeof(unsigned)));
  }
  FinArgs fa{};
  fa.mode = FIN_DOT;
  dot_device(n, V, U, 0, 1, g_dot.partial, g_dot.ticket, g_dot.out, nullptr, fa, nullptr, nullptr, nullptr, nullptr,
             nullptr, s);
  SVG_HIP(hipGetLastError());
  SVG_HIP(hipMemcpyAsync(out, g_dot.out, sizeof(double), hipMemcpyDeviceToHost, s));
  SVG_HIP(hipStreamSynchronize(s));
  return SVG_OK;
}

int svg_dot_v(int dof, int nNo, const double* U, const double* V, double* out, void* stream)
{
  if (!U || !V || !out || dof <= 0 || nNo < 0) return set_error(SVG_ERR_ARG, "svg_dot_v: bad argument");
  return dot_api(dof * nNo, U, V, out, (hipStream_t)stream);
}

int svg_norm_v(int dof, int nNo, const double* U, double* out, void* stream)
{
  if (!U || !out || dof <= 0 || nNo < 0) return set_error(SVG_ERR_ARG, "svg_norm_v: bad argument");
  double d;
  int rc = dot_api(dof * nNo, U, U, &d, (hipStream_t)stream);
  if (rc) return rc;
  *out = sqrt(d);
  return SVG_OK;
}

int svg_fsils_solve(svg_lhs* L, svg_ls* ls, int dof, double* Ri, double* Val, int prec, int nIncL, const int* incL,
                    const double* res, void* stream)
{
  if (!L || !ls || !Ri || !Val) return set_error(SVG_ERR_ARG, "svg_fsils_solve: null argument");
  hipStream_t s = (hipStream_t)stream;
  // Face flags, solve.cpp:72-112.
  if (L->nFaces != 0) {
    for (auto& f : L->face) f.incFlag = true;
    if (nIncL != 0) {
      if (!incL || nIncL < L->nFaces) return set_error(SVG_ERR_ARG, "svg_fsils_solve: incL shorter than nFaces");
      for (int f = 0; f < L->nFaces; f++)
        if (incL[f] == 0) L->face[f].incFlag = false;
    }
    bool neu = false;
    for (auto& f : L->face)
      if (f.bGrp == SVG_BC_NEU) {
        neu = true;
        break;
      }
    if (!res && neu) return set_error(SVG_ERR_NEU_RES, "[fsils_solve] res is required for Neu surfaces");
    for (int f = 0; f < L->nFaces; f++) {
      auto& F = L->face[f];
      F.coupledFlag = false;
      if (!F.incFlag) continue;
      if (F.bGrp == SVG_BC_NEU && res[f] != 0.0) {
        return set_error(SVG_ERR_UNSUPPORTED, "fsils_solve: coupled (resistance) Neumann faces are not built");
      }
    }
  }
  if (dof != 4) return set_error(SVG_ERR_UNSUPPORTED, "fsils_solve: only dof = 4 is built");
  int rc;
  int sDneed = ls->LS_type == SVG_LS_GMRES ? ls->RI.sD : 1;
  if ((rc = ensure_work(L, dof, sDneed < 1 ? 1 : sDneed))) return rc;
  auto& k = L->kw;
  const int n = dof * L->nNo;
  if (prec == SVG_PREC_FSILS) {
    if ((rc = precond_diag_dev(L, dof, Val, Ri, k.W, s))) return rc;
  } else if (prec == SVG_PREC_NONE) {
    svg_fill(k.W, (size_t)n, 1.0, s);
  } else {
    return set_error(SVG_ERR_UNSUPPORTED, "fsils_solve: preconditioner not built");
  }
  switch (ls->LS_type) {
    case SVG_LS_GMRES:
      rc = gmres_v_dev(L, &ls->RI, dof, Val, Ri, s);
      break;
    case SVG_LS_NS:
      rc = set_error(SVG_ERR_UNSUPPORTED, "fsils_solve: NS solver not built yet");
      break;
    case SVG_LS_CG:
    case SVG_LS_BICGS:
      rc = set_error(SVG_ERR_UNSUPPORTED, "fsils_solve: CG/BICGS are not part of the fluid path");
      break;
    default:
      rc = set_error(SVG_ERR_LSTYPE, "FSILS: LS_type not defined");
  }
  if (rc) return rc;
  launch("wc_scale", k_mul, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, (const double*)k.W, (const double*)Ri, Ri);
  SVG_HIP(hipGetLastError());
  SVG_HIP(hipStreamSynchronize(s));
  return SVG_OK;
}

}  // extern "C"
