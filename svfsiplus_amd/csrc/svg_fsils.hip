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

#include <algorithm>
#include <chrono>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <mutex>

namespace svg {

constexpr int RED_BS = 256;   // threads per reduction workgroup (4 wavefronts)
constexpr int RED_EPT = 4;    // elements per thread, at stride RED_BS inside the workgroup's span
constexpr int RED_NW = RED_BS / 64;
constexpr int RED_SPAN = RED_BS * RED_EPT;
constexpr int LOOKAHEAD = 3;  // Arnoldi steps enqueued ahead of the host's convergence check

// ------------------------------------------------------------------ SpMV --

// KU = K U (MODE 0) or KU = B - K U (MODE 1) for BR x BC blocks stored
// row-major (K[BR*BC*j + BC*i + c]): vv is BR = BC = dof, sv is BR = dof,
// BC = 1, vs is BR = 1, BC = dof, ss is 1 x 1 (spar_mul.cpp:46-258). One lane
// per (row, i) accumulates the reference's left-to-right sum; UNR blocks'
// loads are issued before their adds, NT streams the matrix with non-temporal
// loads, xcd remaps workgroups to contiguous row bands per XCD. Every variant
// performs the same additions in the same order.
template <int BR, int BC, int MODE, int UNR, bool NT>
__global__ __launch_bounds__(256) void k_spmv(int nNo, const int* __restrict__ rowPtr, const int* __restrict__ colPtr,
                                              const double* __restrict__ K, const double* __restrict__ U,
                                              double* __restrict__ KU, const double* __restrict__ B,
                                              const int* __restrict__ skip, int xcd)
{
#pragma clang fp contract(off)
  if (skip && *skip) return;
  int bx = blockIdx.x;
  if (xcd) {
    bx = xcd_block(blockIdx.x, gridDim.x);
    if (bx < 0) return;
  }
  const int gid = bx * blockDim.x + threadIdx.x;
  const int row = gid / BR;
  const int i = gid - row * BR;
  if (row >= nNo) return;
  const int s = rowPtr[row], e = rowPtr[row + 1];
  double acc = 0.0;
  int j = s;
  for (; j + UNR <= e; j += UNR) {
    int c[UNR];
    double kv[UNR][BC], uv[UNR][BC];
#pragma unroll
    for (int q = 0; q < UNR; q++) c[q] = NT ? __builtin_nontemporal_load(colPtr + j + q) : colPtr[j + q];
#pragma unroll
    for (int q = 0; q < UNR; q++) {
      const double* kp = K + (size_t)BR * BC * (j + q) + BC * i;
      if constexpr (BC == 4) {
        double2 a0, a1;
        if (NT) {
          a0.x = __builtin_nontemporal_load(kp + 0);
          a0.y = __builtin_nontemporal_load(kp + 1);
          a1.x = __builtin_nontemporal_load(kp + 2);
          a1.y = __builtin_nontemporal_load(kp + 3);
        } else {
          a0 = reinterpret_cast<const double2*>(kp)[0];
          a1 = reinterpret_cast<const double2*>(kp)[1];
        }
        kv[q][0] = a0.x;
        kv[q][1] = a0.y;
        kv[q][2] = a1.x;
        kv[q][3] = a1.y;
        const double2 u0 = reinterpret_cast<const double2*>(U + (size_t)BC * c[q])[0];
        const double2 u1 = reinterpret_cast<const double2*>(U + (size_t)BC * c[q])[1];
        uv[q][0] = u0.x;
        uv[q][1] = u0.y;
        uv[q][2] = u1.x;
        uv[q][3] = u1.y;
      } else {
#pragma unroll
        for (int cc = 0; cc < BC; cc++) {
          kv[q][cc] = NT ? __builtin_nontemporal_load(kp + cc) : kp[cc];
          uv[q][cc] = U[(size_t)BC * c[q] + cc];
        }
      }
    }
#pragma unroll
    for (int q = 0; q < UNR; q++) {
      double t = acc + kv[q][0] * uv[q][0];
#pragma unroll
      for (int cc = 1; cc < BC; cc++) t = t + kv[q][cc] * uv[q][cc];
      acc = t;
    }
  }
  for (; j < e; j++) {
    const int col = colPtr[j];
    const double* k = K + (size_t)BR * BC * j + BC * i;
    const double* u = U + (size_t)BC * col;
    double t = acc + k[0] * u[0];
#pragma unroll
    for (int cc = 1; cc < BC; cc++) t = t + k[cc] * u[cc];
    acc = t;
  }
  if (MODE == 0)
    KU[gid] = acc;
  else
    KU[gid] = B[gid] - acc;
}

// Launch variants of the 4x4 SpMV chosen by the autotuner (svg_spmv_tune).
struct SpmvVariant {
  int unroll;
  bool nt;
  bool xcd;
};
static const SpmvVariant kSpmvVariants[] = {{1, false, false}, {2, false, false}, {4, false, false},
                                            {1, false, true},  {2, false, true},  {4, false, true},
                                            {2, true, false},  {4, true, false},  {2, true, true},
                                            {4, true, true},   {8, false, false}, {8, false, true}};
constexpr int kNumSpmvVariants = sizeof(kSpmvVariants) / sizeof(kSpmvVariants[0]);
constexpr int kDefaultSpmvVariant = 0;

template <int BR, int BC, int MODE, int UNR, bool NT>
static void spmv_launch(const svg_lhs* L, const double* K, const double* U, double* KU, const double* B,
                        const int* skip, hipStream_t s, const char* name, bool xcd)
{
  size_t nt = (size_t)BR * L->nNo;
  unsigned nb = blocks_for(nt, 256);
  const int use_xcd = (xcd && nb >= 64) ? 1 : 0;
  if (use_xcd) nb = (nb + 7) & ~7u;
  const double bytes = (8.0 * BR * BC + 4.0) * L->nnz + (4.0 + 8.0 * BC + 8.0 * BR * (MODE ? 2 : 1)) * L->nNo;
  launch_b(name, bytes, k_spmv<BR, BC, MODE, UNR, NT>, dim3(nb), dim3(256), 0, s, L->nNo, (const int*)L->d_rowPtr,
           (const int*)L->d_colPtr, K, U, KU, B, skip, use_xcd);
}

template <int BR, int BC, int MODE>
static void spmv_variant(int v, const svg_lhs* L, const double* K, const double* U, double* KU, const double* B,
                         const int* skip, hipStream_t s, const char* name)
{
  const SpmvVariant& sv = kSpmvVariants[v];
  switch (sv.unroll * 2 + (sv.nt ? 1 : 0)) {
    case 2: spmv_launch<BR, BC, MODE, 1, false>(L, K, U, KU, B, skip, s, name, sv.xcd); break;
    case 4: spmv_launch<BR, BC, MODE, 2, false>(L, K, U, KU, B, skip, s, name, sv.xcd); break;
    case 5: spmv_launch<BR, BC, MODE, 2, true>(L, K, U, KU, B, skip, s, name, sv.xcd); break;
    case 8: spmv_launch<BR, BC, MODE, 4, false>(L, K, U, KU, B, skip, s, name, sv.xcd); break;
    case 9: spmv_launch<BR, BC, MODE, 4, true>(L, K, U, KU, B, skip, s, name, sv.xcd); break;
    default: spmv_launch<BR, BC, MODE, 8, false>(L, K, U, KU, B, skip, s, name, sv.xcd); break;
  }
}

template <int BR, int BC, int MODE>
static void spmv(const svg_lhs* L, const double* K, const double* U, double* KU, const double* B, const int* skip,
                 hipStream_t s, const char* name = "spmv_bcsr")
{
  if (BR == 4 && BC == 4) {
    spmv_variant<BR, BC, MODE>(L->spmvVariant >= 0 ? L->spmvVariant : kDefaultSpmvVariant, L, K, U, KU, B, skip, s,
                               name);
  } else {
    spmv_launch<BR, BC, MODE, 2, false>(L, K, U, KU, B, skip, s, name, true);
  }
}

// ------------------------------------------------------ mixed precision --
// SpMV of the mixed-precision solve: W = K32 U32 with FP32 copies of the
// scaled matrix and of the current Arnoldi vector, FP32 accumulation in the
// reference's block order, one lane per (row, block-row i), W stored in FP64.
template <int UNR, typename TW = double>
__device__ __forceinline__ void spmv32_row(int row, int i, const int* __restrict__ rowPtr,
                                           const int* __restrict__ colPtr, const float* __restrict__ K,
                                           const float* __restrict__ U, TW* __restrict__ W, size_t k)
{
  const int s = rowPtr[row], e = rowPtr[row + 1];
  float acc = 0.0f;
  int j = s;
  for (; j + UNR <= e; j += UNR) {
    int c[UNR];
    float4 a[UNR], u[UNR];
#pragma unroll
    for (int q = 0; q < UNR; q++) c[q] = colPtr[j + q];
#pragma unroll
    for (int q = 0; q < UNR; q++) {
      a[q] = reinterpret_cast<const float4*>(K + 16 * (size_t)(j + q) + 4 * i)[0];
      u[q] = reinterpret_cast<const float4*>(U + 4 * (size_t)c[q])[0];
    }
#pragma unroll
    for (int q = 0; q < UNR; q++) {
      acc = __builtin_fmaf(a[q].x, u[q].x, acc);
      acc = __builtin_fmaf(a[q].y, u[q].y, acc);
      acc = __builtin_fmaf(a[q].z, u[q].z, acc);
      acc = __builtin_fmaf(a[q].w, u[q].w, acc);
    }
  }
  for (; j < e; j++) {
    const float4 a = reinterpret_cast<const float4*>(K + 16 * (size_t)j + 4 * i)[0];
    const float4 u = reinterpret_cast<const float4*>(U + 4 * (size_t)colPtr[j])[0];
    acc = __builtin_fmaf(a.x, u.x, acc);
    acc = __builtin_fmaf(a.y, u.y, acc);
    acc = __builtin_fmaf(a.z, u.z, acc);
    acc = __builtin_fmaf(a.w, u.w, acc);
  }
  W[k] = (TW)acc;
}

// SpMV of the mixed-precision solve: W = K32 U32 with FP32 copies of the
// scaled matrix and of the current Arnoldi vector, FP32 accumulation in the
// reference's block order, one lane per (row, block-row i), W stored in FP64
// or, where no FP64 term is added to it, in FP32 (the same value: the sum is
// an FP32 number).
template <int UNR, typename TW = double>
__global__ __launch_bounds__(256) void k_spmv32(int nNo, const int* __restrict__ rowPtr,
                                                const int* __restrict__ colPtr, const float* __restrict__ K,
                                                const float* __restrict__ U, TW* __restrict__ W,
                                                const int* __restrict__ skip, int xcd)
{
  if (skip && *skip) return;
  int bx = blockIdx.x;
  if (xcd) {
    bx = xcd_block(blockIdx.x, gridDim.x);
    if (bx < 0) return;
  }
  const size_t k = (size_t)bx * 256 + threadIdx.x;
  const int row = (int)(k >> 2), i = (int)(k & 3);
  if (row >= nNo) return;
  spmv32_row<UNR, TW>(row, i, rowPtr, colPtr, K, U, W, k);
}

// The batched solve's FP32 SpMV: rows of systems that stopped iterating are skipped.
template <int UNR>
__global__ __launch_bounds__(256) void k_bspmv32(int nNo, const int* __restrict__ rowPtr,
                                                 const int* __restrict__ colPtr, const float* __restrict__ K,
                                                 const float* __restrict__ U, double* __restrict__ W,
                                                 const int* __restrict__ sysOfNode, const GmState* __restrict__ st,
                                                 int xcd)
{
  int bx = blockIdx.x;
  if (xcd) {
    bx = xcd_block(blockIdx.x, gridDim.x);
    if (bx < 0) return;
  }
  const size_t k = (size_t)bx * 256 + threadIdx.x;
  const int row = (int)(k >> 2), i = (int)(k & 3);
  if (row >= nNo) return;
  const GmState& S = st[sysOfNode[row]];
  if (!S.act || S.done) return;
  spmv32_row<UNR>(row, i, rowPtr, colPtr, K, U, W, k);
}

__global__ __launch_bounds__(256) void k_d2f(size_t n, const double* __restrict__ a, float* __restrict__ b)
{
  size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; k < n; k += stride) b[k] = (float)a[k];
}

// Launch variants of the FP32 block SpMV (unroll depth, per-XCD row bands),
// chosen per lhs by spmv32_tune; every variant computes the same sums in the
// same order. Variant 0 is the untuned default.
struct Spmv32Variant {
  int unroll;
  bool xcd;
};
static const Spmv32Variant kSpmv32Variants[] = {{2, false}, {1, false}, {4, false}, {2, true}, {1, true}, {4, true}};
constexpr int kNumSpmv32Variants = sizeof(kSpmv32Variants) / sizeof(kSpmv32Variants[0]);

template <int UNR>
static void spmv32_launch(const svg_lhs* L, const float* K32, const float* U32, double* W, const int* skip,
                          const int* sysOfNode, const GmState* st, bool xcd, hipStream_t s, const char* name)
{
  unsigned nb = blocks_for((size_t)4 * L->nNo, 256);
  const int use_xcd = (xcd && nb >= 64) ? 1 : 0;
  if (use_xcd) nb = (nb + 7) & ~7u;
  const double bytes = 68.0 * L->nnz + 52.0 * L->nNo;  // block + col; rowPtr, U (FP32), W (FP64)
  if (sysOfNode)
    launch_b(name, bytes, k_bspmv32<UNR>, dim3(nb), dim3(256), 0, s, L->nNo, (const int*)L->d_rowPtr,
             (const int*)L->d_colPtr, K32, U32, W, sysOfNode, st, use_xcd);
  else
    launch_b(name, bytes, k_spmv32<UNR>, dim3(nb), dim3(256), 0, s, L->nNo, (const int*)L->d_rowPtr,
             (const int*)L->d_colPtr, K32, U32, W, skip, use_xcd);
}

// W = K32 U32 with the lhs's tuned variant; sysOfNode/st: only the rows of
// batched systems still iterating.
static void spmv32(const svg_lhs* L, const float* K32, const float* U32, double* W, const int* skip,
                   hipStream_t s, const char* name = "spmv32", const int* sysOfNode = nullptr,
                   const GmState* st = nullptr, int variant = -1)
{
  const int v = variant >= 0 ? variant : (L->spmv32Variant >= 0 ? L->spmv32Variant : 0);
  const Spmv32Variant& sv = kSpmv32Variants[v];
  switch (sv.unroll) {
    case 1: spmv32_launch<1>(L, K32, U32, W, skip, sysOfNode, st, sv.xcd, s, name); break;
    case 4: spmv32_launch<4>(L, K32, U32, W, skip, sysOfNode, st, sv.xcd, s, name); break;
    default: spmv32_launch<2>(L, K32, U32, W, skip, sysOfNode, st, sv.xcd, s, name); break;
  }
}

// W = K32 U32 with an FP32 W (the lhs's tuned variant).
static void spmv32_f(const svg_lhs* L, const float* K32, const float* U32, float* W, const int* skip, hipStream_t s)
{
  const Spmv32Variant& sv = kSpmv32Variants[L->spmv32Variant >= 0 ? L->spmv32Variant : 0];
  unsigned nb = blocks_for((size_t)4 * L->nNo, 256);
  const int use_xcd = (sv.xcd && nb >= 64) ? 1 : 0;
  if (use_xcd) nb = (nb + 7) & ~7u;
  const double bytes = 68.0 * L->nnz + 36.0 * L->nNo;  // block + col; rowPtr, U (FP32), W (FP32)
  switch (sv.unroll) {
    case 1:
      launch_b("spmv32", bytes, k_spmv32<1, float>, dim3(nb), dim3(256), 0, s, L->nNo, (const int*)L->d_rowPtr,
               (const int*)L->d_colPtr, K32, U32, W, skip, use_xcd);
      break;
    case 4:
      launch_b("spmv32", bytes, k_spmv32<4, float>, dim3(nb), dim3(256), 0, s, L->nNo, (const int*)L->d_rowPtr,
               (const int*)L->d_colPtr, K32, U32, W, skip, use_xcd);
      break;
    default:
      launch_b("spmv32", bytes, k_spmv32<2, float>, dim3(nb), dim3(256), 0, s, L->nNo, (const int*)L->d_rowPtr,
               (const int*)L->d_colPtr, K32, U32, W, skip, use_xcd);
      break;
  }
}

// The autotuners' choice: the fastest variant, except that a variant that
// walks its rows in per-XCD bands wins over a faster one when it is within
// kTuneTie of it. Timings of the fastest few variants differ by less than the
// launch-to-launch spread, while the banded variants move measurably fewer
// bytes through HBM (each XCD's L2 keeps the vector entries its rows gather;
// measured on the 1.2M-tet pipe: FP64 1.03-1.07x the algorithmic bytes banded
// against 1.11-1.15x round robin, FP32 1.13-1.19x against 1.24-1.32x).
constexpr double kTuneTie = 0.02;
static int pick_variant(const double* t, const bool* low_traffic, int n)
{
  int best = 0;
  for (int v = 1; v < n; v++)
    if (t[v] < t[best]) best = v;
  int pick = best;
  for (int v = 0; v < n; v++)
    if (low_traffic[v] && t[v] <= t[best] * (1.0 + kTuneTie) && (!low_traffic[pick] || t[v] < t[pick])) pick = v;
  return pick;
}

// The autotuner of the FP32 SpMV: times every variant (one warm launch, then
// 3) on K32 with the probe vector U32 (zeroed here) into W, keeps the fastest.
static int spmv32_tune(svg_lhs* L, const float* K32, float* U32, double* W, hipStream_t s)
{
  SVG_HIP(hipMemsetAsync(U32, 0, sizeof(float) * 4 * L->nNo, s));
  hipEvent_t a, b;
  SVG_HIP(hipEventCreate(&a));
  SVG_HIP(hipEventCreate(&b));
  double t[kNumSpmv32Variants];
  bool banded[kNumSpmv32Variants];
  for (int v = 0; v < kNumSpmv32Variants; v++) {
    spmv32(L, K32, U32, W, nullptr, s, "spmv32_tune", nullptr, nullptr, v);
    SVG_HIP(hipEventRecord(a, s));
    for (int r = 0; r < 3; r++) spmv32(L, K32, U32, W, nullptr, s, "spmv32_tune", nullptr, nullptr, v);
    SVG_HIP(hipEventRecord(b, s));
    SVG_HIP(hipEventSynchronize(b));
    float ms = 0.f;
    SVG_HIP(hipEventElapsedTime(&ms, a, b));
    t[v] = ms;
    banded[v] = kSpmv32Variants[v].xcd;
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  SVG_HIP(hipGetLastError());
  L->spmv32Variant = pick_variant(t, banded, kNumSpmv32Variants);
  return SVG_OK;
}

// ------------------------------------------------------------ reductions --
enum FinMode { FIN_DOT = 0, FIN_NORM_R = 1, FIN_RESTART = 2, FIN_ARNOLDI = 3 };

struct FinArgs {
  int mode;
  int i;      // Arnoldi column
  int sD;
  int inc;    // FIN_RESTART: itr increment (the first cycle of the inner GMRES adds none)
  int first;  // FIN_RESTART: first cycle of the inner GMRES sets iNorm and eps from err(0)
  int tcheck; // FIN_RESTART: stop when the (FP64, true) restart residual is below eps
  int fbwin;  // FIN_ARNOLDI: fallback after this many non-decreasing iterations of |err| (0: off)
  double absTol, relTol;
};

constexpr int MAXM = 512;  // sums per reduction (Krylov dimension + 2)
constexpr double FB_DROP = 1e-4;  // learned preconditioner: smallest relative decrease that counts

// Pass 1 of the dots of w against m vectors base + j*ld (j < m) and, with
// self, w.w as column m, over one span [k0, kend) of 1024 elements: each
// lane sums its 4 elements (stride 256) in order, each wavefront butterflies,
// and lane 0 of wavefront wv leaves column j's sum in lds[wv * mc + j].
// Loads are unconditional (indices clamped into the span; the clamped values
// are not used) and G columns are loaded and reduced together, so one memory
// latency is spent per G columns (FP32 columns, half the bytes of FP64 ones,
// need more in flight to keep HBM busy).
// Basis entries of the dots. The FP32 (compressed) basis of the mixed and
// flexible solves is streamed with non-temporal loads (and written with
// non-temporal stores by the update): its passes then leave the memory-side
// cache to the FP32 block matrix, which the next Arnoldi step's SpMV reads
// again (in-solve SpMV 45 -> 39 us on the 1.2M-tet pipe). The FP64 basis of
// the plain solve keeps ordinary loads (its matrix does not fit that cache).
__device__ __forceinline__ float basis_load(const float* p) { return __builtin_nontemporal_load(p); }
__device__ __forceinline__ double basis_load(const double* p) { return *p; }

template <int G, typename T, typename TWv = double>
__device__ __forceinline__ void span_dots(const TWv* __restrict__ w, const T* __restrict__ base, size_t ld, int m,
                                          bool self, size_t k0, size_t kend, double* lds, int mc)
{
#pragma clang fp contract(off)
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (kend == 0) {  // an empty vector (no element to clamp to): every sum is 0
    if (lane == 0)
      for (int j = 0; j < mc; j++) lds[wv * mc + j] = 0.0;
    return;
  }
  size_t kc[RED_EPT];
  bool in[RED_EPT];
  double wr[RED_EPT];
#pragma unroll
  for (int r = 0; r < RED_EPT; r++) {
    const size_t k = k0 + (size_t)r * RED_BS;
    in[r] = k < kend;
    kc[r] = in[r] ? k : kend - 1;
    const double v = (double)w[kc[r]];
    wr[r] = in[r] ? v : 0.0;
  }
  int j = 0;
  for (; G > 1 && j + G <= m; j += G) {
    T c[G][RED_EPT];
#pragma unroll
    for (int q = 0; q < G; q++)
#pragma unroll
      for (int r = 0; r < RED_EPT; r++) c[q][r] = basis_load(base + (size_t)(j + q) * ld + kc[r]);
    double acc[G];
#pragma unroll
    for (int q = 0; q < G; q++) {
      acc[q] = 0.0;
#pragma unroll
      for (int r = 0; r < RED_EPT; r++) {
        const double v = in[r] ? (double)c[q][r] : 0.0;
        acc[q] = acc[q] + v * wr[r];
      }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
#pragma unroll
      for (int q = 0; q < G; q++) acc[q] += __shfl_xor(acc[q], o, 64);  // wave_sum's butterfly
    if (lane == 0) {
#pragma unroll
      for (int q = 0; q < G; q++) lds[wv * mc + j + q] = acc[q];
    }
  }
  for (; j < m; j++) {
    T c[RED_EPT];
#pragma unroll
    for (int r = 0; r < RED_EPT; r++) c[r] = basis_load(base + (size_t)j * ld + kc[r]);
    double acc = 0.0;
#pragma unroll
    for (int r = 0; r < RED_EPT; r++) {
      const double v = in[r] ? (double)c[r] : 0.0;
      acc = acc + v * wr[r];
    }
    acc = wave_sum(acc);
    if (lane == 0) lds[wv * mc + j] = acc;
  }
  if (self) {
    double acc = 0.0;
#pragma unroll
    for (int r = 0; r < RED_EPT; r++) acc = acc + wr[r] * wr[r];
    acc = wave_sum(acc);
    if (lane == 0) lds[wv * mc + m] = acc;
  }
}

constexpr int MDOT_CB_G = 4;  // FP32 basis columns per load batch (4: 24.1 us, 8: 25.1 us on the 1.2M-tet pipe)

// Dots of w against m vectors base + j*ld (j < m), pass 1: each workgroup
// sums its 1024-element span (256 lanes x 4 elements at stride 256, in order),
// butterflies inside each wavefront and adds the 4 wavefront sums in order;
// partial[j*nblk + b].
__global__ __launch_bounds__(RED_BS) void k_mdot(int n, const double* __restrict__ w, const double* __restrict__ base,
                                                 size_t ld, int m, double* __restrict__ partial,
                                                 const int* __restrict__ skip)
{
#pragma clang fp contract(off)
  extern __shared__ double lds[];  // [RED_NW waves][m]
  if (skip && *skip) return;
  const int t = threadIdx.x;
  const int nblk = gridDim.x;
  span_dots<1>(w, base, ld, m, false, (size_t)blockIdx.x * RED_SPAN + t, (size_t)n, lds, m);
  __syncthreads();
  for (int j = t; j < m; j += RED_BS) {
    double v = lds[j];
    for (int q = 1; q < RED_NW; q++) v = v + lds[q * m + j];
    partial[(size_t)j * nblk + blockIdx.x] = v;
  }
}

// The mixed-precision solve keeps its Krylov basis in FP32 (a compressed
// basis; the arithmetic stays FP64): dots of w against the m FP32 vectors
// base + j*ld, then w.w as column m, in k_mdot's order; partial[j*nblk + b].
template <typename TWv>
__global__ __launch_bounds__(RED_BS) void k_mdot_cb(int n, const TWv* __restrict__ w, const float* __restrict__ base,
                                                    size_t ld, int m, double* __restrict__ partial,
                                                    const int* __restrict__ skip)
{
#pragma clang fp contract(off)
  extern __shared__ double lds[];  // [RED_NW waves][m + 1]
  if (skip && *skip) return;
  const int t = threadIdx.x;
  const int nblk = gridDim.x, mc = m + 1;
  span_dots<MDOT_CB_G>(w, base, ld, m, true, (size_t)blockIdx.x * RED_SPAN + t, (size_t)n, lds, mc);
  __syncthreads();
  for (int j = t; j < mc; j += RED_BS) {
    double v = lds[j];
    for (int q = 1; q < RED_NW; q++) v = v + lds[q * mc + j];
    partial[(size_t)j * nblk + blockIdx.x] = v;
  }
}

constexpr int FIN_BS = 1024;  // finaliser: one workgroup of 16 wavefronts

// Pass 2 (one workgroup): the partials of every column summed in a fixed
// order (per lane at stride 64, then a butterfly), then the tail selected by
// A.mode:
//   FIN_DOT      out[0] = sum
//   FIN_NORM_R   iNorm = fNorm = sqrt(sum), eps = max(absTol, relTol iNorm)   (gmres.cpp:482-487)
//   FIN_RESTART  err(0) = sqrt(sum), itr += inc; the inner GMRES's first
//                cycle also sets iNorm and eps from err(0) (gmres.cpp:125-148)
//   FIN_ARNOLDI  Hessenberg column i, Pythagorean h(i+1,i), Givens, err, stop (gmres.cpp:550-592)
struct FinLds {
  double sums[MAXM];
  double lh[MAXM], lcs[MAXM], lsn[MAXM];
  double pre_err, pre_eps;
  int pre_itr;
};

// The finaliser's work for one system: column j's partials at
// partial[j * pstride + b], b < nblk.
__device__ __forceinline__ void fin_body(const FinArgs& A, const double* __restrict__ partial, size_t pstride,
                                         int nblk, int m, GmState* st, double* __restrict__ h,
                                         double* __restrict__ cs, double* __restrict__ sn, double* __restrict__ err,
                                         double* __restrict__ coef, double* __restrict__ out,
                                         volatile int* host_done, FinLds& S)
{
#pragma clang fp contract(off)
  double* sums = S.sums;
  double* lh = S.lh;
  double* lcs = S.lcs;
  double* lsn = S.lsn;
  double& pre_err = S.pre_err;
  double& pre_eps = S.pre_eps;
  int& pre_itr = S.pre_itr;
  if (A.mode == FIN_ARNOLDI && st->done) return;
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  if (A.mode == FIN_ARNOLDI) {
    if (t == 64) pre_err = err[A.i];
    if (t == 128) pre_eps = st->eps;
    if (t == 192) pre_itr = st->itr;
  }
  for (int j = wv; j < m; j += FIN_BS / 64) {
    const double* pj = partial + (size_t)j * pstride;
    double acc = 0.0;
    // the lane's partials b = lane, lane + 64, ... added in order; loaded 16
    // at a time (indices clamped, clamped values unused) so that the loads
    // overlap instead of costing one memory latency each
    for (int b0 = 0; b0 < nblk; b0 += 16 * 64) {
      double v[16];
#pragma unroll
      for (int q = 0; q < 16; q++) {
        const int b = b0 + lane + 64 * q;
        v[q] = pj[b < nblk ? b : nblk - 1];
      }
#pragma unroll
      for (int q = 0; q < 16; q++)
        if (b0 + lane + 64 * q < nblk) acc = acc + v[q];
    }
    acc = wave_sum(acc);
    if (lane == 0) sums[j] = acc;
  }
  if (A.mode == FIN_ARNOLDI) {
    for (int j = t; j < A.i; j += FIN_BS) {
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
      st->early = 0;
      st->tconv = 0;
      st->last_i = 0;
      st->fallback = 0;
      st->nondec = 0;
      st->best = DBL_MAX;
      st->act = nrm > A.absTol ? 1 : 0;
      st->cyc = 0;
      st->fbitr = -1;
      st->fStart = nrm;
    }
    return;
  }
  if (A.mode == FIN_RESTART) {
    if (t == 0) {
      double e0 = sqrt(sums[0]);
      err[0] = e0;
      st->itr = st->itr + A.inc;
      st->done = 0;
      st->scale = e0;
      if (A.first) {
        st->early = 0;
        if (e0 <= A.absTol) {
          st->early = 1;
          st->done = 1;
          if (host_done) {
            *host_done = 1;
            __threadfence_system();
          }
          return;
        }
        st->iNorm = e0;
        st->fNorm = e0;
        st->eps = fmax(A.absTol, A.relTol * e0);
      }
      st->tconv = 0;
      if (A.tcheck && e0 < st->eps) {
        st->tconv = 1;
        st->suc = 1;
        st->done = 1;
        st->fNorm = e0;
      } else if (e0 == 0.0) {
        st->divzero = 1;
        st->done = 1;
      }
      // a cycle that starts done is skipped by every launch: the host stops
      // enqueueing it at its first look at the flag
      if (st->done && host_done) {
        *host_done = 1;
        __threadfence_system();
      }
    }
    return;
  }
  // FIN_ARNOLDI: serial part on LDS by one lane, then parallel stores
  const int i = A.i, ldh = A.sD + 1;
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
    double ei = pre_err;
    double e1 = -sv * ei;
    err[i + 1] = e1;
    err[i] = c * ei;
    st->itr = pre_itr + 1;
    st->last_i = i;
    if (fabs(e1) < pre_eps) {
      st->suc = 1;
      st->done = 1;
      if (host_done) {
        *host_done = 1;
        __threadfence_system();
      }
    } else if (A.fbwin > 0 && !st->fallback) {
      // an iteration counts as decreasing when it lowers the smallest residual
      // estimate so far by at least FB_DROP (relative); a system that fell back
      // keeps iterating without the check
      if (fabs(e1) < st->best * (1.0 - FB_DROP)) {
        st->best = fabs(e1);
        st->nondec = 0;
      } else if (++st->nondec >= A.fbwin) {
        st->fallback = 1;
        st->done = 1;
        if (host_done) {
          *host_done = 1;
          __threadfence_system();
        }
      }
    }
  }
  __syncthreads();
  double* hc = h + (size_t)ldh * i;
  for (int j = t; j <= i + 1; j += FIN_BS) hc[j] = lh[j];
  for (int j = t; j <= i; j += FIN_BS) coef[j] = sums[j];
}

__global__ __launch_bounds__(FIN_BS) void k_fin(FinArgs A, const double* __restrict__ partial, int nblk, int m,
                                                GmState* st, double* __restrict__ h, double* __restrict__ cs,
                                                double* __restrict__ sn, double* __restrict__ err,
                                                double* __restrict__ coef, double* __restrict__ out,
                                                volatile int* host_done)
{
  __shared__ FinLds S;
  fin_body(A, partial, (size_t)nblk, nblk, m, st, h, cs, sn, err, coef, out, host_done, S);
}

// u_{i+1} = (u_{i+1} - sum_j coef_j u_j) * scale, sequentially in j (omp_sum_v, omp_mul_v)
__global__ __launch_bounds__(256) void k_gs_update(int n, double* __restrict__ u, size_t ld, int i,
                                                   const double* __restrict__ coef, const GmState* __restrict__ st,
                                                   float* __restrict__ u32, double* __restrict__ z)
{
#pragma clang fp contract(off)
  if (st->done) return;
  const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (size_t)n) return;
  double* w = u + (size_t)(i + 1) * ld;
  double v = w[k];
  // eight vectors' loads issued before their (sequential) updates
  for (int j0 = 0; j0 <= i; j0 += 8) {
    double b[8];
#pragma unroll
    for (int q = 0; q < 8; q++) b[q] = j0 + q <= i ? u[(size_t)(j0 + q) * ld + k] : 0.0;
#pragma unroll
    for (int q = 0; q < 8; q++)
      if (j0 + q <= i) v = v + (-coef[j0 + q]) * b[q];
  }
  v = st->scale * v;
  w[k] = v;
  if (u32) u32[k] = (float)v;
  if (z) z[k] = v;  // flexible solve: the next step is one the network does not precondition
}

// Compressed basis (mixed precision): u_{i+1} = (w - sum_j coef_j v_j) * scale
// in FP64 with the FP32 basis vectors v_j, two elements per lane (n is even);
// the result is stored as the FP32 basis vector v_{i+1} (the next SpMV input)
// and, with keep64, in FP64 over w for a step that reads it (the network's
// input, a coupled face's rank-one term).
template <typename TWv>
__global__ __launch_bounds__(256) void k_gs_update_cb(int n, TWv* __restrict__ w, float* __restrict__ v32,
                                                      size_t ld, int i, const double* __restrict__ coef,
                                                      const GmState* __restrict__ st, int keep64)
{
#pragma clang fp contract(off)
  if (st->done) return;
  const size_t k = 2 * ((size_t)blockIdx.x * blockDim.x + threadIdx.x);
  if (k >= (size_t)n) return;
  double2 v = make_double2((double)w[k], (double)w[k + 1]);
  // the FP32 basis streamed non-temporally (see basis_load), eight vectors'
  // loads issued before their (sequential) updates
  for (int j0 = 0; j0 <= i; j0 += 8) {
    unsigned long long bb[8];
#pragma unroll
    for (int q = 0; q < 8; q++)
      bb[q] = j0 + q <= i ? __builtin_nontemporal_load(
                                reinterpret_cast<const unsigned long long*>(v32 + (size_t)(j0 + q) * ld + k))
                          : 0ull;
#pragma unroll
    for (int q = 0; q < 8; q++)
      if (j0 + q <= i) {
        const double c = -coef[j0 + q];
        v.x = v.x + c * (double)__uint_as_float((unsigned)bb[q]);
        v.y = v.y + c * (double)__uint_as_float((unsigned)(bb[q] >> 32));
      }
  }
  v.x = st->scale * v.x;
  v.y = st->scale * v.y;
  const unsigned long long o = (unsigned long long)__float_as_uint((float)v.x) |
                               ((unsigned long long)__float_as_uint((float)v.y) << 32);
  __builtin_nontemporal_store(o, reinterpret_cast<unsigned long long*>(v32 + (size_t)(i + 1) * ld + k));
  if constexpr (sizeof(TWv) == sizeof(double))
    if (keep64) *reinterpret_cast<double2*>(w + k) = v;
}

// z_i = u_i (and its FP32 copy): an Arnoldi step of the flexible solve that
// the network does not precondition.
__global__ __launch_bounds__(256) void k_zcopy(int n, const double* __restrict__ u, double* __restrict__ z,
                                               float* __restrict__ z32, const GmState* __restrict__ st)
{
  if (st->done) return;
  const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (size_t)n) return;
  const double v = u[k];
  z[k] = v;
  if (z32) z32[k] = (float)v;
}

// u0 = u0 / err0 (Array division, Array.h:793-805)
__global__ __launch_bounds__(256) void k_div(int n, double* __restrict__ u, const GmState* __restrict__ st,
                                             float* __restrict__ u32)
{
  if (st->done) return;
  const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (size_t)n) return;
  const double v = u[k] / st->scale;
  u[k] = v;
  if (u32) u32[k] = (float)v;
}

__global__ void k_reset_state(GmState* st)
{
  if (threadIdx.x == 0) {
    st->done = st->last_i = st->itr = st->suc = st->divzero = st->early = st->tconv = 0;
  }
}

// Back substitution, gmres.cpp:595-608; the triangle is staged in LDS when it fits.
constexpr int BS_LDS = 6144;  // doubles of the staged Hessenberg (48 KB)
__device__ __forceinline__ void backsolve_body(int sD, const double* __restrict__ h, const double* __restrict__ err,
                                               double* __restrict__ y, GmState* st, double* lh, double* ly)
{
#pragma clang fp contract(off)
  if (st->early || st->divzero || st->tconv) return;
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
      int k = j + 1;
      if (staged) {
        for (; k + 4 <= last_i + 1; k += 4) {
          const double h0 = lh[j + nc * k], h1 = lh[j + nc * (k + 1)], h2 = lh[j + nc * (k + 2)],
                       h3 = lh[j + nc * (k + 3)];
          const double y0 = ly[k], y1 = ly[k + 1], y2 = ly[k + 2], y3 = ly[k + 3];
          yj = yj - h0 * y0;
          yj = yj - h1 * y1;
          yj = yj - h2 * y2;
          yj = yj - h3 * y3;
        }
      }
      for (; k <= last_i; k++) yj = yj - (staged ? lh[j + nc * k] : h[j + (size_t)ldh * k]) * ly[k];
      ly[j] = yj / (staged ? lh[j + nc * j] : h[j + (size_t)ldh * j]);
    }
    st->fNorm = fabs(err[last_i + 1]);
  }
  __syncthreads();
  for (int q = threadIdx.x; q < nc; q += blockDim.x) y[q] = ly[q];
}

__global__ __launch_bounds__(256) void k_backsolve(int sD, const double* __restrict__ h, const double* __restrict__ err,
                                                   double* __restrict__ y, GmState* st)
{
  __shared__ double lh[BS_LDS];
  __shared__ double ly[MAXM];
  backsolve_body(sD, h, err, y, st, lh, ly);
}

// X += sum_j y_j u_j, sequentially in j (gmres.cpp:610-612)
__global__ __launch_bounds__(256) void k_xupdate(int n, double* __restrict__ X, const double* __restrict__ u, size_t ld,
                                                 const double* __restrict__ y, const GmState* __restrict__ st)
{
#pragma clang fp contract(off)
  if (st->early || st->divzero || st->tconv) return;
  const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (size_t)n) return;
  int last_i = st->last_i;
  double x = X[k];
  for (int j = 0; j <= last_i; j++) x = x + y[j] * u[(size_t)j * ld + k];
  X[k] = x;
}

// Compressed basis: X += sum_j y_j b_j, sequentially in j, with b_j = z32_j
// (FP32) when the whole flexible basis is FP32 (z32 != null), else z_j (FP64)
// on the steps the network preconditioned (z != null, j % period == 0) and the
// FP32 basis vector v_j otherwise; two elements per lane.
__global__ __launch_bounds__(256) void k_xupdate_cb(int n, double* __restrict__ X, const float* __restrict__ v32,
                                                    const double* __restrict__ z, int period, size_t ld,
                                                    const double* __restrict__ y, const GmState* __restrict__ st,
                                                    const float* __restrict__ z32)
{
#pragma clang fp contract(off)
  if (st->early || st->divzero || st->tconv) return;
  const size_t k = 2 * ((size_t)blockIdx.x * blockDim.x + threadIdx.x);
  if (k >= (size_t)n) return;
  const int last_i = st->last_i;
  double2 x = *reinterpret_cast<const double2*>(X + k);
  for (int j = 0; j <= last_i; j++) {
    double2 b;
    if (z32) {
      const float2 f = *reinterpret_cast<const float2*>(z32 + (size_t)j * ld + k);
      b = make_double2((double)f.x, (double)f.y);
    } else if (z && (period <= 1 || j % period == 0)) {
      b = *reinterpret_cast<const double2*>(z + (size_t)j * ld + k);
    } else {
      const float2 f = *reinterpret_cast<const float2*>(v32 + (size_t)j * ld + k);
      b = make_double2((double)f.x, (double)f.y);
    }
    x.x = x.x + y[j] * b.x;
    x.y = x.y + y[j] * b.y;
  }
  *reinterpret_cast<double2*>(X + k) = x;
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

// ------------------------------------------------- NS / Schur kernels ---
// depart (ns_solver.cpp:91-159): split each 4x4 block into mK(9), mG(3), mD(3),
// mL(1); Gt(:,l) = -mG(:,j) where block l is the transpose of block j.
__global__ __launch_bounds__(256) void k_depart(int nnz, const double* __restrict__ Val, const int* __restrict__ trans,
                                                double* __restrict__ mK, double* __restrict__ mG,
                                                double* __restrict__ mD, double* __restrict__ mL,
                                                double* __restrict__ Gt)
{
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nnz) return;
  const double* t = Val + 16 * (size_t)j;
  double* k = mK + 9 * (size_t)j;
  k[0] = t[0]; k[1] = t[1]; k[2] = t[2];
  k[3] = t[4]; k[4] = t[5]; k[5] = t[6];
  k[6] = t[8]; k[7] = t[9]; k[8] = t[10];
  mG[3 * (size_t)j + 0] = t[3];
  mG[3 * (size_t)j + 1] = t[7];
  mG[3 * (size_t)j + 2] = t[11];
  mD[3 * (size_t)j + 0] = t[12];
  mD[3 * (size_t)j + 1] = t[13];
  mD[3 * (size_t)j + 2] = t[14];
  mL[j] = t[15];
  const double* tt = Val + 16 * (size_t)trans[j];
  Gt[3 * (size_t)j + 0] = -tt[3];
  Gt[3 * (size_t)j + 1] = -tt[7];
  Gt[3 * (size_t)j + 2] = -tt[11];
}

// Fixed-order sum of a block-partial array (per lane at stride 64, then a
// butterfly) computed by the first wavefront of every workgroup that needs it:
// each workgroup obtains the identical value without atomics or a launch.
__device__ __forceinline__ double sum_partials(const double* __restrict__ partial, int nblk, double* shared_out)
{
  const int t = threadIdx.x, lane = t & 63;
  if (t < 64) {
    double acc = 0.0;
    for (int b = lane; b < nblk; b += 64) acc = acc + partial[b];
    acc = wave_sum(acc);
    if (lane == 0) *shared_out = acc;
  }
  __syncthreads();
  return *shared_out;
}

__device__ __forceinline__ void block_partial(double v, double* __restrict__ partial)
{
  __shared__ double wsum[RED_NW];
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  v = wave_sum(v);
  if (lane == 0) wsum[wv] = v;
  __syncthreads();
  if (t == 0) {
    double v2 = wsum[0];
    for (int q = 1; q < RED_NW; q++) v2 = v2 + wsum[q];
    partial[blockIdx.x] = v2;
  }
}

// Schur CG start (cgrad.cpp:66-79): X = 0, P = R, partial sums of R^2.
__global__ __launch_bounds__(RED_BS) void k_cg_init(int nNo, const double* __restrict__ R, double* __restrict__ X,
                                                    double* __restrict__ P, double* __restrict__ partial)
{
#pragma clang fp contract(off)
  double sq = 0.0;
  for (int r = 0; r < RED_EPT; r++) {
    const int row = blockIdx.x * RED_SPAN + threadIdx.x + r * RED_BS;
    if (row >= nNo) continue;
    const double v = R[row];
    X[row] = 0.0;
    P[row] = v;
    sq = sq + v * v;
  }
  block_partial(sq, partial);
}

// iNorm = |R|, eps = max(absTol, relTol iNorm)^2, errO = err = iNorm^2, stop test.
__global__ __launch_bounds__(RED_BS) void k_cg_start(const double* __restrict__ partial, int nblk, CgState* cg,
                                                     double absTol, double relTol)
{
#pragma clang fp contract(off)
  __shared__ double tot;
  const double acc = sum_partials(partial, nblk, &tot);
  if (threadIdx.x == 0) {
    const double iN = sqrt(acc);
    const double e = fmax(absTol, relTol * iN);
    cg->iNorm = iN;
    cg->eps = e * e;
    cg->errO[0] = iN * iN;
    cg->errO_last = cg->errO[0];
    cg->err = cg->errO[0];
    cg->suc = 0;
    cg->stop = 0;
    cg->last_i = 0;
    if (cg->err < cg->eps) {
      cg->suc = 1;
      cg->stop = 1;
    }
  }
}

// Part 1 (cgrad.cpp:101-118): DGP = Gt . GP (vs), SP = mL P (ss), SP = SP - DGP;
// partial sums of P.SP.
__global__ __launch_bounds__(RED_BS) void k_cg_sp(int nNo, const int* __restrict__ rowPtr,
                                                  const int* __restrict__ colPtr, const double* __restrict__ Gt,
                                                  const double* __restrict__ mL, const double* __restrict__ GP,
                                                  const double* __restrict__ P, double* __restrict__ SP,
                                                  double* __restrict__ partial, const CgState* __restrict__ cg)
{
#pragma clang fp contract(off)
  if (cg->stop) return;
  double prod = 0.0;
  for (int r = 0; r < RED_EPT; r++) {
    const int row = blockIdx.x * RED_SPAN + threadIdx.x + r * RED_BS;
    if (row >= nNo) continue;
    double dgp = 0.0, sp = 0.0;
    for (int j = rowPtr[row]; j < rowPtr[row + 1]; j++) {
      const int col = colPtr[j];
      dgp = dgp + Gt[3 * (size_t)j] * GP[3 * (size_t)col] + Gt[3 * (size_t)j + 1] * GP[3 * (size_t)col + 1] +
            Gt[3 * (size_t)j + 2] * GP[3 * (size_t)col + 2];
      sp = sp + mL[j] * P[col];
    }
    sp = sp + (-1.0) * dgp;
    SP[row] = sp;
    prod = prod + P[row] * sp;
  }
  block_partial(prod, partial);
}

// Part 2 (cgrad.cpp:120-135): alpha = errO / <P,SP>, X += alpha P, R -= alpha SP;
// partial sums of R^2 into partial2.
__global__ __launch_bounds__(RED_BS) void k_cg_upd(int nNo, double* __restrict__ X, const double* __restrict__ P,
                                                   double* __restrict__ R, const double* __restrict__ SP,
                                                   const double* __restrict__ partial, double* __restrict__ partial2,
                                                   int nblk, const CgState* __restrict__ cg, int i)
{
#pragma clang fp contract(off)
  __shared__ double tot;
  if (cg->stop) return;
  const double alpha = cg->errO[i & 1] / sum_partials(partial, nblk, &tot);
  double sq = 0.0;
  for (int r = 0; r < RED_EPT; r++) {
    const int row = blockIdx.x * RED_SPAN + threadIdx.x + r * RED_BS;
    if (row >= nNo) continue;
    X[row] = X[row] + alpha * P[row];
    const double rr = R[row] + (-alpha) * SP[row];
    R[row] = rr;
    sq = sq + rr * rr;
  }
  block_partial(sq, partial2);
}

// Part 3 (cgrad.cpp:137-152): err = |R|^2 (sqrt then square), P = c2 (P + c1 R)
// with c1 = errO/err, c2 = err/errO; workgroup 0 records err and the stop test
// of the next trip (cgrad.cpp:84-88).
__global__ __launch_bounds__(RED_BS) void k_cg_p(int nNo, double* __restrict__ P, const double* __restrict__ R,
                                                 const double* __restrict__ partial2, int nblk, CgState* cg, int i,
                                                 int mItr, volatile int* host_done)
{
#pragma clang fp contract(off)
  __shared__ double tot;
  if (cg->stop) return;
  double e = sqrt(sum_partials(partial2, nblk, &tot));
  e = e * e;
  const double errO = cg->errO[i & 1];
  const double c1 = errO / e, c2 = e / errO;
  for (int r = 0; r < RED_EPT; r++) {
    const int row = blockIdx.x * RED_SPAN + threadIdx.x + r * RED_BS;
    if (row >= nNo) continue;
    P[row] = c2 * (P[row] + c1 * R[row]);
  }
  // Workgroup 0 records the trip. It writes only err, last_i, errO[(i+1)&1] and
  // the stop flag; stop guards the remaining P updates of this trip, which are
  // dead once the iteration has converged.
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    cg->err = e;
    cg->errO_last = errO;
    cg->last_i = i;
    if (i + 1 < mItr) {
      cg->last_i = i + 1;
      cg->errO[(i + 1) & 1] = e;
      if (e < cg->eps) {
        cg->suc = 1;
        cg->stop = 1;
        *host_done = 1;
        __threadfence_system();
      }
    } else {
      cg->stop = 1;
    }
  }
}

// Dot products of a list of vector pairs: out[p] = sum_k a_p[k] b_p[k].
struct DotPair {
  const double* a;
  const double* b;
  int n;
};
constexpr int MAXPAIRS = 64;
struct DotList {
  DotPair p[MAXPAIRS];
  int np;
};

__global__ __launch_bounds__(RED_BS) void k_dot_list(DotList L, double* __restrict__ partial)
{
#pragma clang fp contract(off)
  __shared__ double wsum[RED_NW];
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const int pr = blockIdx.y;
  const DotPair P = L.p[pr];
  double acc = 0.0;
  for (int r = 0; r < RED_EPT; r++) {
    size_t k = (size_t)blockIdx.x * RED_SPAN + t + (size_t)r * RED_BS;
    if (k < (size_t)P.n) acc = acc + P.a[k] * P.b[k];
  }
  acc = wave_sum(acc);
  if (lane == 0) wsum[wv] = acc;
  __syncthreads();
  if (t == 0) {
    double v2 = wsum[0];
    for (int q = 1; q < RED_NW; q++) v2 = v2 + wsum[q];
    partial[(size_t)pr * gridDim.x + blockIdx.x] = v2;
  }
}

__global__ __launch_bounds__(FIN_BS) void k_dot_list_fin(int np, int nblk, const double* __restrict__ partial,
                                                         double* __restrict__ out)
{
#pragma clang fp contract(off)
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  for (int q = wv; q < np; q += FIN_BS / 64) {
    double s = 0.0;
    for (int b = lane; b < nblk; b += 64) s = s + partial[(size_t)q * nblk + b];
    s = wave_sum(s);
    if (lane == 0) out[q] = s;
  }
}

// y = x0 - sum_{j<m} c_j V_j, sequentially in j (ns_solver.cpp:420-425).
__global__ __launch_bounds__(256) void k_axpy_sub(int n, const double* __restrict__ x0, const double* __restrict__ V,
                                                  size_t ld, int m, const double* __restrict__ c,
                                                  double* __restrict__ y)
{
#pragma clang fp contract(off)
  const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (size_t)n) return;
  double v = x0[k] - c[0] * V[k];
  for (int j = 1; j < m; j++) v = v - c[j] * V[(size_t)j * ld + k];
  y[k] = v;
}

// y = c[i0] V_0 + sum_{i=1..itr} c[2i+off] V_i (ns_solver.cpp:480-488).
__global__ __launch_bounds__(256) void k_combine(int n, const double* __restrict__ V, size_t ld, int itr,
                                                 const double* __restrict__ c, int off, double* __restrict__ y)
{
#pragma clang fp contract(off)
  const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (size_t)n) return;
  double v = c[off] * V[k];
  for (int i = 1; i <= itr; i++) v = v + c[2 * i + off] * V[(size_t)i * ld + k];
  y[k] = v;
}

// Ri(4,nNo) = [Rm(3,nNo); Rc(nNo)] and the split of Ri into Rm/Rc.
__global__ __launch_bounds__(256) void k_interleave(int nNo, const double* __restrict__ m3, const double* __restrict__ c1,
                                                    double* __restrict__ R4)
{
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= nNo) return;
  R4[4 * (size_t)a + 0] = m3[3 * (size_t)a + 0];
  R4[4 * (size_t)a + 1] = m3[3 * (size_t)a + 1];
  R4[4 * (size_t)a + 2] = m3[3 * (size_t)a + 2];
  R4[4 * (size_t)a + 3] = c1[a];
}

__global__ __launch_bounds__(256) void k_split(int nNo, const double* __restrict__ R4, double* __restrict__ m3,
                                               double* __restrict__ c1)
{
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= nNo) return;
  m3[3 * (size_t)a + 0] = R4[4 * (size_t)a + 0];
  m3[3 * (size_t)a + 1] = R4[4 * (size_t)a + 1];
  m3[3 * (size_t)a + 2] = R4[4 * (size_t)a + 2];
  c1[a] = R4[4 * (size_t)a + 3];
}

__global__ __launch_bounds__(256) void k_sub(int n, const double* __restrict__ a, const double* __restrict__ b,
                                             double* __restrict__ out)
{
#pragma clang fp contract(off)
  const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (size_t)n) return;
  out[k] = a[k] - b[k];
}

// --------------------------------------------------------- workspace ----
void krylov_free(KrylovWork& k)
{
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
  dev_free(k.u32);
  dev_free(k.v32);
  dev_free(k.z);
  dev_free(k.z32);
  dev_free(k.w32);
  dev_free(k.r2);
  if (k.h_st) (void)hipHostFree(k.h_st);
  k.h_st = nullptr;
  if (k.h_done) {
    (void)hipHostFree(k.h_done);
    for (auto& e : k.ev) (void)hipEventDestroy(e);
  }
  k.h_done = nullptr;
  k.d_done = nullptr;
  k.n = 0;
  k.sD = 0;
}

static int ensure_work(KrylovWork& k, int n, int sD)
{
  if (k.u && k.n == n && k.sD >= sD) return SVG_OK;
  krylov_free(k);
  k.n = n;
  k.sD = sD;
  k.ld = ((size_t)n + 63) & ~(size_t)63;
  k.nblk = (int)(((n + RED_SPAN - 1) / RED_SPAN + 7) & ~7);
  int rc;
  if ((rc = dev_alloc(&k.u, k.ld * (sD + 1))) || (rc = dev_alloc(&k.X, (size_t)n)) || (rc = dev_alloc(&k.W, (size_t)n)) ||
      (rc = dev_alloc(&k.h, (size_t)(sD + 1) * sD)) || (rc = dev_alloc(&k.cs, sD)) || (rc = dev_alloc(&k.sn, sD)) ||
      (rc = dev_alloc(&k.err, sD + 1)) || (rc = dev_alloc(&k.coef, sD + 2)) || (rc = dev_alloc(&k.y, sD)) ||
      (rc = dev_alloc(&k.partial, (size_t)k.nblk * (sD + 2))) || (rc = dev_alloc(&k.st, 1)) ||
      (rc = dev_alloc(&k.scal, 8)))
    return rc;
  SVG_HIP(hipHostMalloc((void**)&k.h_st, sizeof(GmState), hipHostMallocDefault));
  if (!k.h_done) {
    SVG_HIP(hipHostMalloc((void**)&k.h_done, sizeof(int), hipHostMallocMapped | hipHostMallocCoherent));
    SVG_HIP(hipHostGetDevicePointer((void**)&k.d_done, k.h_done, 0));
    for (auto& e : k.ev) SVG_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  }
  SVG_HIP(hipMemset(k.h, 0, sizeof(double) * (sD + 1) * sD));
  SVG_HIP(hipMemset(k.st, 0, sizeof(GmState)));
  return SVG_OK;
}

// Scratch for the standalone dot/norm entry points (not tied to an lhs).
struct DotScratch {
  double* partial = nullptr;
  int cap = 0;
  double* out = nullptr;
};
static DotScratch g_dot;

// Scratch of svg_dot_v_noalloc: a file-scope device array (each GPU holds its
// own copy for the life of the process), so that entry point never allocates.
// It holds one partial sum per reduction workgroup (RED_SPAN entries each) of
// a vector of at most kNoAllocMaxN entries, then the result slot. Every slot
// a call reads was written earlier in the same call, so nothing is cleared.
constexpr long kNoAllocMaxN = 1L << 28;
constexpr int kNoAllocPartials = (int)((kNoAllocMaxN + RED_SPAN - 1) / RED_SPAN);
constexpr int kNoAllocSlots = kNoAllocPartials + 32;  // result at kNoAllocPartials (256-byte aligned)
static __device__ __align__(256) double scratch[kNoAllocSlots];

// Host code never uses `scratch` itself (its host-side name is no device
// address): each device's address is looked up once, with that device
// current, and kept here.
constexpr int kMaxDevices = 64;
static std::mutex g_scratchMu;
static double* g_scratchPtr[kMaxDevices] = {};

static int noalloc_scratch(double** out)
{
  int dev = -1;
  SVG_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return set_error(SVG_ERR_ARG, "svg_dot_v_noalloc: device index out of range");
  std::lock_guard<std::mutex> g(g_scratchMu);
  if (!g_scratchPtr[dev]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_check(e, "hipGetSymbolAddress(scratch)");
    if (!p) return set_error(SVG_ERR_HIP, "hipGetSymbolAddress(scratch) returned a null address");
    g_scratchPtr[dev] = (double*)p;
  }
  *out = g_scratchPtr[dev];
  return SVG_OK;
}

static int dot_device(int n, const double* w, const double* base, size_t ld, int m, double* partial, double* out,
                      GmState* st, FinArgs fa, double* h, double* cs, double* sn, double* err, double* coef,
                      hipStream_t s, int* host_done = nullptr)
{
  int nblk = (n + RED_SPAN - 1) / RED_SPAN;
  if (nblk < 1) nblk = 1;
  if (m > MAXM) return set_error(SVG_ERR_ARG, "reduction: too many sums");
  const int* skip = (fa.mode == FIN_ARNOLDI) ? &st->done : nullptr;
  // algorithmic bytes: w and the m - 1 basis vectors read, the partials written; the finaliser reads them
  launch_b("mdot", 8.0 * n * m + 8.0 * m * nblk, k_mdot, dim3(nblk), dim3(RED_BS), (size_t)RED_NW * m * sizeof(double),
           s, n, w, base, ld, m, partial, skip);
  launch_b("mdot_fin", 8.0 * m * nblk, k_fin, dim3(1), dim3(FIN_BS), 0, s, fa, (const double*)partial, nblk, m, st,
           h, cs, sn, err, coef, out, (volatile int*)host_done);
  return SVG_OK;
}

// Arnoldi dots of the compressed-basis solve: w against v_0..v_{m-1} (FP32)
// and w.w, then the finaliser over the m + 1 sums.
template <typename TWv>
static int dot_device_cb(int n, const TWv* w, const float* base, size_t ld, int m, KrylovWork& k, FinArgs fa,
                         hipStream_t s)
{
  int nblk = (n + RED_SPAN - 1) / RED_SPAN;
  if (nblk < 1) nblk = 1;
  if (m + 1 > MAXM) return set_error(SVG_ERR_ARG, "reduction: too many sums");
  // algorithmic bytes: w (FP64 or FP32) and the m FP32 basis vectors read, the partials written
  launch_b("mdot_cb", (double)sizeof(TWv) * n + 4.0 * n * m + 8.0 * (m + 1) * nblk, k_mdot_cb<TWv>, dim3(nblk),
           dim3(RED_BS), (size_t)RED_NW * (m + 1) * sizeof(double), s, n, w, base, ld, m, k.partial,
           (const int*)&k.st->done);
  launch_b("mdot_fin", 8.0 * (m + 1) * nblk, k_fin, dim3(1), dim3(FIN_BS), 0, s, fa, (const double*)k.partial, nblk,
           m + 1, k.st, k.h, k.cs, k.sn, k.err, k.coef, k.scal, (volatile int*)k.d_done);
  return SVG_OK;
}

// add_bc_mul (add_bc_mul.cpp:53-121), op ADD, for one coupled face:
// S = res valM^T X in one workgroup (lane t sums the entries q = t, t+256, ...
// of q = a nsd + i in order, a butterfly per wavefront, the 4 wavefront sums
// in order), then Y(i, glob(a)) += valM(i, a) S.
// X may be Y (op PRE in place): every read of X precedes the barrier.
__global__ __launch_bounds__(256) void k_add_bc_mul(int nNo, int fdof, int nsd, int dof, const int* __restrict__ glob,
                                                    const double* __restrict__ valM, double res, const double* X,
                                                    double* Y, const int* __restrict__ skip)
{
#pragma clang fp contract(off)
  __shared__ double ws[4];
  __shared__ double Ssh;
  if (skip && *skip) return;
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const long nq = (long)nNo * nsd;
  double acc = 0.0;
  for (long q = t; q < nq; q += 256) {
    const int a = (int)(q / nsd), i = (int)(q - (long)a * nsd);
    acc = acc + valM[i + (size_t)fdof * a] * X[i + (size_t)dof * glob[a]];
  }
  acc = wave_sum(acc);
  if (lane == 0) ws[wv] = acc;
  __syncthreads();
  if (t == 0) {
    double S = ws[0];
    for (int w = 1; w < 4; w++) S = S + ws[w];
    Ssh = res * S;
  }
  __syncthreads();
  const double S = Ssh;
  for (long q = t; q < nq; q += 256) {
    const int a = (int)(q / nsd), i = (int)(q - (long)a * nsd);
    double* y = Y + i + (size_t)dof * glob[a];
    *y = *y + valM[i + (size_t)fdof * a] * S;
  }
}

// valM(i, a) = val(i, a) W(i, glob(a)) for i < nsd (precond.cpp:246-256)
__global__ __launch_bounds__(256) void k_face_valM(int nNo, int fdof, int nsd, int dof, const int* __restrict__ glob,
                                                   const double* __restrict__ val, const double* __restrict__ W,
                                                   double* __restrict__ valM)
{
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (long)nNo * nsd) return;
  const int a = (int)(q / nsd), i = (int)(q - (long)a * nsd);
  valM[i + (size_t)fdof * a] = val[i + (size_t)fdof * a] * W[i + (size_t)dof * glob[a]];
}

// The learned preconditioner's stagnation check armed again: no fallback yet,
// no non-decreasing iterations, no best residual estimate.
__global__ __launch_bounds__(64) void k_fb_rearm(GmState* st)
{
  if (threadIdx.x == 0) {
    st->fallback = 0;
    st->nondec = 0;
    st->best = DBL_MAX;
  }
}

// y = b - y
__global__ __launch_bounds__(256) void k_rsub(int n, const double* __restrict__ b, double* __restrict__ y)
{
  const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < (size_t)n) y[k] = b[k] - y[k];
}

static bool has_coupled(const svg_lhs* L)
{
  for (const auto& F : L->face)
    if (F.coupledFlag) return true;
  return false;
}

// op PRE (add_bc_mul.cpp:60-64): Y += coef valM (valM^T X), coef = -res / (1 + res nS),
// face after face (X may be Y: each face then reads the previous faces' update).
static void add_bc_pre_dev(const svg_lhs* L, int dof, const double* X, double* Y, const int* skip, hipStream_t s)
{
  for (const auto& F : L->face) {
    if (!F.coupledFlag) continue;
    const int nsd = F.dof < dof ? F.dof : dof;
    const double coef = -F.res / (1.0 + (F.res * F.nS));
    launch("add_bc_pre", k_add_bc_mul, dim3(1), dim3(256), 0, s, F.nNo, F.dof, nsd, dof, (const int*)F.d_glob,
           (const double*)F.d_valM, coef, X, Y, skip);
  }
}

// Y += res valM (valM^T X) for every coupled face (a no-op without one).
static void add_bc_mul_dev(const svg_lhs* L, int dof, const double* X, double* Y, const int* skip, hipStream_t s)
{
  for (const auto& F : L->face) {
    if (!F.coupledFlag) continue;
    const int nsd = F.dof < dof ? F.dof : dof;
    launch("add_bc_mul", k_add_bc_mul, dim3(1), dim3(256), 0, s, F.nNo, F.dof, nsd, dof, (const int*)F.d_glob,
           (const double*)F.d_valM, F.res, X, Y, skip);
  }
}

// The operator of a Krylov solve: the dof x dof block matrix K over lhs's pattern.
struct Op {
  const svg_lhs* L;
  const double* K;
  int dof;
  bool pre = false;  // coupled faces: op PRE after the operator (the NS solver's inner gmres, gmres.cpp:144-150)
};

static void op_apply(const Op& A, const double* x, double* y, const double* b, const int* skip, hipStream_t s)
{
  if (has_coupled(A.L)) {
    // K x + add_bc_mul (gmres.cpp:508-512, 537-539; 134-139, 186-189); the residual form as
    // b - (K x + res v v^T x); then, for the inner gmres, op PRE in place (gmres.cpp:142-150, 191-199)
    if (A.dof == 4)
      spmv<4, 4, 0>(A.L, A.K, x, y, nullptr, skip, s);
    else
      spmv<3, 3, 0>(A.L, A.K, x, y, nullptr, skip, s, "spmv_k3");
    add_bc_mul_dev(A.L, A.dof, x, y, skip, s);
    const int n = A.dof * A.L->nNo;
    if (b) launch("rsub", k_rsub, dim3(blocks_for((size_t)n, 256)), dim3(256), 0, s, n, b, y);
    if (A.pre) add_bc_pre_dev(A.L, A.dof, y, y, skip, s);
    return;
  }
  if (A.dof == 4) {
    if (b)
      spmv<4, 4, 1>(A.L, A.K, x, y, b, skip, s);
    else
      spmv<4, 4, 0>(A.L, A.K, x, y, nullptr, skip, s);
  } else {
    if (b)
      spmv<3, 3, 1>(A.L, A.K, x, y, b, skip, s, "spmv_k3");
    else
      spmv<3, 3, 0>(A.L, A.K, x, y, nullptr, skip, s, "spmv_k3");
  }
}

int spmv_residual(const svg_lhs* L, const double* K, const double* U, const double* B, double* out, hipStream_t s)
{
  spmv<4, 4, 1>(L, K, U, out, B, nullptr, s, "spmv_res");
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? SVG_OK : hip_check(e, "spmv_residual");
}

// One restart cycle (gmres.cpp:529-612 / :533-592): up to sD Arnoldi steps with
// at most LOOKAHEAD steps enqueued beyond the host's last convergence check,
// then the back substitution and X += sum y_j u_j. Returns after the cycle's
// state has been copied to the host. twolevel: every step is preconditioned by
// the lhs's two-level preconditioner, z_i = P zc + B(u_i - A P zc) with B the
// block Jacobi smoother, and on the network's steps z_i = P zc + N(B(u_i - A P zc)).
static int run_cycle(const Op& A, KrylovWork& k, int sD, FinArgs fa, double* X, hipStream_t s,
                     const float* K32 = nullptr, svg_lpc* net = nullptr, int net_period = 1, bool twolevel = false)
{
  const bool mixed = K32 != nullptr;
  const bool tl = twolevel;
  const bool flex = net != nullptr || tl;
  const int n = k.n;
  const unsigned nbv = blocks_for(n, 256), nbv2 = blocks_for(n / 2, 256);
  // mixed precision keeps the basis in FP32 (k.v32); the FP64 u_i is kept
  // only where a step reads it: the network's input, a coupled face's term
  const bool coupled = mixed && has_coupled(A.L);
  auto netStep = [&](int j) { return net && (net_period <= 1 || j % net_period == 0); };
  // k.h_done was cleared by the caller before its restart check, which raises
  // it when the cycle starts done (converged true residual, zero residual)
  const size_t mark0 = prof_mark();
  size_t postFin[MAXM];  // timed-launch positions before each step's Gram-Schmidt update
  for (int i = 0; i < sD; i++) {
    if (i >= LOOKAHEAD) {
      SVG_HIP(hipEventSynchronize(k.ev[(i - LOOKAHEAD) % 8]));
      if (*(volatile int*)k.h_done) break;
    }
    fa.mode = FIN_ARNOLDI;
    fa.i = i;
    // flexible GMRES: the SpMV acts on z_i = M(u_i) (and its FP32 copy in u32)
    const double* sin = k.u + k.ld * i;
    double* w = k.u + k.ld * (i + 1);
    if (tl && netStep(i)) {
      // the network corrects the block-Jacobi smoothing of the coarse-corrected
      // residual, z = N(B r2); then z += P zc
      const int* skip = (const int*)&k.st->done;
      if (mixed)
        coarse_smooth_dev(A.L, (const float*)(k.v32 + k.ld * i), k.r2, skip, s);
      else
        coarse_smooth_dev(A.L, (const double*)(k.u + k.ld * i), k.r2, skip, s);
      int rc = lpc_apply_dev(net, A.L, k.r2, k.z + k.ld * i, mixed ? k.z32 + k.ld * i : nullptr, skip, s);
      if (rc) return rc;
      coarse_add_dev(A.L, k.z + k.ld * i, mixed ? k.z32 + k.ld * i : nullptr, skip, s);
      sin = k.z + k.ld * i;
    } else if (tl) {
      // z_i = M(u_i): from the FP32 basis vector into the FP32 flexible basis
      // (and FP64 where a coupled face's term reads it), or FP64 throughout
      if (mixed)
        coarse_apply_dev(A.L, (const float*)(k.v32 + k.ld * i), coupled ? k.z + k.ld * i : nullptr,
                         k.z32 + k.ld * i, (const int*)&k.st->done, s);
      else
        coarse_apply_dev(A.L, (const double*)(k.u + k.ld * i), k.z + k.ld * i, nullptr, (const int*)&k.st->done, s);
      sin = k.z + k.ld * i;
    } else if (netStep(i)) {
      int rc = lpc_apply_dev(net, A.L, k.u + k.ld * i, k.z + k.ld * i, mixed ? k.u32 : nullptr, &k.st->done, s);
      if (rc) return rc;
      sin = k.z + k.ld * i;
    } else if (flex && !mixed) {
      if (i == 0)
        launch("zcopy", k_zcopy, dim3(nbv), dim3(256), 0, s, n, (const double*)(k.u + k.ld * i), k.z + k.ld * i,
               (float*)nullptr, (const GmState*)k.st);
      // else: the previous step's Gram-Schmidt update wrote z_i = u_i
      sin = k.z + k.ld * i;
    }
    if (mixed && tl && !coupled && k.w32) {
      // the SpMV's output in FP32 (no FP64 term is added to it here), read by
      // the dots and the update as FP64 values: the same arithmetic, half the bytes
      spmv32_f(A.L, K32, (const float*)(k.z32 + k.ld * i), k.w32, (const int*)&k.st->done, s);
      int rc = dot_device_cb(n, (const float*)k.w32, k.v32, k.ld, i + 1, k, fa, s);
      if (rc) return rc;
      postFin[i] = prof_mark();
      // algorithmic bytes: w (FP32) and v_0..v_i read, v_{i+1} (FP32) written
      launch_b("gs_update_cb", 4.0 * n + 4.0 * n * (i + 2), k_gs_update_cb<float>, dim3(nbv2), dim3(256), 0, s, n,
               k.w32, k.v32, k.ld, i, (const double*)k.coef, (const GmState*)k.st, 0);
    } else if (mixed) {
      spmv32(A.L, K32,
             tl ? (const float*)(k.z32 + k.ld * i)
                : (netStep(i) ? (const float*)k.u32 : (const float*)(k.v32 + k.ld * i)),
             w, (const int*)&k.st->done, s);
      if (coupled) add_bc_mul_dev(A.L, 4, sin, w, (const int*)&k.st->done, s);  // FP64 rank-one term
      int rc = dot_device_cb(n, w, k.v32, k.ld, i + 1, k, fa, s);
      if (rc) return rc;
      postFin[i] = prof_mark();
      const int keep64 = (!tl && (coupled || (i + 1 < sD && netStep(i + 1)))) ? 1 : 0;
      // algorithmic bytes: w and v_0..v_i read, v_{i+1} (FP32) written, w rewritten with keep64
      launch_b("gs_update_cb", 8.0 * n + 4.0 * n * (i + 2) + (keep64 ? 8.0 * n : 0.0), k_gs_update_cb<double>,
               dim3(nbv2), dim3(256), 0, s, n, w, k.v32, k.ld, i, (const double*)k.coef, (const GmState*)k.st, keep64);
    } else {
      op_apply(A, sin, w, nullptr, &k.st->done, s);
      int rc = dot_device(n, w, k.u, k.ld, i + 2, k.partial, k.scal, k.st, fa, k.h, k.cs, k.sn, k.err, k.coef, s,
                          k.d_done);
      if (rc) return rc;
      postFin[i] = prof_mark();
      // flexible: when the next step has no network, this update also writes z_{i+1} = u_{i+1}
      const bool nextPlain = flex && !tl && i + 1 < sD && !netStep(i + 1);
      // algorithmic bytes: w and u_0..u_i read, u_{i+1} written (and z_{i+1})
      launch_b("gs_update", 8.0 * n * (i + 3) + (nextPlain ? 8.0 * n : 0.0), k_gs_update, dim3(nbv), dim3(256), 0, s,
               n, k.u, k.ld, i, (const double*)k.coef, (const GmState*)k.st, (float*)nullptr,
               nextPlain ? k.z + k.ld * (i + 1) : (double*)nullptr);
    }
    SVG_HIP(hipEventRecord(k.ev[i % 8], s));
  }
  const size_t markEnd = prof_mark();
  launch("backsolve", k_backsolve, dim3(1), dim3(256), 0, s, sD, (const double*)k.h, (const double*)k.err, k.y, k.st);
  if (mixed)
    launch("xupdate_cb", k_xupdate_cb, dim3(nbv2), dim3(256), 0, s, n, X, (const float*)k.v32,
           (flex && !tl) ? (const double*)k.z : (const double*)nullptr, net_period, k.ld, (const double*)k.y,
           (const GmState*)k.st, tl ? (const float*)k.z32 : (const float*)nullptr);
  else
    launch("xupdate", k_xupdate, dim3(nbv), dim3(256), 0, s, n, X, (const double*)(flex ? k.z : k.u), k.ld,
           (const double*)k.y, (const GmState*)k.st);
  SVG_HIP(hipGetLastError());
  SVG_HIP(hipMemcpyAsync(k.h_st, k.st, sizeof(GmState), hipMemcpyDeviceToHost, s));
  SVG_HIP(hipStreamSynchronize(s));
  // timed launches that exited at once are not kept: the whole loop when the
  // cycle started done, else those from the converging (or falling back)
  // step's update on
  const GmState& h = *k.h_st;
  if (h.tconv || h.early || h.divzero)
    prof_drop(mark0, markEnd);
  else if (h.done && h.last_i >= 0 && h.last_i < sD)
    prof_drop(postFin[h.last_i], markEnd);
  return SVG_OK;
}

// gmres_v (gmres.cpp:450-629); R holds the (preconditioned) right-hand side on
// entry and the solution on exit.
static int gmres_v_dev(svg_lhs* L, svg_subls* ls, int dof, const double* Val, double* R, hipStream_t s)
{
  if (dof != 4) return set_error(SVG_ERR_UNSUPPORTED, "gmres_v: only dof = 4 is built");
  if (ls->sD < 1 || ls->sD > 250) return set_error(SVG_ERR_ARG, "gmres_v: Krylov dimension must be in [1, 250]");
  auto& k = L->kw;
  int rc = ensure_work(k, dof * L->nNo, ls->sD);
  if (rc) return rc;
  const int n = k.n, sD = ls->sD;
  const Op A{L, Val, 4};
  if (L->spmvVariant < 0 && (rc = svg_spmv_tune(L, Val, nullptr, s))) return rc;
  auto t0 = std::chrono::steady_clock::now();
  FinArgs fa{};
  fa.sD = sD;
  fa.absTol = ls->absTol;
  fa.relTol = ls->relTol;
  fa.mode = FIN_NORM_R;
  if ((rc = dot_device(n, R, R, 0, 1, k.partial, k.scal, k.st, fa, k.h, k.cs, k.sn, k.err, k.coef, s)))
    return rc;
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
    *k.h_done = 0;
    op_apply(A, k.X, k.u, R, nullptr, s);  // u0 = R - A X
    fa.mode = FIN_RESTART;
    fa.inc = 1;
    fa.first = 0;
    if ((rc = dot_device(n, k.u, k.u, 0, 1, k.partial, k.scal, k.st, fa, k.h, k.cs, k.sn, k.err, k.coef, s,
                         k.d_done)))
      return rc;
    launch("div", k_div, dim3(nbv), dim3(256), 0, s, n, k.u, (const GmState*)k.st, (float*)nullptr);
    if ((rc = run_cycle(A, k, sD, fa, k.X, s))) return rc;
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

// Mixed-precision gmres_v: Arnoldi cycles with the FP32 SpMV (FP32 copies of
// the scaled Val and of the current Arnoldi vector), FP64 Gram-Schmidt, Givens,
// least squares and solution update; every restart recomputes the FP64 true
// residual u0 = R - A X (iterative refinement) and the solve stops only when
// that true residual is below eps = max(absTol, relTol |R|).
static int gmres_v_mixed(svg_lhs* L, svg_subls* ls, const double* Val, double* R, hipStream_t s)
{
  if (ls->sD < 1 || ls->sD > 250) return set_error(SVG_ERR_ARG, "gmres_v: Krylov dimension must be in [1, 250]");
  auto& k = L->kw;
  int rc = ensure_work(k, 4 * L->nNo, ls->sD);
  if (rc) return rc;
  const int n = k.n, sD = ls->sD;
  if (!k.v32 && (rc = dev_alloc(&k.v32, k.ld * (k.sD + 1)))) return rc;
  if (!L->d_val32 && (rc = dev_alloc(&L->d_val32, (size_t)16 * L->nnz))) return rc;
  const Op A{L, Val, 4};
  if (L->spmvVariant < 0 && (rc = svg_spmv_tune(L, Val, nullptr, s))) return rc;
  auto t0 = std::chrono::steady_clock::now();
  launch("d2f", k_d2f, dim3(8192), dim3(256), 0, s, (size_t)16 * L->nnz, Val, L->d_val32);
  if (L->spmv32Variant < 0 && (rc = spmv32_tune(L, L->d_val32, k.v32, k.X, s))) return rc;
  FinArgs fa{};
  fa.sD = sD;
  fa.absTol = ls->absTol;
  fa.relTol = ls->relTol;
  fa.mode = FIN_NORM_R;
  if ((rc = dot_device(n, R, R, 0, 1, k.partial, k.scal, k.st, fa, k.h, k.cs, k.sn, k.err, k.coef, s))) return rc;
  SVG_HIP(hipMemcpyAsync(k.h_st, k.st, sizeof(GmState), hipMemcpyDeviceToHost, s));
  SVG_HIP(hipStreamSynchronize(s));
  ls->suc = 0;
  ls->iNorm = k.h_st->iNorm;
  ls->fNorm = k.h_st->fNorm;
  ls->itr = 0;
  if (ls->iNorm <= ls->absTol) {
    ls->callD = DBL_EPSILON;
    ls->dB = 0.0;
    return SVG_OK;
  }
  svg_fill(k.X, (size_t)n, 0.0, s);
  const unsigned nbv = blocks_for(n, 256);
  double dB = ls->fNorm;
  for (int l = 0; l < ls->mItr; l++) {
    dB = ls->fNorm;
    *k.h_done = 0;
    op_apply(A, k.X, k.u, R, nullptr, s);  // FP64 true residual
    fa.mode = FIN_RESTART;
    fa.inc = 1;
    fa.first = 0;
    fa.tcheck = 1;
    if ((rc = dot_device(n, k.u, k.u, 0, 1, k.partial, k.scal, k.st, fa, k.h, k.cs, k.sn, k.err, k.coef, s,
                         k.d_done)))
      return rc;
    launch("div", k_div, dim3(nbv), dim3(256), 0, s, n, k.u, (const GmState*)k.st, k.v32);  // u_0 and v_0
    if ((rc = run_cycle(A, k, sD, fa, k.X, s, L->d_val32))) return rc;
    if (k.h_st->divzero) return set_error(SVG_ERR_ARG, "Array Divide by zero.");
    ls->itr = k.h_st->itr;
    ls->fNorm = k.h_st->fNorm;
    if (k.h_st->tconv) {
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

// Flexible GMRES with the learned preconditioner on the diagonally scaled
// system: z_i = M(u_i), w = A z_i, the reference's Gram-Schmidt and Givens
// steps unchanged, X += sum y_i z_i; every restart recomputes the FP64 true
// residual R - A X and convergence is decided on it. With `mixed` the SpMV
// uses the FP32 matrix copy and the decoder's FP32 copy of z_i. When |err|
// has reached no new minimum for `window` iterations the cycle stops and the
// solve goes on from the current X with the plain (diagonal-only) GMRES
// restarts of the same precision.
static int gmres_v_learned(svg_lhs* L, svg_subls* ls, const double* Val, double* R, bool mixed, svg_lpc* net,
                           int window, int net_period, int coarse, bool coarse_keep, int* fallback_itr, hipStream_t s)
{
  if (ls->sD < 1 || ls->sD > 250) return set_error(SVG_ERR_ARG, "gmres_v: Krylov dimension must be in [1, 250]");
  auto& k = L->kw;
  int rc = ensure_work(k, 4 * L->nNo, ls->sD);
  if (rc) return rc;
  const int n = k.n, sD = ls->sD;
  // z is sized for the workspace's capacity k.sD (>= sD), like v32: the
  // workspace outlives this call and a later solve may use a larger sD.
  if (!k.z && (rc = dev_alloc(&k.z, k.ld * k.sD))) return rc;
  if (mixed) {
    if (!k.u32 && (rc = dev_alloc(&k.u32, (size_t)n))) return rc;
    if (!k.v32 && (rc = dev_alloc(&k.v32, k.ld * (k.sD + 1)))) return rc;
    if (!L->d_val32 && (rc = dev_alloc(&L->d_val32, (size_t)16 * L->nnz))) return rc;
    if (coarse > 0 && !k.z32 && (rc = dev_alloc(&k.z32, k.ld * k.sD))) return rc;
    if (coarse > 0 && !k.w32 && (rc = dev_alloc(&k.w32, (size_t)n))) return rc;
  }
  if (coarse > 0 && net && !k.r2 && (rc = dev_alloc(&k.r2, (size_t)n))) return rc;
  const Op A{L, Val, 4};
  if (L->spmvVariant < 0 && (rc = svg_spmv_tune(L, Val, nullptr, s))) return rc;
  if (net && (rc = lpc_reserve(net, L, s))) return rc;
  auto t0 = std::chrono::steady_clock::now();
  if (net && (rc = lpc_prepare_dev(L, Val, s))) return rc;
  if (coarse > 0 && (rc = coarse_prepare_dev(L, Val, coarse, s, coarse_keep))) return rc;
  if (mixed) {
    launch("d2f", k_d2f, dim3(8192), dim3(256), 0, s, (size_t)16 * L->nnz, Val, L->d_val32);
    if (L->spmv32Variant < 0 && (rc = spmv32_tune(L, L->d_val32, k.u32, k.X, s))) return rc;
  }
  FinArgs fa{};
  fa.sD = sD;
  fa.absTol = ls->absTol;
  fa.relTol = ls->relTol;
  fa.mode = FIN_NORM_R;
  if ((rc = dot_device(n, R, R, 0, 1, k.partial, k.scal, k.st, fa, k.h, k.cs, k.sn, k.err, k.coef, s))) return rc;
  SVG_HIP(hipMemcpyAsync(k.h_st, k.st, sizeof(GmState), hipMemcpyDeviceToHost, s));
  SVG_HIP(hipStreamSynchronize(s));
  ls->suc = 0;
  ls->iNorm = k.h_st->iNorm;
  ls->fNorm = k.h_st->fNorm;
  ls->itr = 0;
  *fallback_itr = -1;
  if (ls->iNorm <= ls->absTol) {
    ls->callD = DBL_EPSILON;
    ls->dB = 0.0;
    return SVG_OK;
  }
  svg_fill(k.X, (size_t)n, 0.0, s);
  const unsigned nbv = blocks_for(n, 256);
  double dB = ls->fNorm;
  svg_lpc* pc = net;
  bool tl = coarse > 0;
  fa.fbwin = window;
  for (int l = 0; l < ls->mItr; l++) {
    dB = ls->fNorm;
    *k.h_done = 0;
    op_apply(A, k.X, k.u, R, nullptr, s);  // FP64 true residual
    fa.mode = FIN_RESTART;
    fa.inc = 1;
    fa.first = 0;
    fa.tcheck = 1;
    if ((rc = dot_device(n, k.u, k.u, 0, 1, k.partial, k.scal, k.st, fa, k.h, k.cs, k.sn, k.err, k.coef, s,
                         k.d_done)))
      return rc;
    launch("div", k_div, dim3(nbv), dim3(256), 0, s, n, k.u, (const GmState*)k.st,
           mixed ? k.v32 : (float*)nullptr);
    if ((rc = run_cycle(A, k, sD, fa, k.X, s, mixed ? L->d_val32 : nullptr, pc, net_period, tl))) return rc;
    if (k.h_st->divzero) return set_error(SVG_ERR_ARG, "Array Divide by zero.");
    ls->itr = k.h_st->itr;
    ls->fNorm = k.h_st->fNorm;
    if (k.h_st->tconv) {
      ls->suc = 1;
      break;
    }
    if ((pc || tl) && k.h_st->fallback) {
      // stagnation: the network is dropped first (the two-level preconditioner
      // with block Jacobi continues and may fall back in turn), then the
      // two-level preconditioner (the diagonally preconditioned GMRES continues)
      *fallback_itr = ls->itr;
      if (pc && tl) {
        pc = nullptr;
        // re-arm the stagnation check for the two-level stage, measured from
        // its own first cycle (the best estimate of the network stage came
        // from FP32 products and may lie below the true residual restarted from)
        launch("rearm", k_fb_rearm, dim3(1), dim3(64), 0, s, k.st);
      } else {
        pc = nullptr;
        tl = false;
        fa.fbwin = 0;
      }
    }
  }
  SVG_HIP(hipMemcpyAsync(R, k.X, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
  auto t1 = std::chrono::steady_clock::now();
  ls->callD = std::chrono::duration<double>(t1 - t0).count();
  ls->dB = 10.0 * log(ls->fNorm / dB);
  return SVG_OK;
}

// gmres (gmres.cpp:91-280), the inner solver of the NS solver: X = A^-1 R with
// X starting at 0, the first restart vector R itself, ls->itr accumulating.
static int gmres_dev(const Op& A, KrylovWork& k, svg_subls* ls, const double* R, double* X, hipStream_t s)
{
  if (ls->sD < 1 || ls->sD > MAXM - 2) return set_error(SVG_ERR_ARG, "gmres: Krylov dimension must be in [1, 510]");
  const int n = A.dof * A.L->nNo, sD = ls->sD;
  int rc = ensure_work(k, n, sD);
  if (rc) return rc;
  auto t0 = std::chrono::steady_clock::now();
  const unsigned nbv = blocks_for(n, 256);
  ls->suc = 0;
  svg_fill(X, (size_t)n, 0.0, s);
  FinArgs fa{};
  fa.sD = sD;
  fa.absTol = ls->absTol;
  fa.relTol = ls->relTol;
  launch("reset_state", k_reset_state, dim3(1), dim3(64), 0, s, k.st);
  int itr0 = ls->itr;
  for (int l = 0; l < ls->mItr; l++) {
    *k.h_done = 0;
    if (l == 0) {
      SVG_HIP(hipMemcpyAsync(k.u, R, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
      if (A.pre) add_bc_pre_dev(A.L, A.dof, k.u, k.u, nullptr, s);  // gmres.cpp:142-150
    } else {
      op_apply(A, X, k.u, R, nullptr, s);  // u0 = R - A X (with op PRE)
    }
    fa.mode = FIN_RESTART;
    fa.inc = l == 0 ? 0 : 1;
    fa.first = l == 0 ? 1 : 0;
    if ((rc = dot_device(n, k.u, k.u, 0, 1, k.partial, k.scal, k.st, fa, k.h, k.cs, k.sn, k.err, k.coef, s,
                         k.d_done)))
      return rc;
    launch("div", k_div, dim3(nbv), dim3(256), 0, s, n, k.u, (const GmState*)k.st, (float*)nullptr);
    double dBprev = ls->fNorm;
    if ((rc = run_cycle(A, k, sD, fa, X, s))) return rc;
    if (l == 0 && k.h_st->early) {
      ls->callD = DBL_EPSILON;
      ls->dB = 0.0;
      return SVG_OK;
    }
    if (k.h_st->divzero) return set_error(SVG_ERR_ARG, "Array Divide by zero.");
    if (l == 0) {
      ls->iNorm = k.h_st->iNorm;
      dBprev = k.h_st->iNorm;
    }
    ls->dB = dBprev;
    ls->itr = itr0 + k.h_st->itr;
    ls->fNorm = k.h_st->fNorm;
    if (k.h_st->suc) {
      ls->suc = 1;
      break;
    }
  }
  auto t1 = std::chrono::steady_clock::now();
  ls->callD = std::chrono::duration<double>(t1 - t0).count() + ls->callD;
  ls->dB = 10.0 * log(ls->fNorm / ls->dB);
  return SVG_OK;
}

// cgrad::schur (cgrad.cpp:50-160) on P in place: P <- (L - D G)^-1 P.
static int schur_dev(svg_lhs* L, svg_subls* ls, double* Pio, hipStream_t s)
{
  auto& w = L->ns;
  const bool coupled = has_coupled(L);
  const int nNo = L->nNo;
  const int nb = (int)blocks_for(nNo, RED_SPAN);
  double* part1 = w.partial;
  double* part2 = w.partial + nb;
  auto t0 = std::chrono::steady_clock::now();
  ls->suc = 0;
  launch("cg_init", k_cg_init, dim3(nb), dim3(RED_BS), 0, s, nNo, (const double*)Pio, w.cgX, w.cgP, part1);
  launch("cg_start", k_cg_start, dim3(1), dim3(RED_BS), 0, s, (const double*)part1, nb, w.cg, ls->absTol,
         ls->relTol);
  auto& kk = L->kw3;  // its look-ahead events and mapped flag
  SVG_HIP(hipMemcpyAsync(w.h_cg, w.cg, sizeof(CgState), hipMemcpyDeviceToHost, s));
  SVG_HIP(hipStreamSynchronize(s));
  const int* skip = &w.cg->stop;
  if (!w.h_cg->stop) {
    *kk.h_done = 0;
    for (int i = 0; i < ls->mItr; i++) {
      if (i >= LOOKAHEAD) {
        SVG_HIP(hipEventSynchronize(kk.ev[(i - LOOKAHEAD) % 8]));
        if (*(volatile int*)kk.h_done) break;
      }
      // GP = G P (sv); SP = L P - D GP; alpha, X, R; err, P
      spmv<3, 1, 0>(L, w.mG, w.cgP, w.cgGP, nullptr, skip, s, "spmv_sv");
      if (coupled) {  // op PRE on G P, reading a copy of it (cgrad.cpp:103-108)
        SVG_HIP(hipMemcpyAsync(w.cgGP0, w.cgGP, sizeof(double) * 3 * nNo, hipMemcpyDeviceToDevice, s));
        add_bc_pre_dev(L, 3, w.cgGP0, w.cgGP, skip, s);
      }
      launch("cg_sp", k_cg_sp, dim3(nb), dim3(RED_BS), 0, s, nNo, (const int*)L->d_rowPtr, (const int*)L->d_colPtr,
             (const double*)w.Gt, (const double*)w.mL, (const double*)w.cgGP, (const double*)w.cgP, w.cgSP, part1,
             (const CgState*)w.cg);
      launch("cg_upd", k_cg_upd, dim3(nb), dim3(RED_BS), 0, s, nNo, w.cgX, (const double*)w.cgP, Pio,
             (const double*)w.cgSP, (const double*)part1, part2, nb, (const CgState*)w.cg, i);
      launch("cg_p", k_cg_p, dim3(nb), dim3(RED_BS), 0, s, nNo, w.cgP, (const double*)Pio, (const double*)part2, nb,
             w.cg, i, ls->mItr, (volatile int*)kk.d_done);
      SVG_HIP(hipEventRecord(kk.ev[i % 8], s));
    }
  }
  SVG_HIP(hipGetLastError());
  SVG_HIP(hipMemcpyAsync(w.h_cg, w.cg, sizeof(CgState), hipMemcpyDeviceToHost, s));
  SVG_HIP(hipMemcpyAsync(Pio, w.cgX, sizeof(double) * nNo, hipMemcpyDeviceToDevice, s));
  SVG_HIP(hipStreamSynchronize(s));
  const CgState& c = *w.h_cg;
  ls->iNorm = c.iNorm;
  ls->suc = c.suc;
  ls->fNorm = sqrt(c.err);
  auto t1 = std::chrono::steady_clock::now();
  ls->callD = std::chrono::duration<double>(t1 - t0).count() + ls->callD;
  ls->itr = ls->itr + c.last_i;
  if (c.errO_last < DBL_EPSILON)
    ls->dB = 0.0;
  else
    ls->dB = 5.0 * log(c.err / c.errO_last);
  return SVG_OK;
}

static int ns_grow(svg_lhs* L, int need)
{
  auto& w = L->ns;
  if (need <= w.cap) return SVG_OK;
  int cap = w.cap ? w.cap : 8;
  while (cap < need) cap *= 2;
  double *U, *MU, *P, *MP;
  int rc;
  if ((rc = dev_alloc(&U, w.ld3 * cap)) || (rc = dev_alloc(&MU, w.ld3 * 2 * cap)) || (rc = dev_alloc(&P, w.ld1 * cap)) ||
      (rc = dev_alloc(&MP, w.ld1 * 2 * cap)))
    return rc;
  if (w.cap) {
    SVG_HIP(hipMemcpy(U, w.U, sizeof(double) * w.ld3 * w.cap, hipMemcpyDeviceToDevice));
    SVG_HIP(hipMemcpy(MU, w.MU, sizeof(double) * w.ld3 * 2 * w.cap, hipMemcpyDeviceToDevice));
    SVG_HIP(hipMemcpy(P, w.P, sizeof(double) * w.ld1 * w.cap, hipMemcpyDeviceToDevice));
    SVG_HIP(hipMemcpy(MP, w.MP, sizeof(double) * w.ld1 * 2 * w.cap, hipMemcpyDeviceToDevice));
  }
  dev_free(w.U);
  dev_free(w.MU);
  dev_free(w.P);
  dev_free(w.MP);
  w.U = U;
  w.MU = MU;
  w.P = P;
  w.MP = MP;
  w.cap = cap;
  double* coef;
  if ((rc = dev_alloc(&coef, (size_t)2 * cap))) return rc;
  dev_free(w.coef);
  w.coef = coef;
  return SVG_OK;
}

static int ns_setup(svg_lhs* L)
{
  auto& w = L->ns;
  const int nNo = L->nNo, nnz = L->nnz;
  if (w.mK && w.nNo == nNo) return SVG_OK;
  int rc0 = lhs_ensure_trans(L);
  if (rc0) return rc0;
  w.nNo = nNo;
  w.ld3 = ((size_t)3 * nNo + 63) & ~(size_t)63;
  w.ld1 = ((size_t)nNo + 63) & ~(size_t)63;
  int rc;
  if ((rc = dev_alloc(&w.mK, (size_t)9 * nnz)) || (rc = dev_alloc(&w.mG, (size_t)3 * nnz)) ||
      (rc = dev_alloc(&w.mD, (size_t)3 * nnz)) || (rc = dev_alloc(&w.mL, (size_t)nnz)) ||
      (rc = dev_alloc(&w.Gt, (size_t)3 * nnz)) || (rc = dev_alloc(&w.Rm, w.ld3)) || (rc = dev_alloc(&w.Rmi, w.ld3)) ||
      (rc = dev_alloc(&w.Rc, w.ld1)) || (rc = dev_alloc(&w.Rci, w.ld1)) || (rc = dev_alloc(&w.cgX, w.ld1)) ||
      (rc = dev_alloc(&w.cgP, w.ld1)) || (rc = dev_alloc(&w.cgSP, w.ld1)) || (rc = dev_alloc(&w.cgGP, w.ld3)) ||
      (rc = dev_alloc(&w.cgGP0, w.ld3)) ||
      (rc = dev_alloc(&w.cg, 1)) || (rc = dev_alloc(&w.dots, MAXPAIRS)))
    return rc;
  w.partialCap = (int)blocks_for(3 * nNo, RED_SPAN) * MAXPAIRS + (int)blocks_for(nNo, RED_BS);
  if ((rc = dev_alloc(&w.partial, (size_t)w.partialCap))) return rc;
  SVG_HIP(hipMemset(w.cg, 0, sizeof(CgState)));
  SVG_HIP(hipHostMalloc((void**)&w.h_cg, sizeof(CgState), hipHostMallocDefault));
  SVG_HIP(hipHostMalloc((void**)&w.h_dots, sizeof(double) * MAXPAIRS, hipHostMallocDefault));
  return ns_grow(L, 8);
}

static int dot_list(svg_lhs* L, const DotList& dl, hipStream_t s)
{
  auto& w = L->ns;
  int nmax = 0;
  for (int p = 0; p < dl.np; p++) nmax = std::max(nmax, dl.p[p].n);
  const unsigned nb = std::max(1u, blocks_for(nmax, RED_SPAN));
  launch("dot_list", k_dot_list, dim3(nb, dl.np), dim3(RED_BS), 0, s, dl, w.partial);
  launch("dot_list_fin", k_dot_list_fin, dim3(1), dim3(FIN_BS), 0, s, dl.np, (int)nb, (const double*)w.partial,
         w.dots);
  SVG_HIP(hipGetLastError());
  SVG_HIP(hipMemcpyAsync(w.h_dots, w.dots, sizeof(double) * dl.np, hipMemcpyDeviceToHost, s));
  SVG_HIP(hipStreamSynchronize(s));
  return SVG_OK;
}

// ge (ge.cpp:40-139): symmetrically scaled Gaussian elimination with partial
// pivoting of the NS outer projection system (N <= 2 RI.mItr); host-side, tiny.
static bool ge_host(int nV, int N, const std::vector<double>& A, std::vector<double>& B)
{
  std::vector<double> W(N > 0 ? N : 1);
  const double tol = DBL_TRUE_MIN, eps = DBL_EPSILON;
  for (int i = 0; i < N; i++) {
    if (fabs(A[i + (size_t)nV * i]) < tol) {
      std::fill(B.begin(), B.end(), 0.0);
      return false;
    }
    W[i] = 1.0 / sqrt(fabs(A[i + (size_t)nV * i]));
  }
  std::vector<double> Cm((size_t)(N > 0 ? N : 1) * (N + 1));
  auto C = [&](int i, int j) -> double& { return Cm[i + (size_t)N * j]; };
  for (int i = 0; i < N; i++) {
    for (int j = 0; j < N; j++) C(i, j) = W[i] * W[j] * A[i + (size_t)nV * j];
    C(i, N) = W[i] * B[i];
  }
  if (N <= 0) return false;
  if (N == 1) {
    B[0] = C(0, 1) / C(0, 0);
    B[0] = B[0] * W[0];
    return true;
  }
  if (N == 2) {
    double pivot = C(0, 0) * C(1, 1) - C(1, 0) * C(0, 1);
    if (fabs(pivot) < eps) {
      std::fill(B.begin(), B.end(), 0.0);
      return false;
    }
    B[0] = (C(0, 2) * C(1, 1) - C(1, 2) * C(0, 1)) / pivot;
    B[1] = (C(1, 2) * C(0, 0) - C(0, 2) * C(1, 0)) / pivot;
    B[0] = W[0] * B[0];
    B[1] = W[1] * B[1];
    return true;
  }
  for (int m = 0; m < N - 1; m++) {
    int ipv = m;
    double pivot = fabs(C(m, m));
    for (int i = m + 1; i < N; i++)
      if (fabs(C(i, m)) > pivot) {
        ipv = i;
        pivot = fabs(C(i, m));
      }
    if (fabs(pivot) < eps) {
      std::fill(B.begin(), B.end(), 0.0);
      return false;
    }
    if (ipv != m)
      for (int j = m; j < N + 1; j++) std::swap(C(m, j), C(ipv, j));
    for (int i = m + 1; i < N; i++) {
      double sv = C(i, m) / C(m, m);
      C(i, m) = 0.0;
      for (int j = m + 1; j < N + 1; j++) C(i, j) = C(i, j) - sv * C(m, j);
    }
  }
  for (int j = N - 1; j >= 0; j--) {
    for (int i = j + 1; i < N; i++) C(j, N) = C(j, N) - C(j, i) * C(i, N);
    C(j, N) = C(j, N) / C(j, j);
  }
  for (int i = 0; i < N; i++) B[i] = W[i] * C(i, N);
  return true;
}

// ns_solver (ns_solver.cpp:167-499) on the preconditioned system; Ri(4,nNo) in/out.
static int ns_solver_dev(svg_lhs* L, svg_ls* ls, const double* Val, double* Ri, hipStream_t s)
{
  const int nNo = L->nNo, nnz = L->nnz, nsd = 3;
  int rc = ns_setup(L);
  if (rc) return rc;
  auto& w = L->ns;
  const int mItr = ls->RI.mItr;
  const int nB = 2 * mItr;
  if (mItr < 1) return set_error(SVG_ERR_ARG, "ns_solver: RI.mItr must be >= 1");
  const unsigned nb1 = blocks_for(nNo, 256), nb3 = blocks_for(3 * nNo, 256);
  auto t0 = std::chrono::steady_clock::now();
  launch("ns_split", k_split, dim3(nb1), dim3(256), 0, s, nNo, (const double*)Ri, w.Rmi, w.Rci);
  SVG_HIP(hipMemcpyAsync(w.Rm, w.Rmi, sizeof(double) * 3 * nNo, hipMemcpyDeviceToDevice, s));
  SVG_HIP(hipMemcpyAsync(w.Rc, w.Rci, sizeof(double) * nNo, hipMemcpyDeviceToDevice, s));
  DotList dl{};
  dl.np = 2;
  dl.p[0] = {w.Rm, w.Rm, 3 * nNo};
  dl.p[1] = {w.Rc, w.Rc, nNo};
  if ((rc = dot_list(L, dl, s))) return rc;
  double eps = sqrt(pow(sqrt(w.h_dots[0]), 2.0) + pow(sqrt(w.h_dots[1]), 2.0));
  ls->RI.iNorm = eps;
  ls->RI.fNorm = eps * eps;
  ls->CG.callD = 0.0;
  ls->GM.callD = 0.0;
  ls->CG.itr = 0;
  ls->GM.itr = 0;
  ls->RI.suc = 0;
  eps = fmax(ls->RI.absTol, ls->RI.relTol * eps);
  launch("depart", k_depart, dim3(blocks_for(nnz, 256)), dim3(256), 0, s, nnz, Val, (const int*)L->d_trans, w.mK,
         w.mG, w.mD, w.mL, w.Gt);
  // bc_pre (ns_solver.cpp:56-89; one rank, no face is shared): nS = |valM|^2 in the
  // reference's order, on the host from the device valM
  for (auto& F : L->face) {
    if (!F.coupledFlag) continue;
    std::vector<double> vm((size_t)F.dof * F.nNo);
    SVG_HIP(hipMemcpyAsync(vm.data(), F.d_valM, sizeof(double) * vm.size(), hipMemcpyDeviceToHost, s));
    SVG_HIP(hipStreamSynchronize(s));
    F.nS = 0.0;
    for (int a = 0; a < F.nNo; a++)
      for (int i = 0; i < 3; i++) F.nS = F.nS + std::pow(vm[i + (size_t)F.dof * a], 2.0);
  }
  const Op K3{L, w.mK, 3, true};
  std::vector<double> A((size_t)nB * nB, 0.0), B(nB, 0.0), xB(nB, 0.0), oldxB(nB, 0.0), tmp;
  int iBB = 0, i_count = 0;
  for (int i = 0; i < mItr; i++) {
    const int iB = 2 * i;
    iBB = 2 * i + 1;
    ls->RI.dB = ls->RI.fNorm;
    i_count = i;
    if ((rc = ns_grow(L, i + 1))) return rc;
    double* Ui = w.U + w.ld3 * i;
    double* Pi = w.P + w.ld1 * i;
    double* MUiB = w.MU + w.ld3 * iB;
    double* MUiBB = w.MU + w.ld3 * iBB;
    double* MPiB = w.MP + w.ld1 * iB;
    double* MPiBB = w.MP + w.ld1 * iBB;
    // U_i = K^-1 Rm; P_i = Rc - D U_i; P_i = S^-1 P_i
    if ((rc = gmres_dev(K3, L->kw3, &ls->GM, w.Rm, Ui, s))) return rc;
    spmv<1, 3, 1>(L, w.mD, Ui, Pi, w.Rc, nullptr, s, "spmv_vs");
    if ((rc = schur_dev(L, &ls->CG, Pi, s))) return rc;
    // MU_iB = G P_i; MU_iBB = Rm - MU_iB; U_i = K^-1 MU_iBB; MU_iBB = K U_i
    spmv<3, 1, 0>(L, w.mG, Pi, MUiB, nullptr, nullptr, s, "spmv_sv");
    launch("sub", k_sub, dim3(nb3), dim3(256), 0, s, 3 * nNo, (const double*)w.Rm, (const double*)MUiB, MUiBB);
    if ((rc = gmres_dev(K3, L->kw3, &ls->GM, MUiBB, Ui, s))) return rc;
    spmv<3, 3, 0>(L, w.mK, Ui, MUiBB, nullptr, nullptr, s, "spmv_k3");
    add_bc_mul_dev(L, 3, Ui, MUiBB, nullptr, s);  // ns_solver.cpp:334
    // MP_iB = L P_i; MP_iBB = D U_i
    spmv<1, 1, 0>(L, w.mL, Pi, MPiB, nullptr, nullptr, s, "spmv_ss");
    spmv<1, 3, 0>(L, w.mD, Ui, MPiBB, nullptr, nullptr, s, "spmv_vs");
    // Gram entries (ns_solver.cpp:350-360): <MU_j,MU_k> + <MP_j,MP_k>, <MU_k,Rmi> + <MP_k,Rci>
    tmp.assign((size_t)2 * (iBB + 2) + 2, 0.0);
    int c = 0;
    std::vector<std::pair<int, int>> idx;  // (dot of MU or Rmi, dot of MP or Rci)
    DotList q{};
    q.np = 0;
    auto flush = [&]() -> int {
      if (q.np == 0) return SVG_OK;
      int r = dot_list(L, q, s);
      if (r) return r;
      for (int p = 0; p < q.np; p++) tmp.push_back(w.h_dots[p]);
      q.np = 0;
      return SVG_OK;
    };
    tmp.clear();
    for (int kk = iB; kk <= iBB; kk++) {
      for (int j = 0; j <= kk + 1; j++) {
        const bool rhs = (j == kk + 1);
        if (q.np + 2 > MAXPAIRS && (rc = flush())) return rc;
        q.p[q.np++] = {rhs ? w.Rmi : w.MU + w.ld3 * j, w.MU + w.ld3 * kk, 3 * nNo};
        q.p[q.np++] = {rhs ? w.Rci : w.MP + w.ld1 * j, w.MP + w.ld1 * kk, nNo};
      }
    }
    if ((rc = flush())) return rc;
    c = 0;
    int tp = 0;
    for (int kk = iB; kk <= iBB; kk++) {
      for (int j = 0; j <= kk; j++) {
        double v = tmp[tp] + tmp[tp + 1];
        tp += 2;
        A[j + (size_t)nB * kk] = v;
        A[kk + (size_t)nB * j] = v;
        c++;
      }
      B[kk] = tmp[tp] + tmp[tp + 1];
      tp += 2;
      c++;
    }
    xB = B;
    if (ge_host(nB, iBB + 1, A, xB)) {
      oldxB = xB;
    } else {
      return set_error(SVG_ERR_SINGULAR, "FSILS: Singular matrix detected");
    }
    double sum = 0.0;
    for (int q2 = 0; q2 <= iBB; q2++) sum += xB[q2] * B[q2];
    ls->RI.fNorm = pow(ls->RI.iNorm, 2.0) - sum;
    if (ls->RI.fNorm < eps * eps) {
      ls->RI.suc = 1;
      break;
    }
    SVG_HIP(hipMemcpyAsync(w.coef, xB.data(), sizeof(double) * (iBB + 1), hipMemcpyHostToDevice, s));
    launch("axpy_sub", k_axpy_sub, dim3(nb3), dim3(256), 0, s, 3 * nNo, (const double*)w.Rmi, (const double*)w.MU,
           w.ld3, iBB + 1, (const double*)w.coef, w.Rm);
    launch("axpy_sub", k_axpy_sub, dim3(nb1), dim3(256), 0, s, nNo, (const double*)w.Rci, (const double*)w.MP, w.ld1,
           iBB + 1, (const double*)w.coef, w.Rc);
  }
  SVG_HIP(hipMemcpyAsync(w.coef, xB.data(), sizeof(double) * (iBB + 1), hipMemcpyHostToDevice, s));
  if (i_count >= mItr) {
    ls->RI.itr = mItr;
  } else {
    ls->RI.itr = i_count;
    launch("axpy_sub", k_axpy_sub, dim3(nb1), dim3(256), 0, s, nNo, (const double*)w.Rci, (const double*)w.MP, w.ld1,
           iBB + 1, (const double*)w.coef, w.Rc);
  }
  DotList rn{};
  rn.np = 1;
  rn.p[0] = {w.Rc, w.Rc, nNo};
  if ((rc = dot_list(L, rn, s))) return rc;
  ls->Resc = (int)(100.0 * pow(sqrt(w.h_dots[0]), 2.0) / ls->RI.fNorm);
  ls->Resm = 100 - ls->Resc;
  // solution (ns_solver.cpp:480-488)
  launch("combine", k_combine, dim3(nb3), dim3(256), 0, s, 3 * nNo, (const double*)w.U, w.ld3, ls->RI.itr,
         (const double*)w.coef, 1, w.Rmi);
  launch("combine", k_combine, dim3(nb1), dim3(256), 0, s, nNo, (const double*)w.P, w.ld1, ls->RI.itr,
         (const double*)w.coef, 0, w.Rci);
  auto t1 = std::chrono::steady_clock::now();
  ls->RI.callD = std::chrono::duration<double>(t1 - t0).count();
  ls->RI.dB = 5.0 * log(ls->RI.fNorm / ls->RI.dB);
  if (ls->Resc < 0.0 || ls->Resm < 0.0) {
    ls->Resc = 0;
    ls->Resm = 0;
    ls->RI.dB = 0;
    ls->RI.fNorm = 0.0;
    return set_error(SVG_ERR_ILLCOND,
                     "FSILS: unexpected behavior in FSILS (likely due to the ill-conditioned LHS matrix)");
  }
  ls->RI.fNorm = sqrt(ls->RI.fNorm);
  launch("ns_join", k_interleave, dim3(nb1), dim3(256), 0, s, nNo, (const double*)w.Rmi, (const double*)w.Rci, Ri);
  SVG_HIP(hipGetLastError());
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


// --------------------------------------------------------- batched GMRES --
// Independent systems s = 0..nsys-1 on one block-diagonal union (rows
// [4 nodeOff[s], 4 nodeOff[s+1])): the union SpMV serves all of them, every
// reduction is per system over 1024-element spans that start at the system's
// first row (the partition a solve of the system alone uses, so each system's
// arithmetic is the single-system gmres_v's, bit for bit), and every system
// has its own Hessenberg, Givens, state and convergence. One restart cycle of
// all systems is captured once into a hipGraph and replayed.
__global__ __launch_bounds__(RED_BS) void k_bmdot(const int* __restrict__ spanStart, const int* __restrict__ spanEnd,
                                                  const int* __restrict__ spanSys, int nsp,
                                                  const double* __restrict__ w, const double* __restrict__ base,
                                                  size_t ld, int m, double* __restrict__ partial,
                                                  const GmState* __restrict__ st, int check)
{
#pragma clang fp contract(off)
  extern __shared__ double lds[];  // [RED_NW waves][m]
  const int b = blockIdx.x;
  if (check) {  // 1: skip systems that stopped iterating; 2: also those done in this cycle
    const GmState& S = st[spanSys[b]];
    if (!S.act || (check == 2 && S.done)) return;
  }
  const int t = threadIdx.x;
  span_dots<1>(w, base, ld, m, false, (size_t)spanStart[b] + t, (size_t)spanEnd[b], lds, m);
  __syncthreads();
  for (int j = t; j < m; j += RED_BS) {
    double v = lds[j];
    for (int q = 1; q < RED_NW; q++) v = v + lds[q * m + j];
    partial[(size_t)j * nsp + b] = v;
  }
}

// k_mdot_cb per span of the batched solve (FP32 basis, then w.w as column m).
__global__ __launch_bounds__(RED_BS) void k_bmdot_cb(const int* __restrict__ spanStart,
                                                     const int* __restrict__ spanEnd, const int* __restrict__ spanSys,
                                                     int nsp, const double* __restrict__ w,
                                                     const float* __restrict__ base, size_t ld, int m,
                                                     double* __restrict__ partial, const GmState* __restrict__ st)
{
#pragma clang fp contract(off)
  extern __shared__ double lds[];  // [RED_NW waves][m + 1]
  const int b = blockIdx.x, mc = m + 1;
  {
    const GmState& S = st[spanSys[b]];
    if (!S.act || S.done) return;
  }
  const int t = threadIdx.x;
  span_dots<MDOT_CB_G>(w, base, ld, m, true, (size_t)spanStart[b] + t, (size_t)spanEnd[b], lds, mc);
  __syncthreads();
  for (int j = t; j < mc; j += RED_BS) {
    double v = lds[j];
    for (int q = 1; q < RED_NW; q++) v = v + lds[q * mc + j];
    partial[(size_t)j * nsp + b] = v;
  }
}

// The FP32 SpMV of the batched solve: rows of systems that converged or
// stopped skip (their basis vectors are no longer read).

// One workgroup per system: the finaliser on the system's span partials.
__global__ __launch_bounds__(FIN_BS) void k_bfin(FinArgs A, const int* __restrict__ sysSpanOff, int nsp,
                                                 const double* __restrict__ partial, int m, GmState* st,
                                                 double* __restrict__ h, double* __restrict__ cs,
                                                 double* __restrict__ sn, double* __restrict__ err,
                                                 double* __restrict__ coef)
{
  __shared__ FinLds S;
  const int sy = blockIdx.x, sD = A.sD;
  GmState* g = st + sy;
  if (A.mode != FIN_NORM_R && !g->act) return;
  const int off = sysSpanOff[sy], nb = sysSpanOff[sy + 1] - off;
  fin_body(A, partial + off, (size_t)nsp, nb, m, g, h + (size_t)sy * (sD + 1) * sD, cs + (size_t)sy * sD,
           sn + (size_t)sy * sD, err + (size_t)sy * (sD + 1), coef + (size_t)sy * (sD + 2), nullptr, nullptr, S);
}

__global__ __launch_bounds__(256) void k_bgs_update(int n, double* __restrict__ u, size_t ld, int i, int sD,
                                                    const double* __restrict__ coef, const GmState* __restrict__ st,
                                                    const int* __restrict__ sysOfNode, float* __restrict__ u32,
                                                    double* __restrict__ z)
{
#pragma clang fp contract(off)
  const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (size_t)n) return;
  const int sy = sysOfNode[k >> 2];
  const GmState& S = st[sy];
  if (!S.act || S.done) return;
  const double* c = coef + (size_t)sy * (sD + 2);
  double* w = u + (size_t)(i + 1) * ld;
  double v = w[k];
  for (int j = 0; j <= i; j++) v = v + (-c[j]) * u[(size_t)j * ld + k];
  v = S.scale * v;
  w[k] = v;
  if (u32) u32[k] = (float)v;
  if (z) z[k] = v;  // flexible solve: the next step is one the network does not precondition
}

// k_gs_update_cb per system of the batched solve (each node's 4 rows belong to one system).
__global__ __launch_bounds__(256) void k_bgs_update_cb(int n, double* __restrict__ w, float* __restrict__ v32,
                                                       size_t ld, int i, int sD, const double* __restrict__ coef,
                                                       const GmState* __restrict__ st,
                                                       const int* __restrict__ sysOfNode, int keep64)
{
#pragma clang fp contract(off)
  const size_t k = 2 * ((size_t)blockIdx.x * blockDim.x + threadIdx.x);
  if (k >= (size_t)n) return;
  const int sy = sysOfNode[k >> 2];
  const GmState& S = st[sy];
  if (!S.act || S.done) return;
  const double* c = coef + (size_t)sy * (sD + 2);
  double2 v = *reinterpret_cast<const double2*>(w + k);
  for (int j = 0; j <= i; j++) {
    const float2 b = *reinterpret_cast<const float2*>(v32 + (size_t)j * ld + k);
    const double cj = -c[j];
    v.x = v.x + cj * (double)b.x;
    v.y = v.y + cj * (double)b.y;
  }
  v.x = S.scale * v.x;
  v.y = S.scale * v.y;
  *reinterpret_cast<float2*>(v32 + (size_t)(i + 1) * ld + k) = make_float2((float)v.x, (float)v.y);
  if (keep64) *reinterpret_cast<double2*>(w + k) = v;
}

__global__ __launch_bounds__(256) void k_bdiv(int n, double* __restrict__ u, const GmState* __restrict__ st,
                                              const int* __restrict__ sysOfNode, float* __restrict__ u32)
{
  const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (size_t)n) return;
  const GmState& S = st[sysOfNode[k >> 2]];
  if (!S.act || S.done) return;
  const double v = u[k] / S.scale;
  u[k] = v;
  if (u32) u32[k] = (float)v;
}

__global__ __launch_bounds__(256) void k_bbacksolve(int sD, const double* __restrict__ h,
                                                    const double* __restrict__ err, double* __restrict__ y,
                                                    GmState* st)
{
  __shared__ double lh[BS_LDS];
  __shared__ double ly[MAXM];
  const int sy = blockIdx.x;
  if (!st[sy].act) return;
  backsolve_body(sD, h + (size_t)sy * (sD + 1) * sD, err + (size_t)sy * (sD + 1), y + (size_t)sy * sD, st + sy, lh,
                 ly);
}

__global__ __launch_bounds__(256) void k_bxupdate(int n, double* __restrict__ X, const double* __restrict__ u,
                                                  size_t ld, int sD, const double* __restrict__ y,
                                                  const GmState* __restrict__ st, const int* __restrict__ sysOfNode)
{
#pragma clang fp contract(off)
  const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (size_t)n) return;
  const int sy = sysOfNode[k >> 2];
  const GmState& S = st[sy];
  if (!S.act || S.early || S.divzero || S.tconv) return;
  const double* ys = y + (size_t)sy * sD;
  double x = X[k];
  for (int j = 0; j <= S.last_i; j++) x = x + ys[j] * u[(size_t)j * ld + k];
  X[k] = x;
}

// k_xupdate_cb per system of the batched solve.
__global__ __launch_bounds__(256) void k_bxupdate_cb(int n, double* __restrict__ X, const float* __restrict__ v32,
                                                     const double* __restrict__ z, int period, size_t ld, int sD,
                                                     const double* __restrict__ y, const GmState* __restrict__ st,
                                                     const int* __restrict__ sysOfNode,
                                                     const float* __restrict__ z32)
{
#pragma clang fp contract(off)
  const size_t k = 2 * ((size_t)blockIdx.x * blockDim.x + threadIdx.x);
  if (k >= (size_t)n) return;
  const int sy = sysOfNode[k >> 2];
  const GmState& S = st[sy];
  if (!S.act || S.early || S.divzero || S.tconv) return;
  const double* ys = y + (size_t)sy * sD;
  double2 x = *reinterpret_cast<const double2*>(X + k);
  for (int j = 0; j <= S.last_i; j++) {
    double2 b;
    if (z32) {
      const float2 f = *reinterpret_cast<const float2*>(z32 + (size_t)j * ld + k);
      b = make_double2((double)f.x, (double)f.y);
    } else if (z && (period <= 1 || j % period == 0)) {
      b = *reinterpret_cast<const double2*>(z + (size_t)j * ld + k);
    } else {
      const float2 f = *reinterpret_cast<const float2*>(v32 + (size_t)j * ld + k);
      b = make_double2((double)f.x, (double)f.y);
    }
    x.x = x.x + ys[j] * b.x;
    x.y = x.y + ys[j] * b.y;
  }
  *reinterpret_cast<double2*>(X + k) = x;
}

// After a cycle: systems that converged stop iterating (plain: the Givens
// estimate, gmres.cpp:585-589; mixed: the FP64 true residual at the restart).
// End of a restart cycle: a system that converged, hit a zero restart norm or
// ran mItr cycles stops; the cycle count and the iteration count after the
// cycle in which the system fell back are kept in its state (the host reads
// the states only between replays, which may hold several cycles).
__global__ void k_bclose(int nsys, GmState* st, int tcheck, int mItr)
{
  const int sy = blockIdx.x * blockDim.x + threadIdx.x;
  if (sy >= nsys) return;
  GmState& S = st[sy];
  const bool ran = S.act != 0;
  if (S.act && (tcheck ? S.tconv : S.suc)) S.act = 0;
  if (S.divzero) S.act = 0;
  if (ran) {
    S.cyc = S.cyc + 1;
    if (S.fallback && S.fbitr < 0) S.fbitr = S.itr;
    if (S.cyc >= mItr) S.act = 0;
  }
}

// Start of a restart cycle: the residual norm the cycle starts from (dB of
// the records is taken against the last cycle's).
__global__ void k_bcycle_start(int nsys, GmState* st)
{
  const int sy = blockIdx.x * blockDim.x + threadIdx.x;
  if (sy >= nsys) return;
  GmState& S = st[sy];
  if (S.act) S.fStart = S.fNorm;
}

// R = X for the systems that iterated; the others keep R (gmres.cpp:489-493).
__global__ __launch_bounds__(256) void k_bresult(int n, double* __restrict__ R, const double* __restrict__ X,
                                                 const GmState* __restrict__ st, const int* __restrict__ sysOfNode,
                                                 double absTol)
{
  const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (size_t)n) return;
  const GmState& S = st[sysOfNode[k >> 2]];
  if (S.iNorm > absTol) R[k] = X[k];
}


static void batch_graph_free(svg_batch* B)
{
  for (auto& gg : B->graph)
    for (auto& g : gg)
      if (g) (void)hipGraphExecDestroy(g), g = nullptr;
  for (hipGraphExec_t* g : {&B->pro, &B->epi})
    if (*g) (void)hipGraphExecDestroy(*g), *g = nullptr;
  B->newtonKey.clear();
}

static void batch_mdot(svg_batch* B, const double* w, const double* base, size_t ld, int m, int check, hipStream_t s)
{
  hipLaunchKernelGGL(k_bmdot, dim3(B->nsp), dim3(RED_BS), (size_t)RED_NW * m * sizeof(double), s,
                     (const int*)B->d_spanStart, (const int*)B->d_spanEnd, (const int*)B->d_spanSys, B->nsp, w, base,
                     ld, m, B->partial, (const GmState*)B->st, check);
}

static void batch_fin(svg_batch* B, const FinArgs& fa, int m, hipStream_t s)
{
  hipLaunchKernelGGL(k_bfin, dim3(B->nsys), dim3(FIN_BS), 0, s, fa, (const int*)B->d_sysSpanOff, B->nsp,
                     (const double*)B->partial, m, B->st, B->h, B->cs, B->sn, B->err, B->coef);
}

// The form of a batched solve: FP32 SpMV (mixed) and the learned
// preconditioner (net, applied at Arnoldi steps i % period == 0, fallback
// after `window` non-decreasing iterations, per system).
struct BatchMode {
  bool mixed = false;
  svg_lpc* net = nullptr;
  int coarse = 0;  // two-level preconditioner: aggregates of about this many nodes (0: none)
  int period = 1, window = 20;
  bool tcheck() const { return mixed || net || coarse; }  // convergence on the FP64 true restart residual
  bool flex() const { return net || coarse; }             // flexible GMRES (per-system fallback)
};

// z_j = u_j (and its FP32 copy) for the systems still iterating in the cycle;
// after a network application (onlyFallback) only for the systems that fell
// back to the diagonal preconditioner, whose flexible basis is then the plain
// one: with the FP32 copy (mixed precision, FP32 basis) z_j is the FP32 basis
// vector's value, as the solve without the network reads it.
__global__ __launch_bounds__(256) void k_bzcopy(int n, const double* __restrict__ u, double* __restrict__ z,
                                                float* __restrict__ z32, const GmState* __restrict__ st,
                                                const int* __restrict__ sysOfNode, int onlyFallback)
{
  const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (size_t)n) return;
  const GmState& S = st[sysOfNode[k >> 2]];
  if (!S.act || S.done || (onlyFallback && !S.fallback)) return;
  const double v = u[k];
  if (z32) {
    const float f = (float)v;
    z32[k] = f;
    z[k] = (double)f;
  } else {
    z[k] = v;
  }
}

// One restart cycle of every system still iterating (the body of gmres_v's
// restart loop, gmres.cpp:529-612; flexible with the learned preconditioner as
// the single-system learned solve), enqueued on s (captured into a graph).
static int batch_cycle(svg_batch* B, const double* Val, const double* R, const BatchMode& md, const FinArgs& base,
                       int mItr, hipStream_t s)
{
  svg_lhs* L = B->L;
  const int n = B->n, sD = B->sD;
  const unsigned nbv = blocks_for(n, 256);
  const Op A{L, Val, 4};
  const bool mixed = md.mixed, flex = md.flex(), tl = md.coarse > 0;
  FinArgs fa = base;
  hipLaunchKernelGGL(k_bcycle_start, dim3(blocks_for(B->nsys, 256)), dim3(256), 0, s, B->nsys, B->st);
  op_apply(A, B->X, B->u, R, nullptr, s);  // u0 = R - A X
  batch_mdot(B, B->u, B->u, 0, 1, 1, s);
  fa.mode = FIN_RESTART;
  fa.inc = 1;
  fa.first = 0;
  fa.tcheck = md.tcheck() ? 1 : 0;
  batch_fin(B, fa, 1, s);
  // mixed precision: the basis in FP32 (B->v32), u_0 and v_0 from the division
  hipLaunchKernelGGL(k_bdiv, dim3(nbv), dim3(256), 0, s, n, B->u, (const GmState*)B->st,
                     (const int*)B->d_sysOfNode, mixed ? B->v32 : (float*)nullptr);
  const unsigned nbv2 = blocks_for(n / 2, 256);
  auto netStep = [&](int j) { return md.net && (md.period <= 1 || j % md.period == 0); };
  for (int i = 0; i < sD; i++) {
    double* ui = B->u + B->ld * i;
    double* un = B->u + B->ld * (i + 1);
    const double* sin = ui;
    if (tl) {
      // z_i = M(u_i) with the systems' dense coarse blocks; the systems that fell
      // back keep z_i = u_i (their flexible basis is the plain one; the apply
      // writes it)
      if (mixed) {
        coarse_apply_batch_dev(L, (const float*)(B->v32 + B->ld * i), nullptr, B->z32 + B->ld * i,
                               (const GmState*)B->st, (const int*)B->d_sysOfNode, s);
      } else {
        double* zi = B->z + B->ld * i;
        coarse_apply_batch_dev(L, (const double*)ui, zi, nullptr, (const GmState*)B->st,
                               (const int*)B->d_sysOfNode, s);
        sin = zi;
      }
    } else if (netStep(i)) {
      double* zi = B->z + B->ld * i;
      float* z32 = mixed ? B->u32 : (float*)nullptr;
      int rc = lpc_apply_dev(md.net, L, ui, zi, z32, nullptr, s, B->d_nsc, B->d_sysOfNode, B->st);
      if (rc) return rc;
      hipLaunchKernelGGL(k_bzcopy, dim3(nbv), dim3(256), 0, s, n, (const double*)ui, zi, z32,
                         (const GmState*)B->st, (const int*)B->d_sysOfNode, 1);
      sin = zi;
    } else if (flex && !mixed) {
      double* zi = B->z + B->ld * i;
      if (i == 0)
        hipLaunchKernelGGL(k_bzcopy, dim3(nbv), dim3(256), 0, s, n, (const double*)ui, zi, (float*)nullptr,
                           (const GmState*)B->st, (const int*)B->d_sysOfNode, 0);
      // else: the previous step's Gram-Schmidt update wrote z_i = u_i
      sin = zi;
    }
    fa.mode = FIN_ARNOLDI;
    fa.i = i;
    if (mixed) {
      spmv32(L, (const float*)L->d_val32,
             tl ? (const float*)(B->z32 + B->ld * i)
                : (netStep(i) ? (const float*)B->u32 : (const float*)(B->v32 + B->ld * i)),
             un, nullptr, s, "bspmv32", (const int*)B->d_sysOfNode, (const GmState*)B->st);
      hipLaunchKernelGGL(k_bmdot_cb, dim3(B->nsp), dim3(RED_BS), (size_t)RED_NW * (i + 2) * sizeof(double), s,
                         (const int*)B->d_spanStart, (const int*)B->d_spanEnd, (const int*)B->d_spanSys, B->nsp,
                         (const double*)un, (const float*)B->v32, B->ld, i + 1, B->partial, (const GmState*)B->st);
      batch_fin(B, fa, i + 2, s);
      hipLaunchKernelGGL(k_bgs_update_cb, dim3(nbv2), dim3(256), 0, s, n, un, B->v32, B->ld, i, sD,
                         (const double*)B->coef, (const GmState*)B->st, (const int*)B->d_sysOfNode,
                         (i + 1 < sD && netStep(i + 1)) ? 1 : 0);
    } else {
      op_apply(A, sin, un, nullptr, nullptr, s);
      batch_mdot(B, un, B->u, B->ld, i + 2, 2, s);
      batch_fin(B, fa, i + 2, s);
      const bool nextPlain = flex && !tl && i + 1 < sD && !netStep(i + 1);
      hipLaunchKernelGGL(k_bgs_update, dim3(nbv), dim3(256), 0, s, n, B->u, B->ld, i, sD, (const double*)B->coef,
                         (const GmState*)B->st, (const int*)B->d_sysOfNode, (float*)nullptr,
                         nextPlain ? B->z + B->ld * (i + 1) : (double*)nullptr);
    }
  }
  hipLaunchKernelGGL(k_bbacksolve, dim3(B->nsys), dim3(256), 0, s, sD, (const double*)B->h, (const double*)B->err,
                     B->y, B->st);
  if (mixed)
    hipLaunchKernelGGL(k_bxupdate_cb, dim3(nbv2), dim3(256), 0, s, n, B->X, (const float*)B->v32,
                       (flex && !tl) ? (const double*)B->z : (const double*)nullptr, md.period, B->ld, sD,
                       (const double*)B->y, (const GmState*)B->st, (const int*)B->d_sysOfNode,
                       tl ? (const float*)B->z32 : (const float*)nullptr);
  else
    hipLaunchKernelGGL(k_bxupdate, dim3(nbv), dim3(256), 0, s, n, B->X, (const double*)(flex ? B->z : B->u), B->ld,
                       sD, (const double*)B->y, (const GmState*)B->st, (const int*)B->d_sysOfNode);
  hipLaunchKernelGGL(k_bclose, dim3(blocks_for(B->nsys, 256)), dim3(256), 0, s, B->nsys, B->st,
                     md.tcheck() ? 1 : 0, mItr);
  SVG_HIP(hipGetLastError());
  return SVG_OK;
}

// The bytes of the arguments a captured graph bakes in (a graph is replayed
// only while they are unchanged).
struct GraphKey {
  std::vector<unsigned char> b;
  template <typename T>
  GraphKey& add(const T& v)
  {
    const unsigned char* p = reinterpret_cast<const unsigned char*>(&v);
    b.insert(b.end(), p, p + sizeof(T));
    return *this;
  }
};

// Captures what enqueue() puts on s into an executable graph; per-launch event
// timing is off while capturing.
template <typename F>
static int capture(hipStream_t s, hipGraphExec_t* out, F enqueue)
{
  const bool prev = prof_set(false);
  hipGraph_t g = nullptr;
  int rc = SVG_OK;
  hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
  if (e == hipSuccess) {
    rc = enqueue();
    hipError_t e2 = hipStreamEndCapture(s, &g);
    if (e == hipSuccess) e = e2;
  }
  prof_set(prev);
  if (!rc && e == hipSuccess) e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
  if (g) (void)hipGraphDestroy(g);
  if (rc) return rc;
  return e == hipSuccess ? SVG_OK : hip_check(e, "graph capture");
}

static FinArgs batch_fin_args(const svg_subls* proto, int sD, const BatchMode& md)
{
  FinArgs fa{};
  fa.sD = sD;
  fa.absTol = proto->absTol;
  fa.relTol = proto->relTol;
  fa.mode = FIN_NORM_R;
  fa.fbwin = md.flex() ? md.window : 0;
  return fa;
}

static int batch_mode(const svg_solve_opts* o, BatchMode* md)
{
  *md = BatchMode{};
  if (!o) return SVG_OK;
  md->mixed = o->precision == SVG_PRECISION_MIXED;
  if (o->method == SVG_METHOD_LEARNED) {
    if (!o->lpc && o->coarse <= 0)
      return set_error(SVG_ERR_ARG, "batched solve: the learned method needs a network (opts.lpc) or a coarse level");
    if (o->lpc && o->coarse > 0)
      return set_error(SVG_ERR_UNSUPPORTED,
                       "batched solve: the network and the two-level preconditioner together are not built");
    md->net = o->lpc;
    md->coarse = o->coarse > 0 ? o->coarse : 0;
    md->period = o->net_period > 1 ? o->net_period : 1;
    md->window = o->fallback_window > 0 ? o->fallback_window : 20;
  } else if (o->method != SVG_METHOD_PLAIN) {
    return set_error(SVG_ERR_ARG, "batched solve: unknown method");
  } else if (o->coarse > 0) {
    // the two-level preconditioner with block-Jacobi smoothing
    md->coarse = o->coarse;
    md->window = o->fallback_window > 0 ? o->fallback_window : 20;
  }
  return SVG_OK;
}

// What must exist before the batched solve is enqueued or captured: the SpMV
// launch variant (tuned once per lhs), the FP32 copies (mixed), the flexible
// basis, per-node scales and network buffers (learned).
static int batch_prepare(svg_batch* B, const double* Val, const BatchMode& md, hipStream_t caller)
{
  svg_lhs* L = B->L;
  int rc;
  if (L->spmvVariant < 0 && (rc = svg_spmv_tune(L, Val, nullptr, caller))) return rc;
  if (md.mixed) {
    if (!L->d_val32 && (rc = dev_alloc(&L->d_val32, (size_t)16 * L->nnz))) return rc;
    if (!B->u32 && (rc = dev_alloc(&B->u32, (size_t)B->n))) return rc;
    if (!B->v32 && (rc = dev_alloc(&B->v32, B->ld * (B->sD + 1)))) return rc;
    // the FP32 SpMV variant for the union (timing only: the values do not matter)
    if (L->spmv32Variant < 0 && (rc = spmv32_tune(L, L->d_val32, B->u32, B->X, caller))) return rc;
  }
  if (md.coarse) {
    if ((rc = coarse_setup(L, md.coarse, B->nodeOff.data(), B->nsys))) return rc;
    if (md.mixed && !B->z32 && (rc = dev_alloc(&B->z32, B->ld * B->sD))) return rc;
    if (!md.mixed && !B->z && (rc = dev_alloc(&B->z, B->ld * B->sD))) return rc;
  }
  if (md.net) {
    if (!B->z && (rc = dev_alloc(&B->z, B->ld * B->sD))) return rc;
    if (!B->d_nsc) {
      std::vector<double2> sc(L->nNo);
      for (int q = 0; q < B->nsys; q++) {
        const double v = std::sqrt((double)(B->nodeOff[q + 1] - B->nodeOff[q]));
        for (int a = B->nodeOff[q]; a < B->nodeOff[q + 1]; a++) sc[a] = make_double2(v, 1.0 / v);
      }
      if ((rc = dev_alloc(&B->d_nsc, (size_t)L->nNo))) return rc;
      SVG_HIP(hipMemcpy(B->d_nsc, sc.data(), sizeof(double2) * L->nNo, hipMemcpyHostToDevice));
    }
    if ((rc = lpc_reserve(md.net, L, caller))) return rc;
  }
  return SVG_OK;
}

// gmres_v before its restart loop (gmres.cpp:466-527) for every system: the
// FP32 matrix copy (mixed), the network's BF16 copies (learned), ||R|| per
// system with eps and the state, X = 0.
static int batch_begin(svg_batch* B, const double* R, const double* Val, const BatchMode& md, const FinArgs& fa,
                       hipStream_t s)
{
  svg_lhs* L = B->L;
  int rc;
  if (md.net && (rc = lpc_prepare_dev(L, Val, s))) return rc;
  if (md.coarse) coarse_enqueue(L, Val, s);
  if (md.mixed) launch("d2f", k_d2f, dim3(8192), dim3(256), 0, s, (size_t)16 * L->nnz, Val, L->d_val32);
  batch_mdot(B, R, R, 0, 1, 0, s);
  batch_fin(B, fa, 1, s);
  int rc2 = svg_fill(B->X, (size_t)B->n, 0.0, s);
  if (rc2) return rc2;
  SVG_HIP(hipGetLastError());
  return SVG_OK;
}

// Restart cycles per replay of the many-cycle graph: the host reads the
// systems' states once per replay. Cycles of systems that have stopped cost
// only their launches (every kernel skips them), so the one-cycle graph takes
// over once no more than an eighth of the systems iterate (on the 256-mesh
// batch: 14 replays per time step instead of 42 cycles, at the same speed;
// four cycles per replay to the end: 12 replays, 2 % slower).
constexpr int kBatchCycles = 4;

// The restart cycle as graphs (R, Val, the tolerances, mItr and the form are
// baked in): graph[p][0] one cycle, graph[p][1] kBatchCycles cycles; p = 0
// serves the FP64 forms, p = 1 the mixed ones.
static int batch_cycle_graph(svg_batch* B, const double* R, const double* Val, const BatchMode& md, const FinArgs& fa,
                             int mItr)
{
  const int gi = md.mixed ? 1 : 0;
  GraphKey k;
  k.add(mItr);
  k.add(R).add(Val).add(fa.absTol).add(fa.relTol).add(fa.fbwin).add(B->L->d_val32).add(B->u32).add(B->v32);
  k.add(B->L->spmvVariant).add(B->L->spmv32Variant).add(md.net).add(md.period).add(B->z).add(B->d_nsc);
  // the network's activation buffers are baked into the captured forward pass
  k.add(lpc_generation(md.net)).add(B->L->msgVariant);
  // the coarse level's buffers are baked in too
  k.add(md.coarse).add(coarse_generation(B->L)).add(B->z32).add(md.window);
  hipStream_t s = B->cs_stream;
  for (int m = 0; m < 2; m++) {
    if (B->graph[gi][m] && B->gKey[gi][m] == k.b) continue;
    if (B->graph[gi][m]) (void)hipGraphExecDestroy(B->graph[gi][m]), B->graph[gi][m] = nullptr;
    const int ncyc = m ? kBatchCycles : 1;
    int rc = capture(s, &B->graph[gi][m], [&]() -> int {
      for (int c = 0; c < ncyc; c++) {
        int r = batch_cycle(B, Val, R, md, fa, mItr, s);
        if (r) return r;
      }
      return SVG_OK;
    });
    if (rc) return rc;
    B->gKey[gi][m] = k.b;
  }
  return SVG_OK;
}

// Replays the restart cycle (enqueued after the begin step on the capture
// stream) until no system iterates or mItr cycles ran; dB[q] is system q's
// residual norm at the start of its last cycle; fbItr[q] the iteration count
// after the cycle in which system q fell back (-1: none).
static int batch_iterate(svg_batch* B, const BatchMode& md, int mItr, std::vector<double>& dB)
{
  hipStream_t s = B->cs_stream;
  const int gi = md.mixed ? 1 : 0;
  SVG_HIP(hipMemcpyAsync(B->h_st, B->st, sizeof(GmState) * B->nsys, hipMemcpyDeviceToHost, s));
  SVG_HIP(hipStreamSynchronize(s));
  // every replay runs at least one cycle of each system still iterating, and a
  // system stops after mItr cycles (k_bclose), so mItr replays suffice
  for (int l = 0; l < mItr; l++) {
    int active = 0;
    for (int q = 0; q < B->nsys; q++) active += B->h_st[q].act;
    if (!active) break;
    SVG_HIP(hipGraphLaunch(B->graph[gi][8 * active > B->nsys ? 1 : 0], s));
    SVG_HIP(hipMemcpyAsync(B->h_st, B->st, sizeof(GmState) * B->nsys, hipMemcpyDeviceToHost, s));
    SVG_HIP(hipStreamSynchronize(s));
    B->replays++;
    for (int q = 0; q < B->nsys; q++)
      if (B->h_st[q].divzero) return set_error(SVG_ERR_ARG, "Array Divide by zero.");
  }
  dB.assign(B->nsys, 0.0);
  B->fbItr.assign(B->nsys, -1);
  for (int q = 0; q < B->nsys; q++) {
    dB[q] = B->h_st[q].fStart;
    if (md.flex()) B->fbItr[q] = B->h_st[q].fbitr;
  }
  return SVG_OK;
}

// The per-system records of ls.RI after a batched solve (h_st is current).
static void batch_records(const svg_batch* B, const svg_subls* proto, const BatchMode& md,
                          const std::vector<double>& dB, double secs, svg_subls* out)
{
  for (int q = 0; q < B->nsys; q++) {
    const GmState& g = B->h_st[q];
    svg_subls& o = out[q];
    o = *proto;
    o.itr = g.itr;
    o.iNorm = g.iNorm;
    o.fNorm = g.fNorm;
    o.suc = md.tcheck() ? g.tconv : g.suc;
    o.callD = secs;
    o.dB = (g.iNorm > proto->absTol && dB[q] > 0.0) ? 10.0 * log(g.fNorm / dB[q]) : 0.0;
  }
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
    case 1: spmv_launch<1, 1, 0, 2, false>(L, K, U, KU, nullptr, nullptr, s, "spmv_bcsr", true); break;
    case 2: spmv_launch<2, 2, 0, 2, false>(L, K, U, KU, nullptr, nullptr, s, "spmv_bcsr", true); break;
    case 3: spmv_launch<3, 3, 0, 2, false>(L, K, U, KU, nullptr, nullptr, s, "spmv_bcsr", true); break;
    case 4: spmv<4, 4, 0>(L, K, U, KU, nullptr, nullptr, s); break;
    default: return set_error(SVG_ERR_UNSUPPORTED, "svg_spar_mul_vv: dof must be 1..4");
  }
  SVG_HIP(hipGetLastError());
  return SVG_OK;
}

int svg_spmv_num_variants(int* count)
{
  if (!count) return set_error(SVG_ERR_ARG, "svg_spmv_num_variants: null argument");
  *count = kNumSpmvVariants;
  return SVG_OK;
}

int svg_spmv_time_variant(const svg_lhs* L, int variant, const double* K, const double* U, double* KU, int launches,
                          double* avg_us, void* stream)
{
  if (!L || !K || !U || !KU || !avg_us || launches < 1 || variant < 0 || variant >= kNumSpmvVariants)
    return set_error(SVG_ERR_ARG, "svg_spmv_time_variant: bad argument");
  hipStream_t s = (hipStream_t)stream;
  hipEvent_t a, b;
  SVG_HIP(hipEventCreate(&a));
  SVG_HIP(hipEventCreate(&b));
  spmv_variant<4, 4, 0>(variant, L, K, U, KU, nullptr, nullptr, s, "spmv_tune");
  SVG_HIP(hipEventRecord(a, s));
  for (int r = 0; r < launches; r++) spmv_variant<4, 4, 0>(variant, L, K, U, KU, nullptr, nullptr, s, "spmv_tune");
  SVG_HIP(hipEventRecord(b, s));
  SVG_HIP(hipEventSynchronize(b));
  float ms = 0.f;
  SVG_HIP(hipEventElapsedTime(&ms, a, b));
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  SVG_HIP(hipGetLastError());
  *avg_us = 1e3 * ms / launches;
  return SVG_OK;
}

int svg_spmv_tune(svg_lhs* L, const double* K, int* best, void* stream)
{
  if (!L || !K) return set_error(SVG_ERR_ARG, "svg_spmv_tune: null argument");
  int rc = ensure_work(L->kw, 4 * L->nNo, L->kw.sD > 0 ? L->kw.sD : 1);
  if (rc) return rc;
  // the Krylov basis holds the probe vectors; results are discarded
  double* U = L->kw.u;
  double* KU = L->kw.X;
  SVG_HIP(hipMemsetAsync(U, 0, sizeof(double) * 4 * L->nNo, (hipStream_t)stream));
  double t[kNumSpmvVariants];
  bool banded[kNumSpmvVariants];
  for (int v = 0; v < kNumSpmvVariants; v++) {
    if ((rc = svg_spmv_time_variant(L, v, K, U, KU, 3, &t[v], stream))) return rc;
    banded[v] = kSpmvVariants[v].xcd && !kSpmvVariants[v].nt;
  }
  const int bv = pick_variant(t, banded, kNumSpmvVariants);
  L->spmvVariant = bv;
  if (best) *best = bv;
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
    if (rc) return rc;
  }
  FinArgs fa{};
  fa.mode = FIN_DOT;
  dot_device(n, V, U, 0, 1, g_dot.partial, g_dot.out, nullptr, fa, nullptr, nullptr, nullptr, nullptr, nullptr, s);
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

int svg_dot_v_noalloc(int dof, int nNo, const double* U, const double* V, double* out, void* stream)
{
  if (!U || !V || !out || dof <= 0 || nNo < 0) return set_error(SVG_ERR_ARG, "svg_dot_v_noalloc: bad argument");
  const long n = (long)dof * (long)nNo;
  if (n > kNoAllocMaxN)
    return set_error(SVG_ERR_ARG, "svg_dot_v_noalloc: dof*nNo exceeds 2^28 entries, the most its static scratch holds");
  double* sp = nullptr;
  int rc = noalloc_scratch(&sp);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  FinArgs fa{};
  fa.mode = FIN_DOT;
  // the kernels and their order are svg_dot_v's: the result is bit-identical
  double* res = sp + kNoAllocPartials;
  if ((rc = dot_device((int)n, V, U, 0, 1, sp, res, nullptr, fa, nullptr, nullptr, nullptr, nullptr, nullptr, s)))
    return rc;
  SVG_HIP(hipGetLastError());
  SVG_HIP(hipMemcpyAsync(out, res, sizeof(double), hipMemcpyDeviceToHost, s));
  SVG_HIP(hipStreamSynchronize(s));
  return SVG_OK;
}

int svg_fsils_solve(svg_lhs* L, svg_ls* ls, int dof, double* Ri, double* Val, int prec, int nIncL, const int* incL,
                    const double* res, void* stream)
{
  return svg_fsils_solve_ex(L, ls, dof, Ri, Val, prec, nIncL, incL, res, nullptr, stream);
}

int svg_fsils_solve_ex(svg_lhs* L, svg_ls* ls, int dof, double* Ri, double* Val, int prec, int nIncL,
                       const int* incL, const double* res, svg_solve_opts* opts, void* stream)
{
  if (opts && opts->precision == SVG_PRECISION_MIXED && ls && ls->LS_type != SVG_LS_GMRES)
    return set_error(SVG_ERR_UNSUPPORTED, "fsils_solve: mixed precision is built for LS type GMRES");
  const bool learned = opts && opts->method == SVG_METHOD_LEARNED;
  if (opts && opts->method != SVG_METHOD_PLAIN && !learned)
    return set_error(SVG_ERR_ARG, "fsils_solve: unknown method");
  if (learned && !opts->lpc && opts->coarse <= 0)
    return set_error(SVG_ERR_UNSUPPORTED, "fsils_solve: the learned preconditioner needs a network or a coarse level");
  // the flexible GMRES: the network, the two-level preconditioner, or both
  const bool flex = learned || (opts && opts->coarse > 0);
  if (flex && ls && ls->LS_type != SVG_LS_GMRES)
    return set_error(SVG_ERR_UNSUPPORTED,
                     "fsils_solve: the learned and two-level preconditioners are built for LS type GMRES");
  if (opts) opts->fallback_itr = -1;
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
        F.res = res[f];
        F.coupledFlag = true;
        if (ls->LS_type != SVG_LS_GMRES && ls->LS_type != SVG_LS_NS)
          return set_error(SVG_ERR_UNSUPPORTED,
                           "fsils_solve: coupled (resistance) Neumann faces are built for LS types GMRES and NS");
      }
    }
  }
  if (dof != 4) return set_error(SVG_ERR_UNSUPPORTED, "fsils_solve: only dof = 4 is built");
  int rc;
  int sDneed = ls->LS_type == SVG_LS_GMRES ? ls->RI.sD : 1;
  if ((rc = ensure_work(L->kw, dof * L->nNo, sDneed < 1 ? 1 : sDneed))) return rc;
  auto& k = L->kw;
  const int n = dof * L->nNo;
  if (prec == SVG_PREC_FSILS) {
    if ((rc = precond_diag_dev(L, dof, Val, Ri, k.W, s))) return rc;
    for (const auto& F : L->face) {
      if (!F.coupledFlag) continue;
      const int nsd = F.dof < dof ? F.dof : dof;
      launch("face_valM", k_face_valM, dim3(blocks_for((size_t)F.nNo * nsd, 256)), dim3(256), 0, s, F.nNo, F.dof,
             nsd, dof, (const int*)F.d_glob, (const double*)F.d_val, (const double*)k.W, F.d_valM);
    }
  } else if (prec == SVG_PREC_NONE) {
    svg_fill(k.W, (size_t)n, 1.0, s);
  } else {
    return set_error(SVG_ERR_UNSUPPORTED, "fsils_solve: preconditioner not built");
  }
  switch (ls->LS_type) {
    case SVG_LS_GMRES:
      if (flex)
        rc = gmres_v_learned(L, &ls->RI, Val, Ri, opts->precision == SVG_PRECISION_MIXED, learned ? opts->lpc : nullptr,
                             opts->fallback_window > 0 ? opts->fallback_window : 20, opts->net_period,
                             opts->coarse, opts->coarse_keep != 0, &opts->fallback_itr, s);
      else
        rc = (opts && opts->precision == SVG_PRECISION_MIXED) ? gmres_v_mixed(L, &ls->RI, Val, Ri, s)
                                                               : gmres_v_dev(L, &ls->RI, dof, Val, Ri, s);
      break;
    case SVG_LS_NS:
      rc = ns_solver_dev(L, ls, Val, Ri, s);
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


int svg_batch_create(svg_lhs* L, int nsys, const int* nodeOff, int sD, svg_batch** out)
{
  if (!L || !nodeOff || !out || nsys < 1) return set_error(SVG_ERR_ARG, "svg_batch_create: bad argument");
  if (sD < 1 || sD > 250) return set_error(SVG_ERR_ARG, "svg_batch_create: Krylov dimension must be in [1, 250]");
  if (nodeOff[0] != 0 || nodeOff[nsys] != L->nNo)
    return set_error(SVG_ERR_ARG, "svg_batch_create: nodeOff must run from 0 to the lhs's nNo");
  for (int q = 0; q < nsys; q++)
    if (nodeOff[q + 1] <= nodeOff[q]) return set_error(SVG_ERR_ARG, "svg_batch_create: empty or unordered system");
  auto* B = new svg_batch();
  B->L = L;
  B->nsys = nsys;
  B->sD = sD;
  B->n = 4 * L->nNo;
  B->ld = ((size_t)B->n + 63) & ~(size_t)63;
  B->nodeOff.assign(nodeOff, nodeOff + nsys + 1);
  std::vector<int> sysOfNode(L->nNo), spanStart, spanEnd, spanSys, sysSpanOff(nsys + 1, 0);
  for (int q = 0; q < nsys; q++) {
    for (int a = nodeOff[q]; a < nodeOff[q + 1]; a++) sysOfNode[a] = q;
    const int r0 = 4 * nodeOff[q], r1 = 4 * nodeOff[q + 1];
    sysSpanOff[q] = (int)spanStart.size();
    for (int k = r0; k < r1; k += RED_SPAN) {
      spanStart.push_back(k);
      spanEnd.push_back(std::min(r1, k + RED_SPAN));
      spanSys.push_back(q);
    }
  }
  sysSpanOff[nsys] = (int)spanStart.size();
  B->nsp = (int)spanStart.size();
  int rc;
  if ((rc = dev_alloc(&B->d_sysOfNode, (size_t)L->nNo)) || (rc = dev_alloc(&B->d_spanStart, (size_t)B->nsp)) ||
      (rc = dev_alloc(&B->d_spanEnd, (size_t)B->nsp)) || (rc = dev_alloc(&B->d_spanSys, (size_t)B->nsp)) ||
      (rc = dev_alloc(&B->d_sysSpanOff, (size_t)nsys + 1)) || (rc = dev_alloc(&B->u, B->ld * (sD + 1))) ||
      (rc = dev_alloc(&B->X, (size_t)B->n)) || (rc = dev_alloc(&B->h, (size_t)nsys * (sD + 1) * sD)) ||
      (rc = dev_alloc(&B->cs, (size_t)nsys * sD)) || (rc = dev_alloc(&B->sn, (size_t)nsys * sD)) ||
      (rc = dev_alloc(&B->err, (size_t)nsys * (sD + 1))) || (rc = dev_alloc(&B->coef, (size_t)nsys * (sD + 2))) ||
      (rc = dev_alloc(&B->y, (size_t)nsys * sD)) || (rc = dev_alloc(&B->partial, (size_t)B->nsp * (sD + 2))) ||
      (rc = dev_alloc(&B->st, (size_t)nsys))) {
    svg_batch_free(B);
    return rc;
  }
  hipError_t e = hipMemcpy(B->d_sysOfNode, sysOfNode.data(), sizeof(int) * L->nNo, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(B->d_spanStart, spanStart.data(), sizeof(int) * B->nsp, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(B->d_spanEnd, spanEnd.data(), sizeof(int) * B->nsp, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(B->d_spanSys, spanSys.data(), sizeof(int) * B->nsp, hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(B->d_sysSpanOff, sysSpanOff.data(), sizeof(int) * (nsys + 1), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(B->h, 0, sizeof(double) * nsys * (sD + 1) * sD);
  if (e == hipSuccess) e = hipMemset(B->st, 0, sizeof(GmState) * nsys);
  if (e == hipSuccess) e = hipHostMalloc((void**)&B->h_st, sizeof(GmState) * nsys, hipHostMallocDefault);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&B->cs_stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&B->ev, hipEventDisableTiming);
  if (e != hipSuccess) {
    svg_batch_free(B);
    return hip_check(e, "svg_batch_create");
  }
  *out = B;
  return SVG_OK;
}

int svg_batch_free(svg_batch* B)
{
  if (!B) return SVG_OK;
  batch_graph_free(B);
  for (int* p : {B->d_sysOfNode, B->d_spanStart, B->d_spanEnd, B->d_spanSys, B->d_sysSpanOff})
    if (p) (void)hipFree(p);
  for (double* p : {B->u, B->X, B->h, B->cs, B->sn, B->err, B->coef, B->y, B->partial})
    if (p) (void)hipFree(p);
  if (B->st) (void)hipFree(B->st);
  if (B->u32) (void)hipFree(B->u32);
  if (B->v32) (void)hipFree(B->v32);
  if (B->z) (void)hipFree(B->z);
  if (B->z32) (void)hipFree(B->z32);
  if (B->d_nsc) (void)hipFree(B->d_nsc);
  if (B->h_st) (void)hipHostFree(B->h_st);
  if (B->ev) (void)hipEventDestroy(B->ev);
  if (B->cs_stream) (void)hipStreamDestroy(B->cs_stream);
  delete B;
  return SVG_OK;
}

int svg_batch_gmres_ex(svg_batch* B, const svg_subls* proto, double* R, const double* Val,
                       const svg_solve_opts* opts, svg_subls* out, void* stream)
{
  if (!B || !proto || !R || !Val || !out) return set_error(SVG_ERR_ARG, "svg_batch_gmres: null argument");
  if (proto->mItr < 1) return set_error(SVG_ERR_ARG, "svg_batch_gmres: mItr must be >= 1");
  BatchMode md;
  int rc;
  if ((rc = batch_mode(opts, &md))) return rc;
  hipStream_t caller = (hipStream_t)stream, s = B->cs_stream;
  auto t0 = std::chrono::steady_clock::now();
  if ((rc = batch_prepare(B, Val, md, caller))) return rc;
  SVG_HIP(hipEventRecord(B->ev, caller));
  SVG_HIP(hipStreamWaitEvent(s, B->ev, 0));
  const FinArgs fa = batch_fin_args(proto, B->sD, md);
  if ((rc = batch_begin(B, R, Val, md, fa, s)) || (rc = batch_cycle_graph(B, R, Val, md, fa, proto->mItr)))
    return rc;
  std::vector<double> dB;
  if ((rc = batch_iterate(B, md, proto->mItr, dB))) return rc;
  if (md.coarse && (rc = coarse_check(B->L, s))) return rc;
  hipLaunchKernelGGL(k_bresult, dim3(blocks_for(B->n, 256)), dim3(256), 0, s, B->n, R, (const double*)B->X,
                     (const GmState*)B->st, (const int*)B->d_sysOfNode, proto->absTol);
  SVG_HIP(hipGetLastError());
  SVG_HIP(hipEventRecord(B->ev, s));
  SVG_HIP(hipStreamWaitEvent(caller, B->ev, 0));
  SVG_HIP(hipStreamSynchronize(s));
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  batch_records(B, proto, md, dB, secs, out);
  return SVG_OK;
}

int svg_batch_gmres(svg_batch* B, const svg_subls* proto, double* R, const double* Val, int precision,
                    svg_subls* out, void* stream)
{
  svg_solve_opts o{};
  o.precision = precision;
  return svg_batch_gmres_ex(B, proto, R, Val, &o, out, stream);
}

int svg_batch_fsils_solve_ex(svg_batch* B, const svg_subls* proto, double* R, double* Val,
                             const svg_solve_opts* opts, svg_subls* out, void* stream)
{
  if (!B || !R || !Val) return set_error(SVG_ERR_ARG, "svg_batch_fsils_solve: null argument");
  svg_lhs* L = B->L;
  hipStream_t s = (hipStream_t)stream;
  for (auto& f : L->face) {
    f.incFlag = true;
    f.coupledFlag = false;  // no resistances in the batched solve
  }
  int rc;
  if ((rc = ensure_work(L->kw, 4 * L->nNo, 1))) return rc;
  if ((rc = precond_diag_dev(L, 4, Val, R, L->kw.W, s))) return rc;
  if ((rc = svg_batch_gmres_ex(B, proto, R, Val, opts, out, stream))) return rc;
  launch("wc_scale", k_mul, dim3(blocks_for(B->n, 256)), dim3(256), 0, s, B->n, (const double*)L->kw.W,
         (const double*)R, R);
  SVG_HIP(hipGetLastError());
  SVG_HIP(hipStreamSynchronize(s));
  return SVG_OK;
}

int svg_batch_fsils_solve(svg_batch* B, const svg_subls* proto, double* R, double* Val, int precision,
                          svg_subls* out, void* stream)
{
  svg_solve_opts o{};
  o.precision = precision;
  return svg_batch_fsils_solve_ex(B, proto, R, Val, &o, out, stream);
}

int svg_batch_fallback_itr(const svg_batch* B, int* itr)
{
  if (!B || !itr) return set_error(SVG_ERR_ARG, "svg_batch_fallback_itr: null argument");
  for (int q = 0; q < B->nsys; q++) itr[q] = q < (int)B->fbItr.size() ? B->fbItr[q] : -1;
  return SVG_OK;
}

int svg_batch_replays(const svg_batch* B, long long* replays)
{
  if (!B || !replays) return set_error(SVG_ERR_ARG, "svg_batch_replays: null argument");
  *replays = B->replays;
  return SVG_OK;
}

int svg_batch_newton_iteration(svg_batch* B, svg_msh* M, svg_face* Fa, const svg_fluid_params* p, double beta,
                               double qmTRI3, const double* Ao, double* An, const double* Yo, double* Yn, double* Ag,
                               double* Yg, double* R, double* Val, const svg_subls* proto, const svg_solve_opts* opts,
                               svg_subls* out, void* stream)
{
  if (!B || !M || !p || !Ao || !An || !Yo || !Yn || !Ag || !Yg || !R || !Val || !proto || !out)
    return set_error(SVG_ERR_ARG, "svg_batch_newton_iteration: null argument");
  if (M->lhs != B->L || (Fa && Fa->msh != M))
    return set_error(SVG_ERR_ARG, "svg_batch_newton_iteration: the mesh, face and batch must share one lhs");
  if (proto->mItr < 1) return set_error(SVG_ERR_ARG, "svg_batch_newton_iteration: mItr must be >= 1");
  BatchMode md;
  int rc;
  if ((rc = batch_mode(opts, &md))) return rc;
  svg_lhs* L = B->L;
  hipStream_t caller = (hipStream_t)stream, s = B->cs_stream;
  auto t0 = std::chrono::steady_clock::now();
  for (auto& f : L->face) {
    f.incFlag = true;
    f.coupledFlag = false;  // no resistances in the batched solve
  }
  if ((rc = ensure_work(L->kw, 4 * L->nNo, 1)) || (rc = build_mask(L, 4)) ||
      (rc = batch_prepare(B, Val, md, caller)))
    return rc;
  const FinArgs fa = batch_fin_args(proto, B->sD, md);
  double* W = L->kw.W;
  GraphKey k;
  k.add(M).add(Fa).add(*p).add(beta).add(qmTRI3).add(Ao).add(An).add(Yo).add(Yn).add(Ag).add(Yg).add(R).add(Val);
  k.add(W).add(md.mixed).add(md.net).add(fa.absTol).add(fa.relTol).add(L->d_val32).add(L->d_val16);
  k.add(md.coarse).add(coarse_generation(L));
  if (B->pro && B->newtonKey != k.b) {
    for (hipGraphExec_t* g : {&B->pro, &B->epi})
      if (*g) (void)hipGraphExecDestroy(*g), *g = nullptr;
  }
  if (!B->pro) {
    // pici, ls_alloc's R = Val = 0, construct_fluid, b_assem_neu_bc,
    // precond_diag and gmres_v up to its restart loop
    rc = capture(s, &B->pro, [&]() -> int {
      int r;
      if ((r = svg_pici(L->nNo, p->am, p->af, Ao, An, Yo, Yn, Ag, Yg, s)) ||
          (r = svg_fill(R, (size_t)4 * L->nNo, 0.0, s)) || (r = svg_fill(Val, (size_t)16 * L->nnz, 0.0, s)) ||
          (r = construct_fluid_enqueue(M, p, Ag, Yg, nullptr, R, Val, s)) ||
          (Fa && (r = svg_b_assem_neu_bc(Fa, p, beta, qmTRI3, nullptr, Yg, R, Val, s))) ||
          (r = precond_diag_dev(L, 4, Val, R, W, s)))
        return r;
      return batch_begin(B, R, Val, md, fa, s);
    });
    // the solutions, R = W R (solve.cpp:184-186) and the picc update
    if (!rc)
      rc = capture(s, &B->epi, [&]() -> int {
        hipLaunchKernelGGL(k_bresult, dim3(blocks_for(B->n, 256)), dim3(256), 0, s, B->n, R, (const double*)B->X,
                           (const GmState*)B->st, (const int*)B->d_sysOfNode, fa.absTol);
        launch("wc_scale", k_mul, dim3(blocks_for(B->n, 256)), dim3(256), 0, s, B->n, (const double*)W,
               (const double*)R, R);
        SVG_HIP(hipGetLastError());
        return svg_picc_update(L->nNo, p->gam, p->dt, R, An, Yn, s);
      });
    if (rc) {
      batch_graph_free(B);
      return rc;
    }
    B->newtonKey = k.b;
  }
  if ((rc = batch_cycle_graph(B, R, Val, md, fa, proto->mItr))) return rc;
  SVG_HIP(hipEventRecord(B->ev, caller));
  SVG_HIP(hipStreamWaitEvent(s, B->ev, 0));
  SVG_HIP(hipGraphLaunch(B->pro, s));
  std::vector<double> dB;
  if ((rc = batch_iterate(B, md, proto->mItr, dB))) return rc;
  SVG_HIP(hipGraphLaunch(B->epi, s));
  SVG_HIP(hipEventRecord(B->ev, s));
  SVG_HIP(hipStreamWaitEvent(caller, B->ev, 0));
  if ((rc = construct_fluid_check(M, s))) return rc;
  if (md.coarse && (rc = coarse_check(L, s))) return rc;
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  batch_records(B, proto, md, dB, secs, out);
  return SVG_OK;
}


int svg_spmv32(const svg_lhs* L, const float* K32, const float* U32, double* KU, void* stream)
{
  if (!L || !K32 || !U32 || !KU) return set_error(SVG_ERR_ARG, "svg_spmv32: null argument");
  spmv32(L, K32, U32, KU, nullptr, (hipStream_t)stream);
  SVG_HIP(hipGetLastError());
  return SVG_OK;
}

int svg_spmv32_tune(svg_lhs* L, const float* K32, float* U32, double* KU, int* best_variant, void* stream)
{
  if (!L || !K32 || !U32 || !KU) return set_error(SVG_ERR_ARG, "svg_spmv32_tune: null argument");
  int rc = spmv32_tune(L, K32, U32, KU, (hipStream_t)stream);
  if (rc) return rc;
  if (best_variant) *best_variant = L->spmv32Variant;
  return SVG_OK;
}

int svg_spmv_variants(const svg_lhs* L, int* fp64_variant, int* fp32_variant)
{
  if (!L) return set_error(SVG_ERR_ARG, "svg_spmv_variants: null lhs");
  if (fp64_variant) *fp64_variant = L->spmvVariant;
  if (fp32_variant) *fp32_variant = L->spmv32Variant;
  return SVG_OK;
}

}  // extern "C"
