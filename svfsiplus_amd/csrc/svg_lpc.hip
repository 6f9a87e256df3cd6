// Learned preconditioner: a message-passing network over mesh nodes that maps
// a residual (and the local diagonal block of the scaled system matrix) to a
// correction, applied inside flexible GMRES (no counterpart in the reference;
// the preconditioner it stands beside is precond_diag, precond.cpp:122-257).
//
// Network (hidden width 64 = 16 channels x 4 dofs, feature index 4c + d):
//   x_i   = v_i * sqrt(nNo)                              (v has unit norm in FGMRES)
//   u_i   = [x_i (4), D_i (16), 1, 0 x 11]                D_i = diagonal block of A
//   h0_i  = W0 u_i                                        W0: 64 x 32
//   for l = 1..3:
//     m_i  = sum_{j in row i} A_ij (x) h_j                m[c][d] = sum_e A_ij[d][e] h_j[c][e]
//     p_i  = W_l [h_i ; m_i] + b_l                        W_l: 64 x 128
//     h_i <- h_i + relu(p_i)
//   z_i   = v_i + (Wd h_i) / sqrt(nNo)                    Wd: 4 x 64
// so with Wd = 0 the network is the identity on the diagonally scaled system
// (= the diagonal preconditioner).
// Precision: weights, activations, messages and the matrix copy in BF16
// (round-to-nearest-even), every product accumulated in FP32; the dense layers
// run on MFMA (v_mfma_f32_16x16x16_bf16) from LDS tiles of 64 nodes.
// Training minimises L = (1/B) sum_b |r_b - A z_b|^2 with z_b = M(r_b), by
// hand-written backward kernels (the same MFMA GEMMs transposed, the transposed
// message passing via the block-transpose map) and Adam on FP32 master weights.
#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <vector>

#include "svg_device.h"

using svg::blocks_for;
using svg::dev_alloc;
using svg::dev_free;
using svg::launch;
using svg::launch_b;
using svg::set_error;

namespace svg {
namespace lpc {

typedef unsigned short bf16;
typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int F = 64;            // hidden width
constexpr int NIN = 32;          // encoder inputs (21 used)
constexpr int NL = 3;            // message-passing layers
constexpr int TILE = 64;         // nodes per workgroup tile
constexpr int BS = 256;          // threads per workgroup (4 wavefronts)
constexpr int XS = 2 * F + 8;    // LDS row stride of [h | m] and of W_l (bf16)
constexpr int US = NIN + 8;      // LDS row stride of u and W0
constexpr int GS = F + 8;        // LDS row stride of the BF16 gradient tile
constexpr int OFF_W0 = 0;        // parameter offsets (floats)
constexpr int OFF_L = F * NIN;
constexpr int LSZ = F * 2 * F + F;  // W_l then b_l
constexpr int OFF_WD = OFF_L + NL * LSZ;
constexpr int NPARAM = OFF_WD + 4 * F;
constexpr int MAXGRID = 512;     // workgroups of the weight-gradient kernels (partials)

__device__ __forceinline__ float bf2f(bf16 h) { return __uint_as_float(((unsigned)h) << 16); }
__device__ __forceinline__ bf16 f2bf(float f)
{
  unsigned u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (bf16)(u >> 16);
}
__device__ __forceinline__ float lo(unsigned u) { return __uint_as_float(u << 16); }
__device__ __forceinline__ float hi(unsigned u) { return __uint_as_float(u & 0xffff0000u); }

// C(M x N) += A(M x K) B(K x N) on MFMA 16x16x16 BF16 tiles, operands in LDS:
// A(m,k) = A[m*sam + k*sak], B(k,n) = B[k*sbk + n*sbn]. Wavefront w owns the
// output tiles q = w + 4*qi of the (M/16) x (N/16) grid (row-major); acc[qi][v]
// holds C(16*tm + 4*(lane/16) + v, 16*tn + lane%16).
template <int M, int N, int K>
__device__ __forceinline__ void mma(const bf16* A, int sam, int sak, const bf16* B, int sbk, int sbn, v4f* acc,
                                    int wave, int lane)
{
  constexpr int TN = N / 16, QW = (M / 16) * (N / 16) / 4;
  const int r = lane & 15, g = lane >> 4;
#pragma unroll
  for (int qi = 0; qi < QW; qi++) {
    const int q = wave + 4 * qi, tm = q / TN, tn = q % TN;
    v4f c = acc[qi];
#pragma unroll
    for (int k0 = 0; k0 < K; k0 += 16) {
      v4s a, b;
#pragma unroll
      for (int v = 0; v < 4; v++) {
        a[v] = (short)A[(16 * tm + r) * sam + (k0 + 4 * g + v) * sak];
        b[v] = (short)B[(k0 + 4 * g + v) * sbk + (16 * tn + r) * sbn];
      }
      c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
    }
    acc[qi] = c;
  }
}

// The same product for operands whose k index is contiguous (A(m,k) =
// A[m*sam + k], B(k,n) = B[n*sbn + k]): each lane's 4 consecutive BF16 of A
// and of B come in one 8-byte LDS read (rows 8-byte aligned).
template <int M, int N, int K>
__device__ __forceinline__ void mma_kc(const bf16* A, int sam, const bf16* B, int sbn, v4f* acc, int wave, int lane)
{
  constexpr int TN = N / 16, QW = (M / 16) * (N / 16) / 4;
  const int r = lane & 15, g = lane >> 4;
#pragma unroll
  for (int qi = 0; qi < QW; qi++) {
    const int q = wave + 4 * qi, tm = q / TN, tn = q % TN;
    v4f c = acc[qi];
    const bf16* Ar = A + (16 * tm + r) * sam + 4 * g;
    const bf16* Br = B + (16 * tn + r) * sbn + 4 * g;
#pragma unroll
    for (int k0 = 0; k0 < K; k0 += 16) {
      const v4s a = *reinterpret_cast<const v4s*>(Ar + k0);
      const v4s b = *reinterpret_cast<const v4s*>(Br + k0);
      c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
    }
    acc[qi] = c;
  }
}

// Row and column of accumulator element (qi, v) of wavefront `wave`.
template <int N>
__device__ __forceinline__ void acc_rc(int wave, int lane, int qi, int v, int& row, int& col)
{
  constexpr int TN = N / 16;
  const int q = wave + 4 * qi;
  row = 16 * (q / TN) + 4 * (lane >> 4) + v;
  col = 16 * (q % TN) + (lane & 15);
}

// Encoder input u of node row r (sample-major rows r = b*n + i) into LDS.
__device__ __forceinline__ void load_u(bf16* U, int t, int r0, int rows, int n, const double* __restrict__ V,
                                       double sc, const double2* __restrict__ nsc, const bf16* __restrict__ Dg)
{
  for (int q = t; q < TILE * NIN; q += BS) {
    const int row = q / NIN, c = q % NIN, r = r0 + row;
    bf16 v = 0;
    if (r < rows) {
      const int i = r % n;
      if (c < 4)
        v = f2bf((float)(V[(size_t)r * 4 + c] * (nsc ? nsc[i].x : sc)));
      else if (c < 20)
        v = Dg[(size_t)i * 16 + (c - 4)];
      else if (c == 20)
        v = (bf16)0x3f80;  // 1.0
    }
    U[row * US + c] = v;
  }
}

// ---------------------------------------------------------------- setup ---
// BF16 copy of the (scaled) block matrix and of its diagonal blocks.
__global__ __launch_bounds__(256) void k_val_bf16(size_t n, const double* __restrict__ Val, bf16* __restrict__ A16)
{
  size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; k < n; k += stride) A16[k] = f2bf((float)Val[k]);
}

__global__ __launch_bounds__(256) void k_diag_bf16(int nNo, const int* __restrict__ diagPtr,
                                                   const double* __restrict__ Val, bf16* __restrict__ Dg)
{
  const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (size_t)nNo * 16) return;
  const int a = (int)(k >> 4), q = (int)(k & 15);
  Dg[k] = f2bf((float)Val[16 * (size_t)diagPtr[a] + q]);
}

// -------------------------------------------------------------- forward ---
// Batched solves: rows of systems that stopped iterating (converged before the
// cycle, or done in it) are skipped; sysOf null: every row is live.
struct RowGate {
  const int* sysOf;
  const GmState* st;
  __device__ __forceinline__ bool live(int node) const
  {
    const GmState& S = st[sysOf[node]];
    return S.act && !S.done;
  }
};

// Whether a row of the 64-row tile at r0 is live (the same answer in every
// thread of the workgroup).
__device__ __forceinline__ bool tile_live(const RowGate& g, int r0, int rows, int n, int t)
{
  __shared__ int any;
  if (t == 0) any = 0;
  __syncthreads();
  if (t < TILE && r0 + t < rows && g.live((r0 + t) % n)) any = 1;
  __syncthreads();
  const bool lv = any != 0;
  __syncthreads();
  return lv;
}

__global__ __launch_bounds__(BS) void k_encode(int n, int rows, const double* __restrict__ V, double sc,
                                               const double2* __restrict__ nsc, const bf16* __restrict__ Dg,
                                               const bf16* __restrict__ W0, bf16* __restrict__ H,
                                               const int* __restrict__ skip, RowGate gate)
{
  constexpr int HS = F + 8;  // LDS row stride of the output tile (bf16; 16-byte aligned rows)
  __shared__ __align__(16) bf16 U[TILE * US];
  __shared__ __align__(16) bf16 Ws[F * US];
  __shared__ __align__(16) bf16 Ho[TILE * HS];
  if (skip && *skip) return;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int ntiles = (rows + TILE - 1) / TILE;
  for (int q = t; q < F * NIN; q += BS) Ws[(q / NIN) * US + q % NIN] = W0[q];  // once per workgroup
  for (int tl = blockIdx.x; tl < ntiles; tl += gridDim.x) {
    const int r0 = tl * TILE;
    if (gate.sysOf && !tile_live(gate, r0, rows, n, t)) continue;
    load_u(U, t, r0, rows, n, V, sc, nsc, Dg);
    __syncthreads();
    v4f acc[4];
#pragma unroll
    for (int qi = 0; qi < 4; qi++) acc[qi] = v4f{0.f, 0.f, 0.f, 0.f};
    mma_kc<TILE, F, NIN>(U, US, Ws, US, acc, wave, lane);  // B(k,n) = W0[n][k]
#pragma unroll
    for (int qi = 0; qi < 4; qi++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        int row, col;
        acc_rc<F>(wave, lane, qi, v, row, col);
        Ho[row * HS + col] = f2bf(acc[qi][v]);
      }
    __syncthreads();
    // the tile's h rows leave LDS as 16-byte stores
    for (int q = t; q < TILE * F / 8; q += BS) {
      const int row = q / (F / 8), c8 = q % (F / 8);
      if (r0 + row < rows)
        reinterpret_cast<uint4*>(H + (size_t)(r0 + row) * F)[c8] = reinterpret_cast<const uint4*>(Ho + row * HS)[c8];
    }
    __syncthreads();
  }
}

typedef __bf16 v2bf __attribute__((ext_vector_type(2)));

// Lane k of each group of four lanes (quad) broadcast to the quad (DPP quad_perm).
template <int K>
__device__ __forceinline__ unsigned quad_bcast(unsigned v)
{
  return (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, K * 0x55, 0xF, 0xF, false);
}

// m[c][d] += sum_q A[d][q] h[c][q] for one block: the block's row d is the
// packed BF16 pairs (a[2d], a[2d+1]), channel c of h the pairs (h[2c], h[2c+1]);
// two packed dot products (FP32 accumulation) per output, q = 0,1 then 2,3.
// The four lanes of a node row (a quad) each load one block row, aq = row
// (lane & 3), and exchange them by DPP, so the block is read once per row.
__device__ __forceinline__ void msg_block(float (&m)[4][4], const uint2& aq, const uint4& h0, const uint4& h1)
{
  const unsigned A[8] = {quad_bcast<0>(aq.x), quad_bcast<0>(aq.y), quad_bcast<1>(aq.x), quad_bcast<1>(aq.y),
                         quad_bcast<2>(aq.x), quad_bcast<2>(aq.y), quad_bcast<3>(aq.x), quad_bcast<3>(aq.y)};
  const unsigned H[8] = {h0.x, h0.y, h0.z, h0.w, h1.x, h1.y, h1.z, h1.w};
#pragma unroll
  for (int c = 0; c < 4; c++)
#pragma unroll
    for (int d = 0; d < 4; d++) {
      float t = __builtin_amdgcn_fdot2_f32_bf16(__builtin_bit_cast(v2bf, A[2 * d]),
                                               __builtin_bit_cast(v2bf, H[2 * c]), m[c][d], false);
      m[c][d] = __builtin_amdgcn_fdot2_f32_bf16(__builtin_bit_cast(v2bf, A[2 * d + 1]),
                                               __builtin_bit_cast(v2bf, H[2 * c + 1]), t, false);
    }
}

// The messages of node i for this lane's 4 channels (16 outputs): m = sum over
// the row's blocks k, in CSR order, of A_k (x) h[colPtr[k]], the next UNR
// blocks' column indices loaded while the current ones are consumed. The four
// lanes of the row (g = lane & 3) run this together: they must form one quad
// and all take part (DPP exchange of the block rows).
template <int UNR>
__device__ __forceinline__ void gather_row(float (&m)[4][4], int s, int e, int g, const int* __restrict__ colPtr,
                                           const bf16* __restrict__ A16, const bf16* __restrict__ Hb)
{
#pragma unroll
  for (int c = 0; c < 4; c++)
#pragma unroll
    for (int d = 0; d < 4; d++) m[c][d] = 0.f;
  int k = s;
  int jn[UNR];
#pragma unroll
  for (int q = 0; q < UNR; q++) jn[q] = k + q < e ? colPtr[k + q] : 0;
  for (; k + UNR <= e; k += UNR) {
    int j[UNR];
#pragma unroll
    for (int q = 0; q < UNR; q++) j[q] = jn[q];
    uint2 aq[UNR];
    uint4 h0[UNR], h1[UNR];
#pragma unroll
    for (int q = 0; q < UNR; q++) {
      aq[q] = reinterpret_cast<const uint2*>(A16 + 16 * (size_t)(k + q))[g];
      h0[q] = reinterpret_cast<const uint4*>(Hb + (size_t)j[q] * F)[0];
      h1[q] = reinterpret_cast<const uint4*>(Hb + (size_t)j[q] * F)[1];
    }
#pragma unroll
    for (int q = 0; q < UNR; q++) jn[q] = k + UNR + q < e ? colPtr[k + UNR + q] : 0;
#pragma unroll
    for (int q = 0; q < UNR; q++) msg_block(m, aq[q], h0[q], h1[q]);
  }
  for (int q = 0; k < e; k++, q++) {
    const int jj = jn[q];
    const uint2 aq = reinterpret_cast<const uint2*>(A16 + 16 * (size_t)k)[g];
    const uint4 h0 = reinterpret_cast<const uint4*>(Hb + (size_t)jj * F)[0];
    const uint4 h1 = reinterpret_cast<const uint4*>(Hb + (size_t)jj * F)[1];
    msg_block(m, aq, h0, h1);
  }
}

// 16 messages rounded to BF16, packed as two 16-byte words (channel c, dof d at 4c + d).
__device__ __forceinline__ void pack_msg(const float (&m)[4][4], uint4& o0, uint4& o1)
{
  o0.x = (unsigned)f2bf(m[0][0]) | ((unsigned)f2bf(m[0][1]) << 16);
  o0.y = (unsigned)f2bf(m[0][2]) | ((unsigned)f2bf(m[0][3]) << 16);
  o0.z = (unsigned)f2bf(m[1][0]) | ((unsigned)f2bf(m[1][1]) << 16);
  o0.w = (unsigned)f2bf(m[1][2]) | ((unsigned)f2bf(m[1][3]) << 16);
  o1.x = (unsigned)f2bf(m[2][0]) | ((unsigned)f2bf(m[2][1]) << 16);
  o1.y = (unsigned)f2bf(m[2][2]) | ((unsigned)f2bf(m[2][3]) << 16);
  o1.z = (unsigned)f2bf(m[3][0]) | ((unsigned)f2bf(m[3][1]) << 16);
  o1.w = (unsigned)f2bf(m[3][2]) | ((unsigned)f2bf(m[3][3]) << 16);
}

// Messages of one layer: M[r] = sum_{j in row i} A_ij (x) H[j] through the
// BF16 block matrix. One lane per (row, group of 4 channels) computing 16
// outputs, so a wavefront keeps 16 rows in flight. No LDS, so the gather runs
// at full occupancy. (The training forward pass; inference fuses the gather
// into the dense layer, k_layer, with the same arithmetic.)
template <int UNR>
__global__ __launch_bounds__(256) void k_msg(int n, int rows, const int* __restrict__ rowPtr,
                                             const int* __restrict__ colPtr, const bf16* __restrict__ A16,
                                             const bf16* __restrict__ H, bf16* __restrict__ M,
                                             const int* __restrict__ skip, RowGate gate, int xcd)
{
  if (skip && *skip) return;
  int bx = blockIdx.x;
  if (xcd) {
    bx = xcd_block(blockIdx.x, gridDim.x);
    if (bx < 0) return;
  }
  const size_t idx = (size_t)bx * 256 + threadIdx.x;
  const int r = (int)(idx >> 2), g = (int)(idx & 3);
  if (r >= rows) return;
  const int b = r / n, i = r - b * n;
  if (gate.sysOf && !gate.live(i)) return;
  float m[4][4];
  gather_row<UNR>(m, rowPtr[i], rowPtr[i + 1], g, colPtr, A16, H + (size_t)b * n * F + 16 * g);
  uint4 o0, o1;
  pack_msg(m, o0, o1);
  uint4* out = reinterpret_cast<uint4*>(M + (size_t)r * F + 16 * g);
  out[0] = o0;
  out[1] = o1;
}

// Dense part of one layer on a tile of 64 node rows: [h | m] and W staged in
// LDS, the 64 x 128 x 64 product on MFMA, residual ReLU update; training
// stores p; the last layer also decodes z = v + (Wd h) / sqrt(nNo).
__global__ __launch_bounds__(BS) void k_dense(int rows, const bf16* __restrict__ Hin, const bf16* __restrict__ M,
                                              bf16* __restrict__ Hout, const bf16* __restrict__ W,
                                              const float* __restrict__ bias, bf16* __restrict__ Pst,
                                              const bf16* __restrict__ Wd, const double* __restrict__ V,
                                              double* __restrict__ Z, float* __restrict__ Z32, double inv_sc,
                                              const double2* __restrict__ nsc, const int* __restrict__ skip,
                                              RowGate gate)
{
  __shared__ __align__(16) bf16 X[TILE * XS];
  if (skip && *skip) return;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int ntiles = (rows + TILE - 1) / TILE;
  // Wavefront w computes output columns 16w..16w+15 of every row tile: its MFMA
  // B operands (W rows 16w + lane%16, k = 4 (lane/16) + 16 kk + 0..3) are held
  // in registers for all of the workgroup's tiles, so LDS holds only [h | m]
  // (17 KB) and twice as many workgroups fit on a CU as with W staged there.
  const int fr = lane & 15, fg = lane >> 4;
  v4s bw[2 * F / 16];
#pragma unroll
  for (int kk = 0; kk < 2 * F / 16; kk++)
    bw[kk] = *reinterpret_cast<const v4s*>(W + (size_t)(16 * wave + fr) * (2 * F) + 4 * fg + 16 * kk);
  float bv[4];
#pragma unroll
  for (int qi = 0; qi < 4; qi++) {
    int row, col;
    acc_rc<F>(wave, lane, qi, 0, row, col);
    bv[qi] = bias[col];
  }
  for (int tl = blockIdx.x; tl < ntiles; tl += gridDim.x) {
    const int r0 = tl * TILE;
    if (gate.sysOf && !tile_live(gate, r0, rows, rows, t)) continue;
    for (int q = t; q < TILE * F / 8; q += BS) {
      const int row = q >> 3, f8 = (q & 7) * 8, r = r0 + row;
      uint4 hv = {0u, 0u, 0u, 0u}, mv = {0u, 0u, 0u, 0u};
      if (r < rows) {
        hv = *reinterpret_cast<const uint4*>(Hin + (size_t)r * F + f8);
        mv = *reinterpret_cast<const uint4*>(M + (size_t)r * F + f8);
      }
      *reinterpret_cast<uint4*>(X + row * XS + f8) = hv;
      *reinterpret_cast<uint4*>(X + row * XS + F + f8) = mv;
    }
    __syncthreads();
    v4f acc[4];
#pragma unroll
    for (int qi = 0; qi < 4; qi++) acc[qi] = v4f{bv[qi], bv[qi], bv[qi], bv[qi]};
    // the 64 x 128 x 64 product in mma_kc's order (row tile qi, column tile = wave)
#pragma unroll
    for (int qi = 0; qi < 4; qi++) {
      v4f c = acc[qi];
      const bf16* Ar = X + (16 * qi + fr) * XS + 4 * fg;
#pragma unroll
      for (int kk = 0; kk < 2 * F / 16; kk++)
        c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(*reinterpret_cast<const v4s*>(Ar + 16 * kk), bw[kk], c, 0, 0, 0);
      acc[qi] = c;
    }
    __syncthreads();
#pragma unroll
    for (int qi = 0; qi < 4; qi++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        int row, col;
        acc_rc<F>(wave, lane, qi, v, row, col);
        const float p = acc[qi][v];
        const bf16 ho = f2bf(bf2f(X[row * XS + col]) + fmaxf(p, 0.f));
        X[row * XS + col] = ho;
        const int r = r0 + row;
        if (r < rows) {
          Hout[(size_t)r * F + col] = ho;
          if (Pst) Pst[(size_t)r * F + col] = f2bf(p);
        }
      }
    __syncthreads();
    if (Wd) {
      const int node = t >> 2, d = t & 3, r = r0 + node;
      if (r < rows) {
        // y = sum_f Wd[d][f] h[f] in f order, 8 BF16 per 16-byte load
        const uint4* xr = reinterpret_cast<const uint4*>(X + node * XS);
        const uint4* wr = reinterpret_cast<const uint4*>(Wd + d * F);
        float y = 0.f;
#pragma unroll
        for (int f8 = 0; f8 < F / 8; f8++) {
          const uint4 xv = xr[f8], wv = wr[f8];
          y = fmaf(lo(wv.x), lo(xv.x), y);
          y = fmaf(hi(wv.x), hi(xv.x), y);
          y = fmaf(lo(wv.y), lo(xv.y), y);
          y = fmaf(hi(wv.y), hi(xv.y), y);
          y = fmaf(lo(wv.z), lo(xv.z), y);
          y = fmaf(hi(wv.z), hi(xv.z), y);
          y = fmaf(lo(wv.w), lo(xv.w), y);
          y = fmaf(hi(wv.w), hi(xv.w), y);
        }
        const double z = V[(size_t)r * 4 + d] + (double)y * (nsc ? nsc[r].y : inv_sc);
        Z[(size_t)r * 4 + d] = z;
        if (Z32) Z32[(size_t)r * 4 + d] = (float)z;
      }
      __syncthreads();
    }
  }
}

// One whole layer for inference on a tile of 64 node rows (one tile per
// workgroup): the rows' messages gathered into LDS beside their h (k_msg's
// arithmetic, rounded to BF16 as k_msg stores them), then k_dense's product,
// residual ReLU update and (last layer) decoder, so the messages never go
// through HBM. Bit for bit what k_msg followed by k_dense computes.
template <int UNR>
__global__ __launch_bounds__(BS) void k_layer(int n, int rows, const int* __restrict__ rowPtr,
                                              const int* __restrict__ colPtr, const bf16* __restrict__ A16,
                                              const bf16* __restrict__ Hin, bf16* __restrict__ Hout,
                                              const bf16* __restrict__ W, const float* __restrict__ bias,
                                              const bf16* __restrict__ Wd, const double* __restrict__ V,
                                              double* __restrict__ Z, float* __restrict__ Z32, double inv_sc,
                                              const double2* __restrict__ nsc, const int* __restrict__ skip,
                                              RowGate gate, int xcd)
{
  __shared__ __align__(16) bf16 X[TILE * XS];
  if (skip && *skip) return;
  int bx = blockIdx.x;
  if (xcd) {
    bx = xcd_block(blockIdx.x, gridDim.x);
    if (bx < 0) return;
  }
  const int r0 = bx * TILE;
  if (r0 >= rows) return;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  if (gate.sysOf && !tile_live(gate, r0, rows, rows, t)) return;
  // gather: thread t = 4 row + g computes channels 4g..4g+3 of its row
  {
    const int row = t >> 2, g = t & 3, r = r0 + row;
    uint4 m0 = {0u, 0u, 0u, 0u}, m1 = {0u, 0u, 0u, 0u}, h0 = {0u, 0u, 0u, 0u}, h1 = {0u, 0u, 0u, 0u};
    if (r < rows) {
      const int b = r / n, i = r - b * n;
      h0 = reinterpret_cast<const uint4*>(Hin + (size_t)r * F + 16 * g)[0];
      h1 = reinterpret_cast<const uint4*>(Hin + (size_t)r * F + 16 * g)[1];
      if (!gate.sysOf || gate.live(i)) {
        float m[4][4];
        gather_row<UNR>(m, rowPtr[i], rowPtr[i + 1], g, colPtr, A16, Hin + (size_t)b * n * F + 16 * g);
        pack_msg(m, m0, m1);
      }
    }
    uint4* xr = reinterpret_cast<uint4*>(X + row * XS + 16 * g);
    xr[0] = h0;
    xr[1] = h1;
    uint4* xm = reinterpret_cast<uint4*>(X + row * XS + F + 16 * g);
    xm[0] = m0;
    xm[1] = m1;
  }
  const int fr = lane & 15, fg = lane >> 4;
  v4s bw[2 * F / 16];
#pragma unroll
  for (int kk = 0; kk < 2 * F / 16; kk++)
    bw[kk] = *reinterpret_cast<const v4s*>(W + (size_t)(16 * wave + fr) * (2 * F) + 4 * fg + 16 * kk);
  v4f acc[4];
#pragma unroll
  for (int qi = 0; qi < 4; qi++) {
    int row, col;
    acc_rc<F>(wave, lane, qi, 0, row, col);
    const float b0 = bias[col];
    acc[qi] = v4f{b0, b0, b0, b0};
  }
  __syncthreads();
#pragma unroll
  for (int qi = 0; qi < 4; qi++) {
    v4f c = acc[qi];
    const bf16* Ar = X + (16 * qi + fr) * XS + 4 * fg;
#pragma unroll
    for (int kk = 0; kk < 2 * F / 16; kk++)
      c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(*reinterpret_cast<const v4s*>(Ar + 16 * kk), bw[kk], c, 0, 0, 0);
    acc[qi] = c;
  }
  __syncthreads();
#pragma unroll
  for (int qi = 0; qi < 4; qi++)
#pragma unroll
    for (int v = 0; v < 4; v++) {
      int row, col;
      acc_rc<F>(wave, lane, qi, v, row, col);
      const float p = acc[qi][v];
      X[row * XS + col] = f2bf(bf2f(X[row * XS + col]) + fmaxf(p, 0.f));
    }
  __syncthreads();
  // the tile's updated h rows leave LDS as 16-byte stores (the accumulator
  // layout would give 2-byte ones)
  for (int q = t; q < TILE * F / 8; q += BS) {
    const int row = q / (F / 8), c8 = q % (F / 8), r = r0 + row;
    if (r < rows) reinterpret_cast<uint4*>(Hout + (size_t)r * F)[c8] = reinterpret_cast<const uint4*>(X + row * XS)[c8];
  }
  if (Wd) {
    const int node = t >> 2, d = t & 3, r = r0 + node;
    if (r < rows) {
      const uint4* xr = reinterpret_cast<const uint4*>(X + node * XS);
      const uint4* wr = reinterpret_cast<const uint4*>(Wd + d * F);
      float y = 0.f;
#pragma unroll
      for (int f8 = 0; f8 < F / 8; f8++) {
        const uint4 xv = xr[f8], wv = wr[f8];
        y = fmaf(lo(wv.x), lo(xv.x), y);
        y = fmaf(hi(wv.x), hi(xv.x), y);
        y = fmaf(lo(wv.y), lo(xv.y), y);
        y = fmaf(hi(wv.y), hi(xv.y), y);
        y = fmaf(lo(wv.z), lo(xv.z), y);
        y = fmaf(hi(wv.z), hi(xv.z), y);
        y = fmaf(lo(wv.w), lo(xv.w), y);
        y = fmaf(hi(wv.w), hi(xv.w), y);
      }
      const double z = V[(size_t)r * 4 + d] + (double)y * (nsc ? nsc[r].y : inv_sc);
      Z[(size_t)r * 4 + d] = z;
      if (Z32) Z32[(size_t)r * 4 + d] = (float)z;
    }
  }
}

// ------------------------------------------------------------- backward ---
// GZ = -(2/B) A^T Res on the FP64 scaled matrix (the transposed block of
// row j's block k sits at trans[k]); one lane per (row, dof).
__global__ __launch_bounds__(256) void k_spmvT(int nNo, const int* __restrict__ rowPtr,
                                               const int* __restrict__ colPtr, const int* __restrict__ trans,
                                               const double* __restrict__ K, const double* __restrict__ U,
                                               double* __restrict__ Y, double alpha)
{
  const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int j = (int)(k >> 2), e = (int)(k & 3);
  if (j >= nNo) return;
  double acc = 0.0;
  for (int q = rowPtr[j]; q < rowPtr[j + 1]; q++) {
    const double* a = K + 16 * (size_t)trans[q] + e;
    const double* u = U + 4 * (size_t)colPtr[q];
    acc += a[0] * u[0] + a[4] * u[1] + a[8] * u[2] + a[12] * u[3];
  }
  Y[k] = alpha * acc;
}

__global__ __launch_bounds__(256) void k_sqsum(size_t n, const double* __restrict__ x, double* __restrict__ part)
{
  __shared__ double red[4];
  double acc = 0.0;
  for (size_t k = (size_t)blockIdx.x * 256 + threadIdx.x; k < n; k += (size_t)gridDim.x * 256) acc += x[k] * x[k];
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// Decoder backward: gy = gz / sqrt(nNo); G = Wd^T gy; partial dWd.
__global__ __launch_bounds__(BS) void k_dec_bwd(int rows, const double* __restrict__ GZ, double inv_sc,
                                                const bf16* __restrict__ Wd, const bf16* __restrict__ H,
                                                float* __restrict__ G, float* __restrict__ part)
{
  __shared__ float gy[TILE * 4];
  __shared__ bf16 Hs[TILE * GS];
  const int t = threadIdx.x, d = t >> 6, f = t & 63;
  const int ntiles = (rows + TILE - 1) / TILE;
  float acc = 0.f;
  const float wdf[4] = {bf2f(Wd[f]), bf2f(Wd[F + f]), bf2f(Wd[2 * F + f]), bf2f(Wd[3 * F + f])};
  for (int tl = blockIdx.x; tl < ntiles; tl += gridDim.x) {
    const int r0 = tl * TILE;
    {
      const int node = t >> 2, dd = t & 3, r = r0 + node;
      gy[t] = r < rows ? (float)(GZ[(size_t)r * 4 + dd] * inv_sc) : 0.f;
    }
    for (int q = t; q < TILE * F; q += BS) {
      const int row = q >> 6, ff = q & 63, r = r0 + row;
      Hs[row * GS + ff] = r < rows ? H[(size_t)r * F + ff] : (bf16)0;
    }
    __syncthreads();
    for (int node = t >> 6; node < TILE; node += 4) {
      const int r = r0 + node;
      if (r < rows) {
        float g = 0.f;
        for (int dd = 0; dd < 4; dd++) g = fmaf(wdf[dd], gy[node * 4 + dd], g);
        G[(size_t)r * F + f] = g;
      }
    }
    for (int node = 0; node < TILE; node++) acc = fmaf(gy[node * 4 + d], bf2f(Hs[node * GS + f]), acc);
    __syncthreads();
  }
  part[(size_t)blockIdx.x * (4 * F) + d * F + f] = acc;
}

// Layer backward on tiles of 64 rows: gp = Gout * (p > 0) (BF16),
// [Gin_direct | GM] = gp W (MFMA), dW += gp^T [h | m] (MFMA, kept in
// registers over the workgroup's tiles), db += sum gp.
__global__ __launch_bounds__(BS) void k_layer_bwd(int rows, const float* __restrict__ Gout,
                                                  const bf16* __restrict__ Hin, const bf16* __restrict__ Mst,
                                                  const bf16* __restrict__ Pst, const bf16* __restrict__ W,
                                                  float* __restrict__ Gin, float* __restrict__ GM,
                                                  float* __restrict__ partW, float* __restrict__ partB)
{
  __shared__ bf16 X[TILE * XS];
  __shared__ bf16 Ws[F * XS];
  __shared__ bf16 GP[TILE * GS];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int ntiles = (rows + TILE - 1) / TILE;
  for (int q = t; q < F * 2 * F; q += BS) Ws[(q >> 7) * XS + (q & 127)] = W[q];
  v4f accW[8];
#pragma unroll
  for (int qi = 0; qi < 8; qi++) accW[qi] = v4f{0.f, 0.f, 0.f, 0.f};
  float accB = 0.f;
  for (int tl = blockIdx.x; tl < ntiles; tl += gridDim.x) {
    const int r0 = tl * TILE;
    for (int q = t; q < TILE * F; q += BS) {
      const int row = q >> 6, f = q & 63, r = r0 + row;
      bf16 h = 0, m = 0, g = 0;
      if (r < rows) {
        const size_t o = (size_t)r * F + f;
        h = Hin[o];
        m = Mst[o];
        g = bf2f(Pst[o]) > 0.f ? f2bf(Gout[o]) : (bf16)0;
      }
      X[row * XS + f] = h;
      X[row * XS + F + f] = m;
      GP[row * GS + f] = g;
    }
    __syncthreads();
    if (t < F)
      for (int node = 0; node < TILE; node++) accB += bf2f(GP[node * GS + t]);
    v4f acc[8];
#pragma unroll
    for (int qi = 0; qi < 8; qi++) acc[qi] = v4f{0.f, 0.f, 0.f, 0.f};
    mma<TILE, 2 * F, F>(GP, GS, 1, Ws, XS, 1, acc, wave, lane);  // A(node,out), B(out,in) = W[out][in]
#pragma unroll
    for (int qi = 0; qi < 8; qi++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        int row, col;
        acc_rc<2 * F>(wave, lane, qi, v, row, col);
        const int r = r0 + row;
        if (r < rows) {
          if (col < F)
            Gin[(size_t)r * F + col] = Gout[(size_t)r * F + col] + acc[qi][v];
          else
            GM[(size_t)r * F + (col - F)] = acc[qi][v];
        }
      }
    mma<F, 2 * F, TILE>(GP, 1, GS, X, XS, 1, accW, wave, lane);  // A(out,node) = GP[node][out], B(node,in)
    __syncthreads();
  }
  float* pw = partW + (size_t)blockIdx.x * (F * 2 * F);
#pragma unroll
  for (int qi = 0; qi < 8; qi++)
#pragma unroll
    for (int v = 0; v < 4; v++) {
      int row, col;
      acc_rc<2 * F>(wave, lane, qi, v, row, col);
      pw[row * 2 * F + col] = accW[qi][v];
    }
  if (t < F) partB[(size_t)blockIdx.x * F + t] = accB;
}

// Transposed message passing: Gin[j] += sum_{i : j in row i} A_ij^T (x) GM[i];
// one lane per (row j, channel c).
__global__ __launch_bounds__(256) void k_spmmT(int n, int rows, const int* __restrict__ rowPtr,
                                               const int* __restrict__ colPtr, const int* __restrict__ trans,
                                               const bf16* __restrict__ A16, const float* __restrict__ GM,
                                               float* __restrict__ Gin)
{
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int r = (int)(idx >> 4), c = (int)(idx & 15);
  if (r >= rows) return;
  const int b = r / n, j = r - b * n;
  const float* Gb = GM + (size_t)b * n * F + 4 * c;
  float e0 = 0.f, e1 = 0.f, e2 = 0.f, e3 = 0.f;
  for (int k = rowPtr[j]; k < rowPtr[j + 1]; k++) {
    const int i = colPtr[k];
    const bf16* a = A16 + 16 * (size_t)trans[k];  // A_ij, row-major [d][e]
    const float4 g = *reinterpret_cast<const float4*>(Gb + (size_t)i * F);
    const float gd[4] = {g.x, g.y, g.z, g.w};
#pragma unroll
    for (int d = 0; d < 4; d++) {
      e0 = fmaf(bf2f(a[4 * d + 0]), gd[d], e0);
      e1 = fmaf(bf2f(a[4 * d + 1]), gd[d], e1);
      e2 = fmaf(bf2f(a[4 * d + 2]), gd[d], e2);
      e3 = fmaf(bf2f(a[4 * d + 3]), gd[d], e3);
    }
  }
  float4* o = reinterpret_cast<float4*>(Gin + (size_t)r * F + 4 * c);
  float4 v = *o;
  v.x += e0;
  v.y += e1;
  v.z += e2;
  v.w += e3;
  *o = v;
}

// Encoder backward: dW0 += bf16(G0)^T u on MFMA.
__global__ __launch_bounds__(BS) void k_enc_bwd(int n, int rows, const double* __restrict__ V, double sc,
                                                const bf16* __restrict__ Dg, const float* __restrict__ G0,
                                                float* __restrict__ part)
{
  __shared__ bf16 U[TILE * US];
  __shared__ bf16 GH[TILE * GS];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int ntiles = (rows + TILE - 1) / TILE;
  v4f acc[2];
  acc[0] = v4f{0.f, 0.f, 0.f, 0.f};
  acc[1] = v4f{0.f, 0.f, 0.f, 0.f};
  for (int tl = blockIdx.x; tl < ntiles; tl += gridDim.x) {
    const int r0 = tl * TILE;
    load_u(U, t, r0, rows, n, V, sc, nullptr, Dg);
    for (int q = t; q < TILE * F; q += BS) {
      const int row = q >> 6, f = q & 63, r = r0 + row;
      GH[row * GS + f] = r < rows ? f2bf(G0[(size_t)r * F + f]) : (bf16)0;
    }
    __syncthreads();
    mma<F, NIN, TILE>(GH, 1, GS, U, US, 1, acc, wave, lane);  // A(out,node), B(node,in)
    __syncthreads();
  }
  float* pw = part + (size_t)blockIdx.x * (F * NIN);
#pragma unroll
  for (int qi = 0; qi < 2; qi++)
#pragma unroll
    for (int v = 0; v < 4; v++) {
      int row, col;
      acc_rc<NIN>(wave, lane, qi, v, row, col);
      pw[row * NIN + col] = acc[qi][v];
    }
}

// out[e] = sum_p part[p*len + e] (p in order).
__global__ __launch_bounds__(256) void k_reduce(int nparts, int len, const float* __restrict__ part,
                                                float* __restrict__ out)
{
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= len) return;
  float s = 0.f;
  for (int p = 0; p < nparts; p++) s += part[(size_t)p * len + e];
  out[e] = s;
}

__global__ __launch_bounds__(256) void k_adam(int np, float* __restrict__ p, const float* __restrict__ g,
                                              float* __restrict__ m, float* __restrict__ v, bf16* __restrict__ pb,
                                              float lr, float b1, float b2, float eps, float bc1, float bc2)
{
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= np) return;
  const float ge = g[e];
  const float mm = b1 * m[e] + (1.f - b1) * ge;
  const float vv = b2 * v[e] + (1.f - b2) * ge * ge;
  m[e] = mm;
  v[e] = vv;
  const float pe = p[e] - lr * (mm / bc1) / (sqrtf(vv / bc2) + eps);
  p[e] = pe;
  pb[e] = f2bf(pe);
}

__global__ __launch_bounds__(256) void k_to_bf16(int np, const float* __restrict__ p, bf16* __restrict__ pb)
{
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e < np) pb[e] = f2bf(p[e]);
}

}  // namespace lpc
}  // namespace svg

using namespace svg::lpc;

struct svg_lpc {
  float* p = nullptr;     // FP32 master parameters
  bf16* pb = nullptr;     // BF16 copy used by the kernels
  float* g = nullptr;     // gradient
  float* am = nullptr;    // Adam moments
  float* av = nullptr;
  int step = 0;
  // activations (rows = nvec * nNo)
  int cap = 0;
  bf16* H[NL + 1] = {};
  bf16* M[NL] = {};
  bf16* P[NL] = {};
  float *G = nullptr, *G2 = nullptr, *GM = nullptr;
  double *Z = nullptr, *Res = nullptr, *GZ = nullptr;
  float* part = nullptr;  // weight-gradient partials
  double* sq = nullptr;   // loss partials
  double* h_sq = nullptr;
  // Changes whenever the buffers a captured forward pass bakes in (the
  // activations) are reallocated, and differs between networks: graphs that
  // contain the forward pass are keyed on it.
  unsigned long long gen = 0;
};

namespace svg {

static std::atomic<unsigned long long> g_lpc_gen{0};

unsigned long long lpc_generation(const svg_lpc* N) { return N ? N->gen : 0ull; }

static void lpc_free_acts(svg_lpc* N)
{
  for (auto& h : N->H) dev_free(h);
  for (auto& m : N->M) dev_free(m);
  for (auto& p : N->P) dev_free(p);
  dev_free(N->G);
  dev_free(N->G2);
  dev_free(N->GM);
  dev_free(N->Z);
  dev_free(N->Res);
  dev_free(N->GZ);
  N->cap = 0;
}

// Activation buffers for `rows` node rows; training also keeps m, p and the
// FP32 gradients.
static int lpc_ensure(svg_lpc* N, int rows, bool train)
{
  const bool have_train = N->G != nullptr;
  if (rows <= N->cap && (!train || have_train)) return SVG_OK;
  const int cap = rows > N->cap ? rows : N->cap;
  lpc_free_acts(N);
  int rc;
  for (auto& h : N->H)
    if ((rc = dev_alloc(&h, (size_t)cap * F))) return rc;
  for (int l = 0; l < (train ? NL : 1); l++)
    if ((rc = dev_alloc(&N->M[l], (size_t)cap * F))) return rc;
  if (train) {
    for (auto& p : N->P)
      if ((rc = dev_alloc(&p, (size_t)cap * F))) return rc;
    if ((rc = dev_alloc(&N->G, (size_t)cap * F)) || (rc = dev_alloc(&N->G2, (size_t)cap * F)) ||
        (rc = dev_alloc(&N->GM, (size_t)cap * F)) || (rc = dev_alloc(&N->Z, (size_t)cap * 4)) ||
        (rc = dev_alloc(&N->Res, (size_t)cap * 4)) || (rc = dev_alloc(&N->GZ, (size_t)cap * 4)))
      return rc;
  }
  N->cap = cap;
  N->gen = ++g_lpc_gen;
  return SVG_OK;
}

static int lhs_trans(svg_lhs* L)
{
  if (L->d_trans) return SVG_OK;
  std::vector<int> tr(L->nnz);
  for (int i = 0; i < L->nNo; i++)
    for (int j = L->rowPtr[i]; j < L->rowPtr[i + 1]; j++) {
      const int k = L->colPtr[j];
      const int* b = L->colPtr.data() + L->rowPtr[k];
      const int* e = L->colPtr.data() + L->rowPtr[k + 1];
      const int* it = std::lower_bound(b, e, i);
      if (it == e || *it != i) return set_error(SVG_ERR_ARG, "the block pattern is not symmetric");
      tr[j] = (int)(it - L->colPtr.data());
    }
  int rc = dev_alloc(&L->d_trans, (size_t)L->nnz);
  if (rc) return rc;
  SVG_HIP(hipMemcpy(L->d_trans, tr.data(), sizeof(int) * L->nnz, hipMemcpyHostToDevice));
  return SVG_OK;
}

int lhs_ensure_trans(svg_lhs* L) { return lhs_trans(L); }

static int lhs_bf16_alloc(svg_lhs* L)
{
  int rc;
  if (!L->d_val16 && (rc = dev_alloc(&L->d_val16, (size_t)16 * L->nnz))) return rc;
  if (!L->d_diag16 && (rc = dev_alloc(&L->d_diag16, (size_t)16 * L->nNo))) return rc;
  return SVG_OK;
}

int lpc_prepare_dev(svg_lhs* L, const double* Val, hipStream_t s)
{
  int rc;
  if ((rc = lhs_bf16_alloc(L))) return rc;
  launch("lpc_prep", k_val_bf16, dim3(4096), dim3(256), 0, s, (size_t)16 * L->nnz, Val, (bf16*)L->d_val16);
  launch("lpc_prep", k_diag_bf16, dim3(blocks_for((size_t)16 * L->nNo, 256)), dim3(256), 0, s, L->nNo,
         (const int*)L->d_diagPtr, Val, (bf16*)L->d_diag16);
  SVG_HIP(hipGetLastError());
  return SVG_OK;
}

// Forward pass on nvec stacked vectors V (nvec, nNo, 4) -> Z; training keeps
// m and p of every layer. nsc (nvec = 1 only): per-node (sc, 1/sc) in place of
// sqrt(nNo) (independent systems of a block-diagonal union, each scaled by its own size).
static int lpc_forward(svg_lpc* N, const svg_lhs* L, int nvec, const double* V, double* Z, float* Z32,
                       bool train, const int* skip, hipStream_t s, const double2* nsc = nullptr,
                       const int* sysOf = nullptr, const GmState* st = nullptr)
{
  const int n = L->nNo, rows = nvec * n;
  if ((nsc || sysOf) && nvec != 1) return set_error(SVG_ERR_ARG, "lpc_forward: per-node data need one vector");
  const lpc::RowGate gate{sysOf, st};
  const double sc = std::sqrt((double)n);
  const unsigned nt = blocks_for(rows, TILE);
  const bf16* pb = N->pb;
  // algorithmic bytes: V and the diagonal blocks read, h written
  launch_b("lpc_encode", (32.0 + 32.0 + 2.0 * F) * rows, k_encode, dim3(std::min(nt, 2048u)), dim3(BS), 0, s, n,
           rows, V, sc, nsc, (const bf16*)L->d_diag16, pb + OFF_W0, N->H[0], skip, gate);
  unsigned nm = blocks_for((size_t)rows * 4, 256);
  const int mx = (L->msgVariant == 1 && nm >= 64) ? 1 : 0;
  if (mx) nm = (nm + 7) & ~7u;
  if (!train) {
    // inference: one fused launch per layer, one 64-row tile per workgroup
    unsigned nl = nt;
    if (mx) nl = (nl + 7) & ~7u;
    for (int l = 0; l < NL; l++) {
      const bool last = l == NL - 1;
      // BF16 blocks + colPtr per block, rowPtr + h in + h out per row (+ V in, z out for the last layer)
      const double db = (32.0 + 4.0) * L->nnz * nvec + (4.0 + 4.0 * F) * rows +
                        (last ? (64.0 + (Z32 ? 16.0 : 0.0)) * rows : 0.0);
      launch_b("lpc_layer", db, k_layer<4>, dim3(nl), dim3(BS), 0, s, n, rows, (const int*)L->d_rowPtr,
               (const int*)L->d_colPtr, (const bf16*)L->d_val16, (const bf16*)N->H[l & 1], N->H[(l + 1) & 1],
               pb + OFF_L + l * LSZ, (const float*)(N->p + OFF_L + l * LSZ + F * 2 * F),
               last ? pb + OFF_WD : (const bf16*)nullptr, V, last ? Z : (double*)nullptr,
               last ? Z32 : (float*)nullptr, 1.0 / sc, nsc, skip, gate, mx);
    }
    SVG_HIP(hipGetLastError());
    return SVG_OK;
  }
  for (int l = 0; l < NL; l++) {
    const bool last = l == NL - 1;
    const bf16* Hin = train ? N->H[l] : N->H[l & 1];
    bf16* Hout = train ? N->H[l + 1] : N->H[(l + 1) & 1];
    bf16* Ml = train ? N->M[l] : N->M[0];
    // BF16 blocks + colPtr per block (per sample), rowPtr + h + m out per row
    launch_b("lpc_msg", (32.0 + 4.0) * L->nnz * nvec + (4.0 + 2.0 * F) * rows, k_msg<4>, dim3(nm), dim3(256), 0, s, n,
             rows, (const int*)L->d_rowPtr, (const int*)L->d_colPtr, (const bf16*)L->d_val16, Hin, Ml, skip, gate,
             mx);
    // h, m in, h out (+ p for training; + V in, z out for the last layer)
    const double db = 6.0 * F * rows + (train ? 2.0 * F * rows : 0.0) + (last ? (64.0 + (Z32 ? 16.0 : 0.0)) * rows : 0.0);
    launch_b("lpc_dense", db, k_dense, dim3(std::min(nt, 2048u)), dim3(BS), 0, s, rows, Hin, (const bf16*)Ml, Hout,
             pb + OFF_L + l * LSZ, (const float*)(N->p + OFF_L + l * LSZ + F * 2 * F),
             train ? N->P[l] : (bf16*)nullptr, last ? pb + OFF_WD : (const bf16*)nullptr, V,
             last ? Z : (double*)nullptr, last ? Z32 : (float*)nullptr, 1.0 / sc, nsc, skip, gate);
  }
  SVG_HIP(hipGetLastError());
  return SVG_OK;
}

int lpc_apply_dev(svg_lpc* N, const svg_lhs* L, const double* V, double* Z, float* Z32, const int* skip,
                  hipStream_t s, const double2* nsc, const int* sysOf, const GmState* st)
{
  int rc = lpc_ensure(N, L->nNo, false);
  if (rc) return rc;
  return lpc_forward(N, L, 1, V, Z, Z32, false, skip, s, nsc, sysOf, st);
}

// The message gather's launch variant (0: round-robin workgroups, 1: one row
// band per XCD, whose L2 then keeps the h rows its band gathers: 155 against
// 278 MB of HBM reads per launch on the 1.2M-tet pipe; identical results),
// timed once per lhs on its buffers: the block copy as it is (its values do
// not change the time) and the network's scratch rows. (Non-temporal loads of
// the streamed blocks raised the reads to 278-440 MB and are not a variant.)
static int lpc_msg_tune(svg_lpc* N, const svg_lhs* L, hipStream_t s)
{
  if (L->msgVariant >= 0) return SVG_OK;
  SVG_HIP(hipMemsetAsync(N->H[0], 0, sizeof(bf16) * (size_t)F * L->nNo, s));
  hipEvent_t a, b;
  SVG_HIP(hipEventCreate(&a));
  SVG_HIP(hipEventCreate(&b));
  const unsigned nm0 = blocks_for((size_t)L->nNo * 4, 256);
  float best = 1e30f;
  int bv = 0;
  for (int v = 0; v < 2; v++) {
    const int x = (v == 1 && nm0 >= 64) ? 1 : 0;
    const unsigned nm = x ? ((nm0 + 7) & ~7u) : nm0;
    for (int r = 0; r < 4; r++) {
      if (r == 1) SVG_HIP(hipEventRecord(a, s));
      launch("lpc_msg_tune", k_msg<4>, dim3(nm), dim3(256), 0, s, L->nNo, L->nNo, (const int*)L->d_rowPtr,
             (const int*)L->d_colPtr, (const bf16*)L->d_val16, (const bf16*)N->H[0], N->M[0], (const int*)nullptr,
             lpc::RowGate{nullptr, nullptr}, x);
    }
    SVG_HIP(hipEventRecord(b, s));
    SVG_HIP(hipEventSynchronize(b));
    float ms = 0.f;
    SVG_HIP(hipEventElapsedTime(&ms, a, b));
    if (ms < best) {
      best = ms;
      bv = v;
    }
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  SVG_HIP(hipGetLastError());
  L->msgVariant = bv;
  return SVG_OK;
}

int lpc_reserve(svg_lpc* N, svg_lhs* L, hipStream_t s)
{
  int rc = lhs_bf16_alloc(L);
  if (rc || (rc = lpc_ensure(N, L->nNo, false))) return rc;
  return lpc_msg_tune(N, L, s);
}

// Loss and gradient of L = (1/B) sum_b |R_b - A M(R_b)|^2 into N->g.
// Loss (1/nvec) sum_b |T_b - A M(R_b)|^2 and its gradient: R the network's
// input, T the target (null: T = R).
static int lpc_loss_grad(svg_lpc* N, svg_lhs* L, const double* Val, int nvec, const double* R, const double* T,
                         double* loss, hipStream_t s)
{
  const int n = L->nNo, rows = nvec * n;
  int rc;
  if ((rc = lhs_trans(L)) || (rc = lpc_ensure(N, rows, true))) return rc;
  if (!N->part && (rc = dev_alloc(&N->part, (size_t)MAXGRID * (F * 2 * F + F)))) return rc;
  if (!N->sq) {
    if ((rc = dev_alloc(&N->sq, 256))) return rc;
    SVG_HIP(hipHostMalloc((void**)&N->h_sq, 256 * sizeof(double), hipHostMallocDefault));
  }
  if ((rc = lpc_forward(N, L, nvec, R, N->Z, nullptr, true, nullptr, s))) return rc;
  const double sc = std::sqrt((double)n);
  const unsigned nbv = blocks_for((size_t)4 * n, 256);
  for (int b = 0; b < nvec; b++) {
    const size_t o = (size_t)b * 4 * n;
    // Res = T - A Z (FP64, the solver's SpMV), GZ = -(2/B) A^T Res
    if ((rc = spmv_residual(L, Val, N->Z + o, (T ? T : R) + o, N->Res + o, s))) return rc;
    launch("lpc_spmvT", k_spmvT, dim3(nbv), dim3(256), 0, s, n, (const int*)L->d_rowPtr, (const int*)L->d_colPtr,
           (const int*)L->d_trans, Val, (const double*)(N->Res + o), N->GZ + o, -2.0 / nvec);
  }
  launch("lpc_loss", k_sqsum, dim3(256), dim3(256), 0, s, (size_t)rows * 4, (const double*)N->Res, N->sq);
  const unsigned ng = std::min<unsigned>(MAXGRID, blocks_for(rows, TILE));
  const bf16* pb = N->pb;
  float* g = N->g;
  launch("lpc_bwd", k_dec_bwd, dim3(ng), dim3(BS), 0, s, rows, (const double*)N->GZ, 1.0 / sc, pb + OFF_WD,
         (const bf16*)N->H[NL], N->G, N->part);
  launch("lpc_bwd", k_reduce, dim3(blocks_for(4 * F, 256)), dim3(256), 0, s, (int)ng, 4 * F,
         (const float*)N->part, g + OFF_WD);
  float* Gout = N->G;
  float* Gin = N->G2;
  for (int l = NL - 1; l >= 0; l--) {
    launch("lpc_bwd", k_layer_bwd, dim3(ng), dim3(BS), 0, s, rows, (const float*)Gout, (const bf16*)N->H[l],
           (const bf16*)N->M[l], (const bf16*)N->P[l], pb + OFF_L + l * LSZ, Gin, N->GM, N->part,
           N->part + (size_t)ng * F * 2 * F);
    launch("lpc_bwd", k_reduce, dim3(blocks_for(F * 2 * F, 256)), dim3(256), 0, s, (int)ng, F * 2 * F,
           (const float*)N->part, g + OFF_L + l * LSZ);
    launch("lpc_bwd", k_reduce, dim3(1), dim3(256), 0, s, (int)ng, F, (const float*)(N->part + (size_t)ng * F * 2 * F),
           g + OFF_L + l * LSZ + F * 2 * F);
    launch("lpc_bwd", k_spmmT, dim3(blocks_for((size_t)rows * 16, 256)), dim3(256), 0, s, n, rows,
           (const int*)L->d_rowPtr, (const int*)L->d_colPtr, (const int*)L->d_trans, (const bf16*)L->d_val16,
           (const float*)N->GM, Gin);
    std::swap(Gout, Gin);
  }
  launch("lpc_bwd", k_enc_bwd, dim3(ng), dim3(BS), 0, s, n, rows, R, sc, (const bf16*)L->d_diag16,
         (const float*)Gout, N->part);
  launch("lpc_bwd", k_reduce, dim3(blocks_for(F * NIN, 256)), dim3(256), 0, s, (int)ng, F * NIN,
         (const float*)N->part, g + OFF_W0);
  SVG_HIP(hipGetLastError());
  SVG_HIP(hipMemcpyAsync(N->h_sq, N->sq, 256 * sizeof(double), hipMemcpyDeviceToHost, s));
  SVG_HIP(hipStreamSynchronize(s));
  double acc = 0.0;
  for (int q = 0; q < 256; q++) acc += N->h_sq[q];
  *loss = acc / nvec;
  return SVG_OK;
}

}  // namespace svg

using namespace svg;

extern "C" {

int svg_lpc_num_params(int* count)
{
  if (!count) return set_error(SVG_ERR_ARG, "svg_lpc_num_params: null argument");
  *count = NPARAM;
  return SVG_OK;
}

int svg_lpc_create(svg_lpc** out)
{
  if (!out) return set_error(SVG_ERR_ARG, "svg_lpc_create: null argument");
  auto* N = new svg_lpc();
  int rc;
  if ((rc = dev_alloc(&N->p, NPARAM)) || (rc = dev_alloc(&N->pb, NPARAM)) || (rc = dev_alloc(&N->g, NPARAM)) ||
      (rc = dev_alloc(&N->am, NPARAM)) || (rc = dev_alloc(&N->av, NPARAM))) {
    svg_lpc_free(N);
    return rc;
  }
  hipError_t e = hipMemset(N->p, 0, sizeof(float) * NPARAM);
  if (e == hipSuccess) e = hipMemset(N->pb, 0, sizeof(bf16) * NPARAM);
  if (e == hipSuccess) e = hipMemset(N->am, 0, sizeof(float) * NPARAM);
  if (e == hipSuccess) e = hipMemset(N->av, 0, sizeof(float) * NPARAM);
  if (e != hipSuccess) {
    svg_lpc_free(N);
    return hip_check(e, "hipMemset");
  }
  N->gen = ++g_lpc_gen;
  *out = N;
  return SVG_OK;
}

int svg_lpc_free(svg_lpc* N)
{
  if (!N) return SVG_OK;
  lpc_free_acts(N);
  dev_free(N->p);
  dev_free(N->pb);
  dev_free(N->g);
  dev_free(N->am);
  dev_free(N->av);
  dev_free(N->part);
  dev_free(N->sq);
  if (N->h_sq) (void)hipHostFree(N->h_sq);
  delete N;
  return SVG_OK;
}

int svg_lpc_set_params(svg_lpc* N, const float* params)
{
  if (!N || !params) return set_error(SVG_ERR_ARG, "svg_lpc_set_params: null argument");
  SVG_HIP(hipMemcpy(N->p, params, sizeof(float) * NPARAM, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_to_bf16, dim3(blocks_for(NPARAM, 256)), dim3(256), 0, 0, NPARAM, (const float*)N->p, N->pb);
  SVG_HIP(hipGetLastError());
  SVG_HIP(hipMemset(N->am, 0, sizeof(float) * NPARAM));
  SVG_HIP(hipMemset(N->av, 0, sizeof(float) * NPARAM));
  N->step = 0;
  SVG_HIP(hipDeviceSynchronize());
  return SVG_OK;
}

int svg_lpc_get_params(const svg_lpc* N, float* params)
{
  if (!N || !params) return set_error(SVG_ERR_ARG, "svg_lpc_get_params: null argument");
  SVG_HIP(hipMemcpy(params, N->p, sizeof(float) * NPARAM, hipMemcpyDeviceToHost));
  return SVG_OK;
}

int svg_lpc_prepare(svg_lhs* L, const double* d_Val, void* stream)
{
  if (!L || !d_Val) return set_error(SVG_ERR_ARG, "svg_lpc_prepare: null argument");
  int rc = lpc_prepare_dev(L, d_Val, (hipStream_t)stream);
  if (rc) return rc;
  SVG_HIP(hipStreamSynchronize((hipStream_t)stream));
  return SVG_OK;
}

int svg_lpc_apply(svg_lpc* N, const svg_lhs* L, int nvec, const double* d_V, double* d_Z, void* stream)
{
  if (!N || !L || !d_V || !d_Z || nvec < 1) return set_error(SVG_ERR_ARG, "svg_lpc_apply: bad argument");
  if (!L->d_val16) return set_error(SVG_ERR_ARG, "svg_lpc_apply: svg_lpc_prepare was not called on this lhs");
  if ((long long)nvec * L->nNo > (1LL << 30)) return set_error(SVG_ERR_ARG, "svg_lpc_apply: too many rows");
  hipStream_t s = (hipStream_t)stream;
  int rc = lpc_ensure(N, nvec * L->nNo, false);
  if (rc || (rc = lpc_msg_tune(N, L, s))) return rc;
  if ((rc = lpc_forward(N, L, nvec, d_V, d_Z, nullptr, false, nullptr, s))) return rc;
  SVG_HIP(hipStreamSynchronize(s));
  return SVG_OK;
}

int svg_lpc_loss_grad(svg_lpc* N, svg_lhs* L, const double* d_Val, int nvec, const double* d_R, double* loss,
                      float* grad, void* stream)
{
  if (!N || !L || !d_Val || !d_R || !loss || nvec < 1) return set_error(SVG_ERR_ARG, "svg_lpc_loss_grad: bad argument");
  if (!L->d_val16) return set_error(SVG_ERR_ARG, "svg_lpc_loss_grad: svg_lpc_prepare was not called on this lhs");
  if ((long long)nvec * L->nNo > (1LL << 28)) return set_error(SVG_ERR_ARG, "svg_lpc_loss_grad: too many rows");
  hipStream_t s = (hipStream_t)stream;
  int rc = lpc_loss_grad(N, L, d_Val, nvec, d_R, nullptr, loss, s);
  if (rc) return rc;
  if (grad) SVG_HIP(hipMemcpy(grad, N->g, sizeof(float) * NPARAM, hipMemcpyDeviceToHost));
  return SVG_OK;
}

int svg_lpc_train_step(svg_lpc* N, svg_lhs* L, const double* d_Val, int nvec, const double* d_R, float lr,
                       double* loss, void* stream)
{
  return svg_lpc_train_step_ex(N, L, d_Val, nvec, d_R, nullptr, lr, loss, stream);
}

int svg_lpc_train_step_ex(svg_lpc* N, svg_lhs* L, const double* d_Val, int nvec, const double* d_V, const double* d_T,
                          float lr, double* loss, void* stream)
{
  if (!N || !L || !d_Val || !d_V || !loss || nvec < 1) return set_error(SVG_ERR_ARG, "svg_lpc_train_step: bad argument");
  if (!L->d_val16) return set_error(SVG_ERR_ARG, "svg_lpc_train_step: svg_lpc_prepare was not called on this lhs");
  if ((long long)nvec * L->nNo > (1LL << 28)) return set_error(SVG_ERR_ARG, "svg_lpc_train_step: too many rows");
  hipStream_t s = (hipStream_t)stream;
  int rc = lpc_loss_grad(N, L, d_Val, nvec, d_V, d_T, loss, s);
  if (rc) return rc;
  N->step++;
  const float b1 = 0.9f, b2 = 0.999f;
  const float bc1 = 1.f - std::pow(b1, (float)N->step), bc2 = 1.f - std::pow(b2, (float)N->step);
  launch("lpc_adam", k_adam, dim3(blocks_for(NPARAM, 256)), dim3(256), 0, s, NPARAM, N->p, (const float*)N->g, N->am,
         N->av, N->pb, lr, b1, b2, 1e-8f, bc1, bc2);
  SVG_HIP(hipGetLastError());
  SVG_HIP(hipStreamSynchronize(s));
  return SVG_OK;
}

}  // extern "C"
