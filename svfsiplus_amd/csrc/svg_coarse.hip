// Two-level preconditioner of the diagonally scaled fluid system (no
// counterpart in the reference; it stands beside precond_diag,
// liner_solver/precond.cpp:122-257, and is used only by the flexible GMRES of
// svg_fsils_solve_ex):
//
//   M(v) = P zc + B (v - A P zc),   zc = Ac^-1 P^T v,   Ac = P^T A P
//
//   P   piecewise-constant prolongation of the pressure (dof 3) over node
//       aggregates: the global pressure modes that a local smoother cannot
//       reach (the Laplacian-like continuity block of the VMS P1-P1 system,
//       fluid.cpp:1662-1688) are solved on the aggregates
//   Ac  dense nAgg x nAgg; on the diagonally scaled system it is symmetric and
//       negative definite, so it is inverted by blocked Gauss-Jordan elimination
//       without pivoting (FP64), and applied as an FP32 matrix-vector product.
//       A union of independent systems (the batched solve) has one dense block
//       per system (its aggregates never cross systems): the elimination runs
//       over all blocks at once, each block with the arithmetic a single
//       system of that size gets.
//   B   block Jacobi: the inverses of the 4x4 diagonal blocks of A (FP32)
//
// Aggregates are grown breadth-first over the lhs's block pattern (a greedy
// aggregation that needs no coordinates, so it sees only what fsils_solve
// sees), with undersized ones merged into the neighbour they share most
// edges with. Every reduction runs in a fixed order, so the preconditioner is
// deterministic.
#include "svg_device.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <map>
#include <numeric>
#include <vector>

namespace svg {

constexpr int GJB = 64;              // Gauss-Jordan block size
constexpr int COARSE_MAX = 8192;     // most aggregates of one system (dense FP64 block: 512 MB)

struct CoarseLevel {
  int aggSize = 0;    // target nodes per aggregate this level was built for
  int nsys = 0;       // systems of a batched union (0: one system)
  int nAgg = 0, nAP = 0, nSeg = 0, nNo = 0;
  // dense blocks, one per system: block q holds the aggregates
  // [aggOff[q], aggOff[q] + n_q) in an N_q x N_q row-major matrix (N_q = n_q
  // rounded up to GJB, identity on the padding) at element offset blkOff[q];
  // its Gauss-Jordan column panel (N_q x GJB) and row panel (GJB x N_q) at
  // panOff[q], its two pivot-block slots at GJB * GJB * 2q
  int nblk = 0, maxN = 0;
  size_t total = 0, panTotal = 0;
  std::vector<int> aggOff, blkN;
  int *d_aggOff = nullptr, *d_blkN = nullptr, *d_aggBlk = nullptr;
  long long *d_blkOff = nullptr, *d_panOff = nullptr;
  int *d_agg = nullptr, *d_aggPtr = nullptr, *d_aggNodes = nullptr;
  int *d_apPtr = nullptr, *d_apAgg = nullptr, *d_apOf = nullptr;
  int *d_segPtr = nullptr, *d_segPerm = nullptr;
  long long* d_segIdx = nullptr;
  float* d_apVal = nullptr;   // (4, nAP): column 3 of A summed over each AP entry's blocks
  double* d_apPP = nullptr;   // (nAP): the pressure-pressure entry of the same sums, FP64
  float* d_dinv = nullptr;    // (16, nNo): inverses of the diagonal blocks, row-major
  double* d_Ac = nullptr;     // the blocks, inverted in place
  float* d_Ai = nullptr;      // FP32 copy of the inverted blocks
  double *d_piv = nullptr, *d_cp = nullptr, *d_rp = nullptr;  // Gauss-Jordan panels
  float *d_rc = nullptr, *d_zc = nullptr;                     // restricted vector, coarse solution
  int* d_flag = nullptr;
  int* h_flag = nullptr;
  // differs between builds (a graph holding this level's launches is keyed on it)
  unsigned long long gen = 0;
  bool inverted = false;  // d_Ai holds the inverse of a system's Ac (a keep set-up may reuse it)
};

static std::atomic<unsigned long long> g_coarse_gen{0};

unsigned long long coarse_generation(const svg_lhs* L) { return L->coarse ? L->coarse->gen : 0ull; }

void coarse_free(CoarseLevel*& c)
{
  if (!c) return;
  for (int* p : {c->d_agg, c->d_aggPtr, c->d_aggNodes, c->d_apPtr, c->d_apAgg, c->d_apOf, c->d_segPtr, c->d_segPerm,
                 c->d_flag, c->d_aggOff, c->d_blkN, c->d_aggBlk})
    if (p) (void)hipFree(p);
  for (long long* p : {c->d_segIdx, c->d_blkOff, c->d_panOff})
    if (p) (void)hipFree(p);
  for (float* p : {c->d_apVal, c->d_dinv, c->d_Ai, c->d_rc, c->d_zc})
    if (p) (void)hipFree(p);
  for (double* p : {c->d_apPP, c->d_Ac, c->d_piv, c->d_cp, c->d_rp})
    if (p) (void)hipFree(p);
  if (c->h_flag) (void)hipHostFree(c->h_flag);
  delete c;
  c = nullptr;
}

// ------------------------------------------------------------- host setup --
// Breadth-first aggregation of the node graph: seeds in node order, each
// aggregate grows over unaggregated neighbours until it holds `target` nodes;
// then every aggregate under target/4 nodes joins the neighbouring aggregate
// (of at least target/4 nodes) it shares most edges with (ties: the lowest
// id). Labels are renumbered in order of first node.
static void aggregate(const svg_lhs* L, int target, std::vector<int>& agg, int& nAgg)
{
  const int nNo = L->nNo;
  const auto& rp = L->rowPtr;
  const auto& cp = L->colPtr;
  agg.assign(nNo, -1);
  std::vector<int> size, q;
  int na = 0;
  for (int s = 0; s < nNo; s++) {
    if (agg[s] >= 0) continue;
    q.clear();
    q.push_back(s);
    agg[s] = na;
    int cnt = 1;
    for (size_t h = 0; h < q.size() && cnt < target; h++) {
      const int v = q[h];
      for (int k = rp[v]; k < rp[v + 1] && cnt < target; k++) {
        const int w = cp[k];
        if (agg[w] < 0) {
          agg[w] = na;
          q.push_back(w);
          cnt++;
        }
      }
    }
    size.push_back(cnt);
    na++;
  }
  const int small = std::max(1, target / 4);
  std::vector<int> dest(na);
  std::iota(dest.begin(), dest.end(), 0);
  // nodes of each aggregate
  std::vector<int> off(na + 1, 0), nodes(nNo);
  for (int a = 0; a < nNo; a++) off[agg[a] + 1]++;
  for (int a = 0; a < na; a++) off[a + 1] += off[a];
  {
    std::vector<int> pos(off.begin(), off.end() - 1);
    for (int a = 0; a < nNo; a++) nodes[pos[agg[a]]++] = a;
  }
  std::map<int, int> edges;
  for (int a = 0; a < na; a++) {
    if (size[a] >= small) continue;
    edges.clear();
    for (int t = off[a]; t < off[a + 1]; t++) {
      const int v = nodes[t];
      for (int k = rp[v]; k < rp[v + 1]; k++) {
        const int b = agg[cp[k]];
        if (b != a && size[b] >= small) edges[b]++;
      }
    }
    int best = -1, bc = 0;
    for (const auto& e : edges)
      if (e.second > bc) best = e.first, bc = e.second;
    if (best >= 0) dest[a] = best;
  }
  std::vector<int> label(na, -1);
  nAgg = 0;
  for (int a = 0; a < nNo; a++) {
    const int d = dest[agg[a]];
    if (label[d] < 0) label[d] = nAgg++;
    agg[a] = label[d];
  }
}

template <typename T>
static int upload(T** d, const std::vector<T>& h)
{
  int rc = dev_alloc(d, h.size());
  if (rc) return rc;
  if (!h.empty()) SVG_HIP(hipMemcpy(*d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice));
  return SVG_OK;
}

// sysNodeOff (nsys + 1 node offsets, or null for one system): the systems of a
// block-diagonal union; each gets its own dense block.
static int coarse_build(const svg_lhs* L, int target, const int* sysNodeOff, int nsys, CoarseLevel** out)
{
  std::vector<int> agg;
  int nAgg = 0;
  aggregate(L, target, agg, nAgg);
  const int nNo = L->nNo, nnz = L->nnz;
  // the blocks: aggregates are numbered by first node and never cross systems
  std::vector<int> aggOff;
  if (sysNodeOff && nsys > 0) {
    aggOff.resize(nsys + 1);
    for (int q = 0; q < nsys; q++) aggOff[q] = agg[sysNodeOff[q]];
    aggOff[nsys] = nAgg;
    for (int q = 0; q < nsys; q++)
      for (int a = sysNodeOff[q]; a < sysNodeOff[q + 1]; a++)
        if (agg[a] < aggOff[q] || agg[a] >= aggOff[q + 1])
          return set_error(SVG_ERR_ARG, "two-level preconditioner: an aggregate crosses systems (node " +
                                            std::to_string(a) + "); the systems must not share block-pattern entries");
  } else {
    aggOff = {0, nAgg};
  }
  const int nblk = (int)aggOff.size() - 1;
  std::vector<int> blkN(nblk), aggBlk(nAgg);
  std::vector<long long> blkOff(nblk + 1, 0), panOff(nblk + 1, 0);
  int maxN = 0;
  for (int q = 0; q < nblk; q++) {
    const int nq = aggOff[q + 1] - aggOff[q];
    if (nq > COARSE_MAX)
      return set_error(SVG_ERR_ARG, "two-level preconditioner: " + std::to_string(nq) +
                                        " aggregates exceed the dense coarse limit of " + std::to_string(COARSE_MAX) +
                                        "; use larger aggregates");
    blkN[q] = (nq + GJB - 1) / GJB * GJB;
    maxN = std::max(maxN, blkN[q]);
    blkOff[q + 1] = blkOff[q] + (long long)blkN[q] * blkN[q];
    panOff[q + 1] = panOff[q] + (long long)blkN[q] * GJB;
    for (int a = aggOff[q]; a < aggOff[q + 1]; a++) aggBlk[a] = q;
  }
  const auto& rp = L->rowPtr;
  const auto& cp = L->colPtr;
  // restriction lists: nodes of each aggregate, ascending
  std::vector<int> aggPtr(nAgg + 1, 0), aggNodes(nNo);
  for (int a = 0; a < nNo; a++) aggPtr[agg[a] + 1]++;
  for (int a = 0; a < nAgg; a++) aggPtr[a + 1] += aggPtr[a];
  {
    std::vector<int> pos(aggPtr.begin(), aggPtr.end() - 1);
    for (int a = 0; a < nNo; a++) aggNodes[pos[agg[a]]++] = a;
  }
  // A P: per row, the distinct aggregates of its column nodes (ascending);
  // apOf[k] = the entry of block k
  std::vector<int> apPtr(nNo + 1, 0), apAgg, apOf(nnz), row;
  apAgg.reserve((size_t)nNo * 4);
  for (int i = 0; i < nNo; i++) {
    row.clear();
    for (int k = rp[i]; k < rp[i + 1]; k++) row.push_back(agg[cp[k]]);
    std::sort(row.begin(), row.end());
    row.erase(std::unique(row.begin(), row.end()), row.end());
    for (int k = rp[i]; k < rp[i + 1]; k++)
      apOf[k] = apPtr[i] + (int)(std::lower_bound(row.begin(), row.end(), agg[cp[k]]) - row.begin());
    apAgg.insert(apAgg.end(), row.begin(), row.end());
    apPtr[i + 1] = (int)apAgg.size();
  }
  const int nAP = (int)apAgg.size();
  // dense Ac entries: the AP entries ordered by (aggregate of the row, column
  // aggregate, entry index); one segment per coarse entry
  std::vector<int> apRowAgg(nAP);
  for (int i = 0; i < nNo; i++)
    for (int e = apPtr[i]; e < apPtr[i + 1]; e++) apRowAgg[e] = agg[i];
  std::vector<int> perm(nAP);
  std::iota(perm.begin(), perm.end(), 0);
  std::stable_sort(perm.begin(), perm.end(), [&](int x, int y) {
    return apRowAgg[x] != apRowAgg[y] ? apRowAgg[x] < apRowAgg[y] : apAgg[x] < apAgg[y];
  });
  std::vector<int> segPtr;
  std::vector<long long> segIdx;
  for (int q = 0; q < nAP; q++) {
    const int e = perm[q];
    if (q == 0 || apRowAgg[e] != apRowAgg[perm[q - 1]] || apAgg[e] != apAgg[perm[q - 1]]) {
      const int b = aggBlk[apRowAgg[e]], a0 = aggOff[b];
      if (aggBlk[apAgg[e]] != b)
        return set_error(SVG_ERR_ARG, "two-level preconditioner: a block-pattern entry couples two systems");
      segPtr.push_back(q);
      segIdx.push_back(blkOff[b] + (long long)(apRowAgg[e] - a0) * blkN[b] + (apAgg[e] - a0));
    }
  }
  segPtr.push_back(nAP);

  auto* c = new CoarseLevel;
  c->aggSize = target;
  c->nsys = (sysNodeOff && nsys > 0) ? nsys : 0;
  c->nAgg = nAgg;
  c->nAP = nAP;
  c->nSeg = (int)segIdx.size();
  c->nNo = nNo;
  c->nblk = nblk;
  c->maxN = maxN;
  c->total = (size_t)blkOff[nblk];
  c->panTotal = (size_t)panOff[nblk];
  c->aggOff = aggOff;
  c->blkN = blkN;
  int rc;
  if ((rc = upload(&c->d_agg, agg)) || (rc = upload(&c->d_aggPtr, aggPtr)) || (rc = upload(&c->d_aggNodes, aggNodes)) ||
      (rc = upload(&c->d_apPtr, apPtr)) || (rc = upload(&c->d_apAgg, apAgg)) || (rc = upload(&c->d_apOf, apOf)) ||
      (rc = upload(&c->d_segPtr, segPtr)) || (rc = upload(&c->d_segIdx, segIdx)) ||
      (rc = upload(&c->d_segPerm, perm)) || (rc = upload(&c->d_aggOff, aggOff)) || (rc = upload(&c->d_blkN, blkN)) ||
      (rc = upload(&c->d_aggBlk, aggBlk)) || (rc = upload(&c->d_blkOff, blkOff)) ||
      (rc = upload(&c->d_panOff, panOff)) || (rc = dev_alloc(&c->d_apVal, (size_t)4 * nAP)) ||
      (rc = dev_alloc(&c->d_apPP, (size_t)nAP)) || (rc = dev_alloc(&c->d_dinv, (size_t)16 * nNo)) ||
      (rc = dev_alloc(&c->d_Ac, c->total)) || (rc = dev_alloc(&c->d_Ai, c->total)) ||
      (rc = dev_alloc(&c->d_piv, (size_t)2 * GJB * GJB * nblk)) || (rc = dev_alloc(&c->d_cp, c->panTotal)) ||
      (rc = dev_alloc(&c->d_rp, c->panTotal)) || (rc = dev_alloc(&c->d_rc, (size_t)nAgg)) ||
      (rc = dev_alloc(&c->d_zc, (size_t)nAgg)) || (rc = dev_alloc(&c->d_flag, 1))) {
    coarse_free(c);
    return rc;
  }
  hipError_t e = hipHostMalloc((void**)&c->h_flag, sizeof(int), hipHostMallocDefault);
  if (e != hipSuccess) {
    coarse_free(c);
    return hip_check(e, "hipHostMalloc");
  }
  c->gen = ++g_coarse_gen;
  *out = c;
  return SVG_OK;
}

// ------------------------------------------------ per-system set-up kernels --
// A P values of row i: apVal(d, e) = sum over the row's blocks k of entry e, in
// CSR order, of A_k(d, 3); apPP(e) the same sum of A_k(3, 3) in FP64.
__global__ __launch_bounds__(256) void k_cl_ap(int nNo, const int* __restrict__ rowPtr, const int* __restrict__ apPtr,
                                               const int* __restrict__ apOf, const double* __restrict__ Val,
                                               float* __restrict__ apVal, double* __restrict__ apPP)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nNo) return;
  const int k0 = rowPtr[i], k1 = rowPtr[i + 1];
  for (int e = apPtr[i]; e < apPtr[i + 1]; e++) {
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    for (int k = k0; k < k1; k++) {
      if (apOf[k] != e) continue;
      const double* b = Val + (size_t)16 * k;
      s0 += b[3];
      s1 += b[7];
      s2 += b[11];
      s3 += b[15];
    }
    *reinterpret_cast<float4*>(apVal + (size_t)4 * e) = make_float4((float)s0, (float)s1, (float)s2, (float)s3);
    apPP[e] = s3;
  }
}

// Inverse of node i's 4x4 diagonal block (Gauss-Jordan with partial pivoting in
// registers); a zero diagonal entry (a Dirichlet row and column, zeroed by the
// scaling) becomes 1 first; a singular block gives the identity.
__global__ __launch_bounds__(256) void k_cl_dinv(int nNo, const int* __restrict__ diagPtr,
                                                 const double* __restrict__ Val, float* __restrict__ dinv)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nNo) return;
  const double* b = Val + (size_t)16 * diagPtr[i];
  double m[4][8];
#pragma unroll
  for (int r = 0; r < 4; r++) {
#pragma unroll
    for (int c = 0; c < 4; c++) {
      m[r][c] = b[4 * r + c];
      m[r][4 + c] = r == c ? 1.0 : 0.0;
    }
    if (m[r][r] == 0.0) m[r][r] = 1.0;
  }
  bool ok = true;
#pragma unroll
  for (int p = 0; p < 4; p++) {
    int piv = p;
    double best = fabs(m[p][p]);
#pragma unroll
    for (int r = p + 1; r < 4; r++)
      if (fabs(m[r][p]) > best) best = fabs(m[r][p]), piv = r;
    if (!(best > 0.0)) ok = false;
#pragma unroll
    for (int r = 0; r < 4; r++)
      if (r == piv && piv != p) {
#pragma unroll
        for (int c = 0; c < 8; c++) {
          const double t = m[p][c];
          m[p][c] = m[r][c];
          m[r][c] = t;
        }
      }
    const double inv = best > 0.0 ? 1.0 / m[p][p] : 0.0;
#pragma unroll
    for (int c = 0; c < 8; c++) m[p][c] *= inv;
#pragma unroll
    for (int r = 0; r < 4; r++) {
      if (r == p) continue;
      const double f = m[r][p];
#pragma unroll
      for (int c = 0; c < 8; c++) m[r][c] -= f * m[p][c];
    }
  }
  float* o = dinv + (size_t)16 * i;
#pragma unroll
  for (int r = 0; r < 4; r++)
    *reinterpret_cast<float4*>(o + 4 * r) =
        ok ? make_float4((float)m[r][4], (float)m[r][5], (float)m[r][6], (float)m[r][7])
           : make_float4(r == 0 ? 1.f : 0.f, r == 1 ? 1.f : 0.f, r == 2 ? 1.f : 0.f, r == 3 ? 1.f : 0.f);
}

// Block blockIdx.y of Ac = 0 except the identity on its padding rows n..N-1.
__global__ __launch_bounds__(256) void k_cl_acinit(const int* __restrict__ aggOff, const int* __restrict__ blkN,
                                                   const long long* __restrict__ blkOff, double* __restrict__ Ac,
                                                   int* __restrict__ flag)
{
  const int b = blockIdx.y;
  const int N = blkN[b], n = aggOff[b + 1] - aggOff[b];
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b == 0 && q == 0) *flag = 0;
  if (q >= (size_t)N * N) return;
  const int r = (int)(q / N), c = (int)(q - (size_t)r * N);
  Ac[blkOff[b] + q] = (r == c && r >= n) ? 1.0 : 0.0;
}

// One coarse entry per thread: the sum of its AP entries in segment order.
__global__ __launch_bounds__(256) void k_cl_ac(int nSeg, const int* __restrict__ segPtr,
                                               const long long* __restrict__ segIdx, const int* __restrict__ perm,
                                               const double* __restrict__ apPP, double* __restrict__ Ac)
{
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= nSeg) return;
  double s = 0.0;
  for (int q = segPtr[g]; q < segPtr[g + 1]; q++) s += apPP[perm[q]];
  Ac[segIdx[g]] = s;
}

// --------------------------------------------- blocked Gauss-Jordan inverse --
// In-place inversion of each row-major N x N block A (block blockIdx.z: N =
// blkN[b], at blkOff[b]), one GJB-wide block column K per step (blocks with
// fewer steps skip the later ones):
//   piv = A[K,K]^-1                       (k_gj_pivot, one workgroup)
//   rp  = piv A[K,:],  cp = A[:,K]         (k_gj_panels)
//   A[I,J] -= cp[I] rp[J]  (I, J != K);  A[I,K] = -cp[I] piv;  A[K,J] = rp[J];  A[K,K] = piv
// No pivoting: the matrix is definite (every pivot a leading principal minor's
// ratio, nonzero); a zero or non-finite pivot raises *flag.

// The pivot block by unpivoted Gauss-Jordan in registers: lane (w, i) of the
// 256 holds row i, columns 16w..16w+15 (v); step p publishes row p and column
// p through LDS, double buffered by the step's parity (a thread writing step
// p + 1's buffer has passed step p's barrier, so every thread has read that
// buffer at step p - 1), one barrier per step. Block column K of the
// elimination: a zero or non-finite pivot raises *flag.
__device__ __forceinline__ void gj_pivot_regs(double (&v)[16], int K, double (&rowb)[2][GJB],
                                              double (&colb)[2][GJB], int* __restrict__ flag)
{
  static_assert(GJB == 64, "gj_pivot_regs: one row per lane");
  const int t = threadIdx.x, i = t & 63, w = t >> 6, j0 = 16 * w;
#pragma unroll
  for (int p = 0; p < GJB; p++) {
    const int buf = p & 1;
    if (i == p) {
#pragma unroll
      for (int jj = 0; jj < 16; jj++) rowb[buf][j0 + jj] = v[jj];
    }
    if (w == p / 16) colb[buf][i] = v[p % 16];
    __syncthreads();
    const double pv = rowb[buf][p];
    // 1 / pv: the hardware reciprocal refined by two Newton steps (a shorter
    // dependent chain than the correctly rounded division; this is the
    // elimination's critical path)
    double inv = __builtin_amdgcn_rcp(pv);
    inv = fma(inv, fma(-pv, inv, 1.0), inv);
    inv = fma(inv, fma(-pv, inv, 1.0), inv);
    const double sip = colb[buf][i];
    // the general entry, then column p (one wavefront, a fixed register) and
    // row p (one lane) overwritten: no per-entry selects
#pragma unroll
    for (int jj = 0; jj < 16; jj++) v[jj] = v[jj] - sip * (rowb[buf][j0 + jj] * inv);
    if (w == p / 16) v[p % 16] = -sip * inv;
    if (i == p) {
#pragma unroll
      for (int jj = 0; jj < 16; jj++) v[jj] = (j0 + jj == p) ? inv : rowb[buf][j0 + jj] * inv;
    }
    if (t == 0 && !(fabs(pv) > 0.0 && isfinite(pv))) *flag = 1 + K * GJB + p;
  }
}

// Pivot blocks are double buffered by the step's parity: step K's panels and
// update read slot K & 1 while the update's first workgroup writes step K + 1's.
__device__ __forceinline__ double* gj_piv_slot(double* piv0, int b, int K)
{
  return piv0 + (size_t)GJB * GJB * (2 * b + (K & 1));
}

// The first step's pivot block (later steps' are inverted by the update).
__global__ __launch_bounds__(256) void k_gj_pivot(const double* __restrict__ A0, const int* __restrict__ blkN,
                                                  const long long* __restrict__ blkOff, int K,
                                                  double* __restrict__ piv0, int* __restrict__ flag)
{
  __shared__ double rowb[2][GJB], colb[2][GJB];
  const int b = blockIdx.z, N = blkN[b];
  if (K >= N / GJB) return;
  const double* A = A0 + blkOff[b] + (size_t)K * GJB * N + (size_t)K * GJB;
  double* piv = gj_piv_slot(piv0, b, K);
  const int t = threadIdx.x, i = t & 63, j0 = 16 * (t >> 6);
  double v[16];
#pragma unroll
  for (int jj = 0; jj < 16; jj++) v[jj] = A[(size_t)i * N + j0 + jj];
  gj_pivot_regs(v, K, rowb, colb, flag);
#pragma unroll
  for (int jj = 0; jj < 16; jj++) piv[i * GJB + j0 + jj] = v[jj];
}

typedef double gj_v4d __attribute__((ext_vector_type(4)));

// C (64 x 64) = L R on the FP64 matrix cores, L(r, k) = Lp[r * ls + k * lk_],
// R(k, c) = Rp[k * rs + c]: wavefront w takes the 32 x 32 quarter (w / 2,
// w % 2) as 2 x 2 16 x 16 blocks over 16 steps of k = 4 (v_mfma_f64_16x16x4:
// lane l holds L[l % 16][l / 16], R[l / 16][l % 16] and C[4 r + l / 16][l % 16],
// r < 4); acc[u][v][r] is C(rb + 16 u + 4 r + l / 16, cb + 16 v + l % 16).
__device__ __forceinline__ void gj_mfma64(const double* __restrict__ Lp, size_t ls, size_t lk_,
                                          const double* __restrict__ Rp, size_t rs, gj_v4d (&acc)[2][2])
{
  const int t = threadIdx.x, w = t >> 6, l = t & 63, lr = l & 15, lk = l >> 4;
  const int rb = 32 * (w >> 1), cb = 32 * (w & 1);
  const double* ap = Lp + (size_t)(rb + lr) * ls + (size_t)lk * lk_;
  const double* bp = Rp + (size_t)lk * rs + cb + lr;
#pragma unroll
  for (int u = 0; u < 2; u++)
#pragma unroll
    for (int v = 0; v < 2; v++) acc[u][v] = gj_v4d{0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int ks = 0; ks < GJB / 4; ks++) {
    const double a0 = ap[(size_t)4 * ks * lk_], a1 = ap[(size_t)4 * ks * lk_ + 16 * ls];
    const double b0 = bp[(size_t)4 * ks * rs], b1 = bp[(size_t)4 * ks * rs + 16];
    acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
    acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
    acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
    acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
  }
}

// Workgroup J: the column panel block cpT[:, J rows] = A[J rows, K]^T (the
// panel stored transposed, GJB x N row-major, so that the update reads it
// along rows) and rp[:, J] = piv A[K, J] (J != K).
__global__ __launch_bounds__(256) void k_gj_panels(const double* __restrict__ A0, const int* __restrict__ blkN,
                                                   const long long* __restrict__ blkOff,
                                                   const long long* __restrict__ panOff, int K,
                                                   const double* __restrict__ piv0, double* __restrict__ rp0,
                                                   double* __restrict__ cp0)
{
  const int b = blockIdx.z, N = blkN[b], nb = N / GJB;
  if (K >= nb || (int)blockIdx.x >= nb) return;
  const double* A = A0 + blkOff[b];
  const double* piv = gj_piv_slot(const_cast<double*>(piv0), b, K);
  double* rp = rp0 + panOff[b];
  double* cpT = cp0 + panOff[b];
  const int J = blockIdx.x, t = threadIdx.x;
  for (int e = t; e < GJB * GJB; e += 256) {
    const int k = e / GJB, r = e % GJB;
    cpT[(size_t)k * N + J * GJB + r] = A[(size_t)(J * GJB + r) * N + K * GJB + k];
  }
  if (J == K) return;
  gj_v4d acc[2][2];
  gj_mfma64(piv, GJB, 1, A + (size_t)K * GJB * N + (size_t)J * GJB, N, acc);
  const int w = t >> 6, l = t & 63, lr = l & 15, lk = l >> 4, rb = 32 * (w >> 1), cb = 32 * (w & 1);
#pragma unroll
  for (int u = 0; u < 2; u++)
#pragma unroll
    for (int v = 0; v < 2; v++)
#pragma unroll
      for (int r = 0; r < 4; r++)
        rp[(size_t)(rb + 16 * u + 4 * r + lk) * N + J * GJB + cb + 16 * v + lr] = acc[u][v][r];
}

// Tile (I, J) of the update, the tiles rotated so that workgroup (0, 0) takes
// (K + 1, K + 1): it finishes its tile first and then inverts it as step
// K + 1's pivot block while the other tiles are updated. The product cp[I]
// rp[J] (or cp[I] piv for J = K) runs on the FP64 matrix cores, both operands
// read straight from the panels (16 consecutive entries per 16 lanes).
__global__ __launch_bounds__(256) void k_gj_update(double* __restrict__ A0, const int* __restrict__ blkN,
                                                   const long long* __restrict__ blkOff,
                                                   const long long* __restrict__ panOff, int K,
                                                   double* __restrict__ piv0, const double* __restrict__ rp0,
                                                   const double* __restrict__ cp0, int* __restrict__ flag)
{
  __shared__ double T[GJB][GJB + 1];
  __shared__ double rowb[2][GJB], colb[2][GJB];
  const int b = blockIdx.z, N = blkN[b], nb = N / GJB;
  if (K >= nb || (int)blockIdx.x >= nb || (int)blockIdx.y >= nb) return;
  double* A = A0 + blkOff[b];
  const double* piv = gj_piv_slot(piv0, b, K);
  const double* rp = rp0 + panOff[b];
  const double* cpT = cp0 + panOff[b];
  const int J = (blockIdx.x + K + 1) % nb, I = (blockIdx.y + K + 1) % nb, t = threadIdx.x;
  double* At = A + (size_t)I * GJB * N + (size_t)J * GJB;
  if (I == K) {
    const int r0 = 4 * (t / 16), c0 = 4 * (t % 16);
#pragma unroll
    for (int u = 0; u < 4; u++)
#pragma unroll
      for (int v = 0; v < 4; v++)
        At[(size_t)(r0 + u) * N + c0 + v] =
            (J == K) ? piv[(r0 + u) * GJB + c0 + v] : rp[(size_t)(r0 + u) * N + J * GJB + c0 + v];
    return;
  }
  const bool diag = (J == K);
  gj_v4d acc[2][2];
  gj_mfma64(cpT + (size_t)I * GJB, 1, N, diag ? piv : rp + (size_t)J * GJB, diag ? GJB : N, acc);
  const int w = t >> 6, l = t & 63, lr = l & 15, lk = l >> 4, rb = 32 * (w >> 1), cb = 32 * (w & 1);
  const bool next = (I == K + 1 && J == K + 1);
#pragma unroll
  for (int u = 0; u < 2; u++)
#pragma unroll
    for (int v = 0; v < 2; v++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int row = rb + 16 * u + 4 * r + lk, col = cb + 16 * v + lr;
        double* a = At + (size_t)row * N + col;
        const double x = diag ? -acc[u][v][r] : *a - acc[u][v][r];
        *a = x;
        if (next) T[row][col] = x;
      }
  if (!next) return;
  __syncthreads();
  const int i = t & 63, j0 = 16 * (t >> 6);
  double vr[16];
#pragma unroll
  for (int jj = 0; jj < 16; jj++) vr[jj] = T[i][j0 + jj];
  gj_pivot_regs(vr, K + 1, rowb, colb, flag);
  double* pn = gj_piv_slot(piv0, b, K + 1);
#pragma unroll
  for (int jj = 0; jj < 16; jj++) pn[i * GJB + j0 + jj] = vr[jj];
}

__global__ __launch_bounds__(256) void k_cl_d2f(size_t n, const double* __restrict__ a, float* __restrict__ b)
{
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
    b[q] = (float)a[q];
}

// ----------------------------------------------------- application kernels --
// rc(a) = sum of v(3, i) over the nodes i of aggregate a (one wavefront per
// aggregate: lanes at stride 64 in node order, then a fixed butterfly).
// The batched solve's gate (st non-null): the work of a system that does not
// iterate in this cycle (converged before it, or during it) is skipped; the
// apply gives a system that fell back to the diagonal preconditioner z = v.
__device__ __forceinline__ bool sys_idle(const GmState* st, int sys)
{
  const GmState& S = st[sys];
  return !S.act || S.done;
}

template <typename T>
__global__ __launch_bounds__(256) void k_cl_restrict(int nAgg, const int* __restrict__ aggPtr,
                                                     const int* __restrict__ aggNodes, const T* __restrict__ v,
                                                     float* __restrict__ rc, const int* __restrict__ skip,
                                                     const GmState* __restrict__ st, const int* __restrict__ aggBlk)
{
  if (skip && *skip) return;
  const int a = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (a >= nAgg) return;
  if (st && sys_idle(st, aggBlk[a])) return;
  double s = 0.0;
  for (int t = aggPtr[a] + lane; t < aggPtr[a + 1]; t += 64) s += (double)v[(size_t)4 * aggNodes[t] + 3];
  s = wave_sum(s);
  if (lane == 0) rc[a] = (float)s;
}

// The streams read once per application (the coarse inverse, the
// block-Jacobi inverses) use non-temporal loads, leaving the memory-side cache
// to the block matrix the Arnoldi step's SpMV reads again.
typedef float cl_f4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ cl_f4 nt_f4(const float4* p)
{
  return __builtin_nontemporal_load(reinterpret_cast<const cl_f4*>(p));
}

// zc = Ai rc block by block: workgroup (x, y) takes rows 4x..4x+3 of block y
// (one wavefront per row; row r of the block's N x N FP32 inverse), with the
// block's rc staged in LDS and zero on the padding columns. Each lane adds its
// float4 chunks q = lane, lane + 64, ... in order, then a fixed butterfly. The
// row's first 16 chunks per lane (N <= 4096) are loaded before the staging,
// so that their latency overlaps it; longer rows continue eight at a time.
__global__ __launch_bounds__(256) void k_cl_gemv(const int* __restrict__ aggOff, const int* __restrict__ blkN,
                                                 const long long* __restrict__ blkOff, const float* __restrict__ Ai,
                                                 const float* __restrict__ rc, float* __restrict__ zc,
                                                 const int* __restrict__ skip, const GmState* __restrict__ st)
{
  extern __shared__ float xs[];
  if (skip && *skip) return;
  const int bk = blockIdx.y, a0 = aggOff[bk], n = aggOff[bk + 1] - a0, N = blkN[bk];
  if ((int)blockIdx.x * 4 >= n) return;
  if (st && sys_idle(st, bk)) return;  // block bk is system bk's
  const int r = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  const bool live = r < n;
  const float4* row = reinterpret_cast<const float4*>(Ai + blkOff[bk] + (size_t)(live ? r : 0) * N);
  const int nq = N / 4;
  cl_f4 a[16];
#pragma unroll
  for (int u = 0; u < 16; u++) a[u] = (live && lane + 64 * u < nq) ? nt_f4(row + lane + 64 * u) : cl_f4{0.f, 0.f, 0.f, 0.f};
  for (int c0 = threadIdx.x; c0 < N; c0 += 256 * 8) {
    float xv[8];
#pragma unroll
    for (int u = 0; u < 8; u++) xv[u] = c0 + 256 * u < n ? rc[a0 + c0 + 256 * u] : 0.f;
#pragma unroll
    for (int u = 0; u < 8; u++)
      if (c0 + 256 * u < N) xs[c0 + 256 * u] = xv[u];
  }
  __syncthreads();
  if (!live) return;
  const float4* x = reinterpret_cast<const float4*>(xs);
  float s = 0.f;
#pragma unroll
  for (int u = 0; u < 16; u++)
    if (lane + 64 * u < nq) {
      const float4 bb = x[lane + 64 * u];
      s += a[u].x * bb.x + a[u].y * bb.y + a[u].z * bb.z + a[u].w * bb.w;
    }
  for (int q0 = lane + 1024; q0 < nq; q0 += 512) {
    cl_f4 c[8];
#pragma unroll
    for (int u = 0; u < 8; u++) c[u] = (q0 + 64 * u < nq) ? nt_f4(row + q0 + 64 * u) : cl_f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int u = 0; u < 8; u++)
      if (q0 + 64 * u < nq) {
        const float4 bb = x[q0 + 64 * u];
        s += c[u].x * bb.x + c[u].y * bb.y + c[u].z * bb.z + c[u].w * bb.w;
      }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if (lane == 0) zc[a0 + r] = s;
}

// Lane K of each quad's value, broadcast to the quad (DPP quad_perm).
template <int K>
__device__ __forceinline__ float quad_bcast_f32(float v)
{
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), K * 0x55, 0xF, 0xF, false));
}

// With r2_i = v_i - sum_e AP(:, e) zc(apAgg(e)) (the coarse-corrected residual):
//   mode 0  z_i = B_i r2_i + (0, 0, 0, zc(agg(i)))   (the two-level preconditioner)
//   mode 1  z_i = r2_i
//   mode 2  z_i = B_i r2_i
// written in FP64 and/or FP32; a quad of lanes per node (lane d computes
// component d of r2 and of z), so that four times as many wavefronts cover the
// loads' latency.
template <typename T>
__global__ __launch_bounds__(256) void k_cl_apply(int nNo, const T* __restrict__ v, const int* __restrict__ apPtr,
                                                  const int* __restrict__ apAgg, const float* __restrict__ apVal,
                                                  const float* __restrict__ zc, const int* __restrict__ agg,
                                                  const float* __restrict__ dinv, double* __restrict__ z64,
                                                  float* __restrict__ z32, const int* __restrict__ skip, int mode,
                                                  const GmState* __restrict__ st, const int* __restrict__ sysOfNode)
{
  if (skip && *skip) return;
  // a quad of lanes per node, lane d for dof d (whole quads leave together)
  const size_t gid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int i = (int)(gid >> 2), d = (int)(gid & 3);
  if (i >= nNo) return;
  const size_t k = (size_t)4 * i + d;
  if (st) {
    const GmState& S = st[sysOfNode[i]];
    if (!S.act || S.done) return;
    if (S.fallback) {
      // a system that fell back to the diagonal preconditioner: its flexible
      // basis vector is the plain one, z_i = v_i
      if (z32) z32[k] = (float)v[k];
      if (z64) z64[k] = (double)v[k];
      return;
    }
  }
  float rd = (float)v[k];
  // row d of the block-Jacobi inverse and the own aggregate's entry do not
  // depend on the loop: their loads are issued before it
  cl_f4 D = {0.f, 0.f, 0.f, 0.f};
  if (mode != 1) D = nt_f4(reinterpret_cast<const float4*>(dinv + (size_t)16 * i) + d);
  const float zown = (mode == 0 && d == 3) ? zc[agg[i]] : 0.f;
  // component d of r2 over the AP entries in order, their loads issued four at a time
  const int eb = apPtr[i], ee = apPtr[i + 1];
  for (int e = eb; e < ee; e += 4) {
    int g[4];
    float a[4], c[4];
#pragma unroll
    for (int u = 0; u < 4; u++)
      if (e + u < ee) {
        g[u] = apAgg[e + u];
        a[u] = apVal[(size_t)4 * (e + u) + d];
      }
#pragma unroll
    for (int u = 0; u < 4; u++)
      if (e + u < ee) c[u] = zc[g[u]];
#pragma unroll
    for (int u = 0; u < 4; u++)
      if (e + u < ee) rd -= a[u] * c[u];
  }
  float z = rd;
  if (mode != 1) {
    // the quad's four components of r2 by DPP
    const float r0 = quad_bcast_f32<0>(rd), r1 = quad_bcast_f32<1>(rd), r2c = quad_bcast_f32<2>(rd),
                r3 = quad_bcast_f32<3>(rd);
    z = D.x * r0 + D.y * r1 + D.z * r2c + D.w * r3;
    if (mode == 0 && d == 3) z += zown;
  }
  if (z32) z32[k] = z;
  if (z64) z64[k] = (double)z;
}

// The same with one lane per node (the batched solve's gated form: most of a
// batch's nodes belong to meshes that stopped iterating, whose lanes leave at
// once, so fewer lanes per node launch fewer idle workgroups).
template <typename T>
__global__ __launch_bounds__(256) void k_cl_apply_node(int nNo, const T* __restrict__ v, const int* __restrict__ apPtr,
                                                  const int* __restrict__ apAgg, const float* __restrict__ apVal,
                                                  const float* __restrict__ zc, const int* __restrict__ agg,
                                                  const float* __restrict__ dinv, double* __restrict__ z64,
                                                  float* __restrict__ z32, const int* __restrict__ skip, int mode,
                                                  const GmState* __restrict__ st, const int* __restrict__ sysOfNode)
{
  if (skip && *skip) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nNo) return;
  if (st) {
    const GmState& S = st[sysOfNode[i]];
    if (!S.act || S.done) return;
    if (S.fallback) {
      // a system that fell back to the diagonal preconditioner: its flexible
      // basis vector is the plain one, z_i = v_i
#pragma unroll
      for (int d = 0; d < 4; d++) {
        if (z32) z32[(size_t)4 * i + d] = (float)v[(size_t)4 * i + d];
        if (z64) z64[(size_t)4 * i + d] = (double)v[(size_t)4 * i + d];
      }
      return;
    }
  }
  float r[4];
#pragma unroll
  for (int d = 0; d < 4; d++) r[d] = (float)v[(size_t)4 * i + d];
  // the block-Jacobi inverse and the own aggregate's entry do not depend on
  // the loop: their loads are issued before it
  cl_f4 D[4];
  const float4* Dp = reinterpret_cast<const float4*>(dinv + (size_t)16 * i);
  if (mode != 1) {
#pragma unroll
    for (int d = 0; d < 4; d++) D[d] = nt_f4(Dp + d);
  }
  const float zown = mode == 0 ? zc[agg[i]] : 0.f;
  // the AP entries in order, their loads issued four at a time
  const int eb = apPtr[i], ee = apPtr[i + 1];
  for (int e = eb; e < ee; e += 4) {
    int g[4];
    float4 a[4];
    float c[4];
#pragma unroll
    for (int u = 0; u < 4; u++)
      if (e + u < ee) {
        g[u] = apAgg[e + u];
        a[u] = *reinterpret_cast<const float4*>(apVal + (size_t)4 * (e + u));
      }
#pragma unroll
    for (int u = 0; u < 4; u++)
      if (e + u < ee) c[u] = zc[g[u]];
#pragma unroll
    for (int u = 0; u < 4; u++)
      if (e + u < ee) {
        r[0] -= a[u].x * c[u];
        r[1] -= a[u].y * c[u];
        r[2] -= a[u].z * c[u];
        r[3] -= a[u].w * c[u];
      }
  }
  float z[4];
  if (mode == 1) {
#pragma unroll
    for (int d = 0; d < 4; d++) z[d] = r[d];
  } else {
#pragma unroll
    for (int d = 0; d < 4; d++) {
      const cl_f4 m = D[d];
      z[d] = m.x * r[0] + m.y * r[1] + m.z * r[2] + m.w * r[3];
    }
    if (mode == 0) z[3] += zown;
  }
  if (z32) *reinterpret_cast<float4*>(z32 + (size_t)4 * i) = make_float4(z[0], z[1], z[2], z[3]);
  if (z64) {
    double2* o = reinterpret_cast<double2*>(z64 + (size_t)4 * i);
    o[0] = make_double2(z[0], z[1]);
    o[1] = make_double2(z[2], z[3]);
  }
}

// z(3, i) += zc(agg(i)) in FP64 and FP32 (the coarse correction added to
// another smoother's output).
__global__ __launch_bounds__(256) void k_cl_addc(int nNo, const float* __restrict__ zc, const int* __restrict__ agg,
                                                 double* __restrict__ z64, float* __restrict__ z32,
                                                 const int* __restrict__ skip)
{
  if (skip && *skip) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nNo) return;
  const float c = zc[agg[i]];
  if (z64) z64[(size_t)4 * i + 3] += (double)c;
  if (z32) z32[(size_t)4 * i + 3] += c;
}

// ------------------------------------------------------------- interfaces --
// The coarse level of L for aggregates of about aggSize nodes, one dense block
// per system of sysNodeOff (null: one system); built once and kept on L.
int coarse_setup(svg_lhs* L, int aggSize, const int* sysNodeOff, int nsys)
{
  if (aggSize < 2) return set_error(SVG_ERR_ARG, "two-level preconditioner: aggregates need at least 2 nodes");
  CoarseLevel*& c = L->coarse;
  const int ns = (sysNodeOff && nsys > 0) ? nsys : 0;
  if (c && (c->aggSize != aggSize || c->nNo != L->nNo || c->nsys != ns)) coarse_free(c);
  if (!c) return coarse_build(L, aggSize, sysNodeOff, nsys, &c);
  return SVG_OK;
}

// This system's A P, block-Jacobi inverses, Ac and its inverse from the scaled
// Val: launches only (no host synchronisation, no allocation; capturable).
void coarse_enqueue(svg_lhs* L, const double* Val, hipStream_t s, bool invert)
{
  CoarseLevel* c = L->coarse;
  const int nNo = L->nNo, nblk = c->nblk, nb = c->maxN / GJB;
  launch("cl_ap", k_cl_ap, dim3(blocks_for(nNo, 256)), dim3(256), 0, s, nNo, (const int*)L->d_rowPtr,
         (const int*)c->d_apPtr, (const int*)c->d_apOf, Val, c->d_apVal, c->d_apPP);
  launch("cl_dinv", k_cl_dinv, dim3(blocks_for(nNo, 256)), dim3(256), 0, s, nNo, (const int*)L->d_diagPtr, Val,
         c->d_dinv);
  if (!invert) return;
  c->inverted = true;
  launch("cl_acinit", k_cl_acinit, dim3(blocks_for((size_t)c->maxN * c->maxN, 256), nblk), dim3(256), 0, s,
         (const int*)c->d_aggOff, (const int*)c->d_blkN, (const long long*)c->d_blkOff, c->d_Ac, c->d_flag);
  launch("cl_ac", k_cl_ac, dim3(blocks_for(c->nSeg, 256)), dim3(256), 0, s, c->nSeg, (const int*)c->d_segPtr,
         (const long long*)c->d_segIdx, (const int*)c->d_segPerm, (const double*)c->d_apPP, c->d_Ac);
  launch("gj_pivot", k_gj_pivot, dim3(1, 1, nblk), dim3(256), 0, s, (const double*)c->d_Ac, (const int*)c->d_blkN,
         (const long long*)c->d_blkOff, 0, c->d_piv, c->d_flag);
  for (int K = 0; K < nb; K++) {
    launch("gj_panels", k_gj_panels, dim3(nb, 1, nblk), dim3(256), 0, s, (const double*)c->d_Ac,
           (const int*)c->d_blkN, (const long long*)c->d_blkOff, (const long long*)c->d_panOff, K,
           (const double*)c->d_piv, c->d_rp, c->d_cp);
    launch("gj_update", k_gj_update, dim3(nb, nb, nblk), dim3(256), 0, s, c->d_Ac, (const int*)c->d_blkN,
           (const long long*)c->d_blkOff, (const long long*)c->d_panOff, K, c->d_piv, (const double*)c->d_rp,
           (const double*)c->d_cp, c->d_flag);
  }
  launch("cl_d2f", k_cl_d2f, dim3(2048), dim3(256), 0, s, c->total, (const double*)c->d_Ac, c->d_Ai);
}

// The elimination's pivot check (synchronises s).
int coarse_check(const svg_lhs* L, hipStream_t s)
{
  const CoarseLevel* c = L->coarse;
  SVG_HIP(hipGetLastError());
  SVG_HIP(hipMemcpyAsync(c->h_flag, c->d_flag, sizeof(int), hipMemcpyDeviceToHost, s));
  SVG_HIP(hipStreamSynchronize(s));
  if (*c->h_flag)
    return set_error(SVG_ERR_SINGULAR, "two-level preconditioner: zero pivot " + std::to_string(*c->h_flag - 1) +
                                           " in the coarse matrix");
  return SVG_OK;
}

int coarse_prepare_dev(svg_lhs* L, const double* Val, int aggSize, hipStream_t s, bool keep)
{
  int rc = coarse_setup(L, aggSize, nullptr, 0);
  if (rc) return rc;
  const bool invert = !(keep && L->coarse->inverted);
  coarse_enqueue(L, Val, s, invert);
  return invert ? coarse_check(L, s) : SVG_OK;
}

template <typename T>
static void coarse_apply_t(const svg_lhs* L, const T* v, double* z64, float* z32, const int* skip, hipStream_t s,
                           int mode = 0, const GmState* st = nullptr, const int* sysOfNode = nullptr)
{
  const CoarseLevel* c = L->coarse;
  const int nAgg = c->nAgg;
  launch("cl_restrict", k_cl_restrict<T>, dim3(blocks_for(nAgg, 4)), dim3(256), 0, s, nAgg,
         (const int*)c->d_aggPtr, (const int*)c->d_aggNodes, v, c->d_rc, skip, st, (const int*)c->d_aggBlk);
  launch("cl_gemv", k_cl_gemv, dim3(blocks_for(c->maxN, 4), c->nblk), dim3(256), sizeof(float) * c->maxN, s,
         (const int*)c->d_aggOff, (const int*)c->d_blkN, (const long long*)c->d_blkOff, (const float*)c->d_Ai,
         (const float*)c->d_rc, c->d_zc, skip, st);
  if (st)
    launch("cl_apply", k_cl_apply_node<T>, dim3(blocks_for(L->nNo, 256)), dim3(256), 0, s, L->nNo, v,
           (const int*)c->d_apPtr, (const int*)c->d_apAgg, (const float*)c->d_apVal, (const float*)c->d_zc,
           (const int*)c->d_agg, (const float*)c->d_dinv, z64, z32, skip, mode, st, sysOfNode);
  else
    launch("cl_apply", k_cl_apply<T>, dim3(blocks_for((size_t)4 * L->nNo, 256)), dim3(256), 0, s, L->nNo, v,
           (const int*)c->d_apPtr, (const int*)c->d_apAgg, (const float*)c->d_apVal, (const float*)c->d_zc,
           (const int*)c->d_agg, (const float*)c->d_dinv, z64, z32, skip, mode, st, sysOfNode);
}

void coarse_apply_batch_dev(const svg_lhs* L, const float* v, double* z64, float* z32, const GmState* st,
                            const int* sysOfNode, hipStream_t s)
{
  coarse_apply_t<float>(L, v, z64, z32, nullptr, s, 0, st, sysOfNode);
}

void coarse_apply_batch_dev(const svg_lhs* L, const double* v, double* z64, float* z32, const GmState* st,
                            const int* sysOfNode, hipStream_t s)
{
  coarse_apply_t<double>(L, v, z64, z32, nullptr, s, 0, st, sysOfNode);
}

void coarse_apply_dev(const svg_lhs* L, const float* v, double* z64, float* z32, const int* skip, hipStream_t s)
{
  coarse_apply_t<float>(L, v, z64, z32, skip, s);
}

void coarse_apply_dev(const svg_lhs* L, const double* v, double* z64, float* z32, const int* skip, hipStream_t s)
{
  coarse_apply_t<double>(L, v, z64, z32, skip, s);
}

void coarse_residual_dev(const svg_lhs* L, const float* v, double* r2, const int* skip, hipStream_t s)
{
  coarse_apply_t<float>(L, v, r2, nullptr, skip, s, 1);
}

void coarse_residual_dev(const svg_lhs* L, const double* v, double* r2, const int* skip, hipStream_t s)
{
  coarse_apply_t<double>(L, v, r2, nullptr, skip, s, 1);
}

void coarse_smooth_dev(const svg_lhs* L, const float* v, double* y, const int* skip, hipStream_t s)
{
  coarse_apply_t<float>(L, v, y, nullptr, skip, s, 2);
}

void coarse_smooth_dev(const svg_lhs* L, const double* v, double* y, const int* skip, hipStream_t s)
{
  coarse_apply_t<double>(L, v, y, nullptr, skip, s, 2);
}

void coarse_add_dev(const svg_lhs* L, double* z64, float* z32, const int* skip, hipStream_t s)
{
  const CoarseLevel* c = L->coarse;
  launch("cl_addc", k_cl_addc, dim3(blocks_for(L->nNo, 256)), dim3(256), 0, s, L->nNo, (const float*)c->d_zc,
         (const int*)c->d_agg, z64, z32, skip);
}

int coarse_num_aggregates(const svg_lhs* L) { return L->coarse ? L->coarse->nAgg : 0; }

}  // namespace svg

using namespace svg;

extern "C" {

int svg_coarse_prepare(svg_lhs* L, int aggSize, const double* Val, void* stream)
{
  if (!L || !Val) return set_error(SVG_ERR_ARG, "svg_coarse_prepare: null argument");
  return coarse_prepare_dev(L, Val, aggSize, (hipStream_t)stream);
}

int svg_coarse_apply(const svg_lhs* L, int mode, int nvec, const double* V, double* Z, void* stream)
{
  if (!L || !V || !Z || nvec < 0) return set_error(SVG_ERR_ARG, "svg_coarse_apply: bad argument");
  if (!L->coarse) return set_error(SVG_ERR_ARG, "svg_coarse_apply: svg_coarse_prepare has not run on this lhs");
  if (mode != SVG_COARSE_APPLY && mode != SVG_COARSE_RESIDUAL && mode != SVG_COARSE_SMOOTH)
    return set_error(SVG_ERR_ARG, "svg_coarse_apply: unknown mode");
  if (V == Z) return set_error(SVG_ERR_ARG, "svg_coarse_apply: V and Z must not overlap");
  hipStream_t s = (hipStream_t)stream;
  const size_t n = (size_t)4 * L->nNo;
  for (int b = 0; b < nvec; b++) {
    coarse_apply_t<double>(L, V + b * n, Z + b * n, nullptr, nullptr, s, mode);
  }
  SVG_HIP(hipGetLastError());
  SVG_HIP(hipStreamSynchronize(s));
  return SVG_OK;
}

int svg_coarse_info(const svg_lhs* L, int* nAgg, int* agg)
{
  if (!L || !nAgg) return set_error(SVG_ERR_ARG, "svg_coarse_info: null argument");
  const CoarseLevel* c = L->coarse;
  *nAgg = c ? c->nAgg : 0;
  if (c && agg) SVG_HIP(hipMemcpy(agg, c->d_agg, sizeof(int) * c->nNo, hipMemcpyDeviceToHost));
  return SVG_OK;
}

}  // extern "C"
