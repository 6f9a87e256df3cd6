// Mesh setup on the device: the block-CSR pattern of the FSILS lhs (what
// lhsa builds, lhsa.cpp:153-380, handed to fsils_lhs_create, lhs.cpp:57-120)
// and the element colouring, colour order and element -> block slot map the
// assembly kernel scatters through (do_assem's search, lhsa.cpp:123-134).
//
// Pattern: every element contributes its 16 node pairs and every node its
// diagonal pair as a 64-bit key row * nNo + col; one radix sort and one
// unique pass leave the keys in CSR order (rows ascending, columns strictly
// increasing in a row) — the same rowPtr/colPtr as the host build.
//
// Colouring: Jones-Plassmann. Each element has a fixed pseudo-random
// priority; in each round an uncoloured element whose priority is the
// largest among the uncoloured elements at all four of its nodes (per-node
// atomicMax) takes the smallest colour not yet used at its nodes (per-node
// 256-bit masks). Winners of a round share no node, so no two elements of a
// colour share a node, and the result is the sequential greedy colouring in
// decreasing priority order. The colour order is a stable radix sort of the
// element ids by colour.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstdint>
#include <cstring>

#include "svg_device.h"

namespace svg {

namespace {

constexpr int CW = 4;  // 64-bit words of a node's used-colour mask (256 colours)

__global__ __launch_bounds__(256) void k_check_ien(long long n, const int* __restrict__ ien, int nNo,
                                                   int* __restrict__ err)
{
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const int a = ien[q];
  if (a < 0 || a >= nNo) atomicOr(err, 1);
}

__global__ __launch_bounds__(256) void k_pair_keys(int nEl, int nNo, const int4* __restrict__ ien,
                                                   unsigned long long* __restrict__ keys)
{
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long nPair = 16LL * nEl;
  if (t < nPair) {
    const int e = (int)(t >> 4), a = (int)((t >> 2) & 3), b = (int)(t & 3);
    const int4 v = ien[e];
    const int na = a == 0 ? v.x : a == 1 ? v.y : a == 2 ? v.z : v.w;
    const int nb = b == 0 ? v.x : b == 1 ? v.y : b == 2 ? v.z : v.w;
    keys[t] = (unsigned long long)na * (unsigned)nNo + (unsigned)nb;
  } else if (t < nPair + nNo) {
    const unsigned long long a = (unsigned long long)(t - nPair);
    keys[t] = a * (unsigned)nNo + a;
  }
}

// Unique sorted keys -> colPtr, rowPtr (first entry of each row; every row
// holds its diagonal), diagPtr.
__global__ __launch_bounds__(256) void k_csr_from_keys(int nnz, int nNo, const unsigned long long* __restrict__ keys,
                                                       int* __restrict__ rowPtr, int* __restrict__ colPtr,
                                                       int* __restrict__ diagPtr)
{
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nnz) return;
  const unsigned long long key = keys[k];
  const int r = (int)(key / (unsigned)nNo), c = (int)(key - (unsigned long long)r * (unsigned)nNo);
  colPtr[k] = c;
  if (k == 0 || (int)(keys[k - 1] / (unsigned)nNo) != r) rowPtr[r] = k;
  if (c == r) diagPtr[r] = k;
  if (k == nnz - 1) rowPtr[nNo] = nnz;
}

__device__ __forceinline__ unsigned mix32(unsigned x)
{
  // a 32-bit integer hash (the murmur3 finaliser)
  x ^= x >> 16;
  x *= 0x85ebca6bu;
  x ^= x >> 13;
  x *= 0xc2b2ae35u;
  x ^= x >> 16;
  return x;
}

__device__ __forceinline__ unsigned long long prio(int e)
{
  return ((unsigned long long)mix32((unsigned)e * 2654435761u + 0x9e3779b9u) << 32) | (unsigned)e;
}

__global__ __launch_bounds__(256) void k_jp_propose(int nEl, const int4* __restrict__ ien,
                                                    const int* __restrict__ color,
                                                    unsigned long long* __restrict__ best)
{
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nEl || color[e] >= 0) return;
  const unsigned long long p = prio(e);
  const int4 v = ien[e];
  atomicMax(best + v.x, p);
  atomicMax(best + v.y, p);
  atomicMax(best + v.z, p);
  atomicMax(best + v.w, p);
}

__global__ __launch_bounds__(256) void k_jp_select(int nEl, const int4* __restrict__ ien, int* __restrict__ color,
                                                   const unsigned long long* __restrict__ best,
                                                   unsigned long long* __restrict__ used, int* __restrict__ nDone,
                                                   int* __restrict__ maxColor, int* __restrict__ err)
{
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nEl || color[e] >= 0) return;
  const unsigned long long p = prio(e);
  const int4 v = ien[e];
  if (best[v.x] != p || best[v.y] != p || best[v.z] != p || best[v.w] != p) return;
  // e holds all four of its nodes this round: no other winner touches their masks
  const int nd[4] = {v.x, v.y, v.z, v.w};
  unsigned long long m[CW] = {0, 0, 0, 0};
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int w = 0; w < CW; w++) m[w] |= used[(size_t)CW * nd[a] + w];
  int c = -1;
  for (int w = 0; w < CW && c < 0; w++)
    if (~m[w]) c = 64 * w + __builtin_ctzll(~m[w]);
  if (c < 0) {
    atomicOr(err, 2);
    color[e] = 0x7fffffff;  // leaves the loop; the error is reported
    atomicAdd(nDone, 1);
    return;
  }
  color[e] = c;
#pragma unroll
  for (int a = 0; a < 4; a++) used[(size_t)CW * nd[a] + c / 64] |= 1ull << (c % 64);
  atomicAdd(nDone, 1);
  atomicMax(maxColor, c);
}

// Colour-sorted element q = order[q0]: its nodes and 16 block slots (binary
// search of each column in the row's sorted colPtr range).
__global__ __launch_bounds__(256) void k_slots(int nEl, int nElPad, const int* __restrict__ order,
                                               const int4* __restrict__ ien, const int* __restrict__ rowPtr,
                                               const int* __restrict__ colPtr, int4* __restrict__ ienSorted,
                                               int* __restrict__ slot, int* __restrict__ err)
{
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nEl) return;
  const int4 v = ien[order[q]];
  ienSorted[q] = v;
  const int nd[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
  for (int a = 0; a < 4; a++) {
    const int s = rowPtr[nd[a]], e = rowPtr[nd[a] + 1];
#pragma unroll
    for (int b = 0; b < 4; b++) {
      int lo = s, hi = e;  // first position with colPtr >= nd[b]
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (colPtr[mid] < nd[b])
          lo = mid + 1;
        else
          hi = mid;
      }
      if (lo == e || colPtr[lo] != nd[b]) {
        atomicOr(err, 4);
        lo = s;
      }
      slot[(size_t)(4 * a + b) * nElPad + q] = lo;
    }
  }
}

__global__ __launch_bounds__(256) void k_iota(int n, int* __restrict__ v)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = i;
}

int bits_for(unsigned long long v)
{
  int b = 1;
  while (b < 64 && (v >> b) != 0) b++;
  return b;
}

// Device scratch released on every return path.
struct Scratch {
  std::vector<void*> p;
  template <typename T>
  int get(T** out, size_t n)
  {
    int rc = dev_alloc(out, n);
    if (rc == SVG_OK) p.push_back((void*)*out);
    return rc;
  }
  ~Scratch()
  {
    for (void* q : p) (void)hipFree(q);
  }
};

// Element-order assembly map: entry k = 16 e + 4 a + b of element e (caller
// order) adds to the block slot (ien[a], ien[b]); key[k] = that slot.
__global__ __launch_bounds__(256) void k_contrib_keys(int nEl, const int4* __restrict__ ien,
                                                      const int* __restrict__ rowPtr, const int* __restrict__ colPtr,
                                                      int* __restrict__ key, int* __restrict__ val,
                                                      int* __restrict__ err)
{
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= 16LL * nEl) return;
  const int e = (int)(k >> 4), a = (int)((k >> 2) & 3), b = (int)(k & 3);
  const int4 v = ien[e];
  const int nd[4] = {v.x, v.y, v.z, v.w};
  const int row = a == 0 ? nd[0] : a == 1 ? nd[1] : a == 2 ? nd[2] : nd[3];
  const int col = b == 0 ? nd[0] : b == 1 ? nd[1] : b == 2 ? nd[2] : nd[3];
  const int s = rowPtr[row], t = rowPtr[row + 1];
  int lo = s, hi = t;  // first position with colPtr >= col
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (colPtr[mid] < col)
      lo = mid + 1;
    else
      hi = mid;
  }
  if (lo == t || colPtr[lo] != col) {
    atomicOr(err, 4);
    lo = s;
  }
  key[k] = lo;
  val[k] = (int)k;
}

// key[s] = (window of slot s's row) * 256 + its entry count (clamped to 255),
// val[s] = s: the gather's lanes walk windows of kGatherRows rows in order and,
// inside a window, the slots by entry count.
constexpr int kGatherRows = 1024;
__global__ __launch_bounds__(256) void k_slot_count(int nnz, const int* __restrict__ cptr,
                                                    const int* __restrict__ clist, const int4* __restrict__ ien,
                                                    int* __restrict__ key, int* __restrict__ val)
{
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nnz) return;
  const int c0 = cptr[s], n = cptr[s + 1] - c0;
  int row = 0;
  if (n > 0) {
    const int code = clist[c0];
    const int4 v = ien[code >> 4];
    const int a = (code >> 2) & 3;
    row = a == 0 ? v.x : a == 1 ? v.y : a == 2 ? v.z : v.w;
  }
  key[s] = (row / kGatherRows) * 256 + min(n, 255);
  val[s] = s;
}

// cptr[s] = first position of slot s in the sorted keys (lower bound), s in [0, nnz].
__global__ __launch_bounds__(256) void k_contrib_ptr(int nnz, long long n, const int* __restrict__ sorted,
                                                     int* __restrict__ cptr)
{
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s > nnz) return;
  long long lo = 0, hi = n;
  while (lo < hi) {
    const long long mid = (lo + hi) >> 1;
    if (sorted[mid] < s)
      lo = mid + 1;
    else
      hi = mid;
  }
  cptr[s] = (int)lo;
}

int check_ien(long long n, const int* d_ien, int nNo, int* d_err, hipStream_t s, const char* who)
{
  SVG_HIP(hipMemsetAsync(d_err, 0, sizeof(int), s));
  hipLaunchKernelGGL(k_check_ien, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, d_ien, nNo, d_err);
  int h = 0;
  SVG_HIP(hipMemcpyAsync(&h, d_err, sizeof(int), hipMemcpyDeviceToHost, s));
  SVG_HIP(hipStreamSynchronize(s));
  if (h) return set_error(SVG_ERR_ARG, std::string(who) + ": IEN entry out of range");
  return SVG_OK;
}

}  // namespace

// The element-order assembly's map and record storage for M (see svg_msh):
// every (element, a, b) entry keyed by its block slot, sorted stably by slot
// (so each slot's entries stay in element order), and the slot offsets.
int msh_gather_build(svg_msh* M, const int* d_IEN, hipStream_t s)
{
  const svg_lhs* L = M->lhs;
  const int nEl = M->nEl, nnz = L->nnz;
  if (nEl >= (1 << 27)) return set_error(SVG_ERR_ARG, "element-order assembly: more than 2^27 elements");
  if (reinterpret_cast<uintptr_t>(d_IEN) & 15)
    return set_error(SVG_ERR_ARG, "element-order assembly: IEN must be 16-byte aligned (read as int4)");
  const long long n = 16LL * nEl;
  Scratch S;
  int rc;
  int *key = nullptr, *sorted = nullptr, *val = nullptr, *d_err = nullptr;
  if ((rc = dev_alloc(&M->d_ien0, (size_t)4 * nEl)) || (rc = dev_alloc(&M->d_cptr, (size_t)nnz + 1)) ||
      (rc = dev_alloc(&M->d_clist, (size_t)n)) || (rc = dev_alloc(&M->d_rec, (size_t)kTet4Rec * nEl)))
    return rc;
  if ((rc = S.get(&key, (size_t)n)) || (rc = S.get(&sorted, (size_t)n)) || (rc = S.get(&val, (size_t)n)) ||
      (rc = S.get(&d_err, 1)))
    return rc;
  SVG_HIP(hipMemcpyAsync(M->d_ien0, d_IEN, sizeof(int) * 4 * (size_t)nEl, hipMemcpyDeviceToDevice, s));
  SVG_HIP(hipMemsetAsync(d_err, 0, sizeof(int), s));
  hipLaunchKernelGGL(k_contrib_keys, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, nEl,
                     reinterpret_cast<const int4*>(M->d_ien0), (const int*)L->d_rowPtr, (const int*)L->d_colPtr, key,
                     val, d_err);
  SVG_HIP(hipGetLastError());
  const int endBit = bits_for((unsigned long long)(nnz > 0 ? nnz - 1 : 0));
  size_t tb = 0;
  SVG_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, (const int*)key, sorted, (const int*)val, M->d_clist,
                                             (int)n, 0, endBit, s));
  unsigned char* tmp = nullptr;
  if ((rc = S.get(&tmp, tb))) return rc;
  SVG_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, tb, (const int*)key, sorted, (const int*)val, M->d_clist, (int)n,
                                             0, endBit, s));
  hipLaunchKernelGGL(k_contrib_ptr, dim3(blocks_for((size_t)nnz + 1, 256)), dim3(256), 0, s, nnz, n,
                     (const int*)sorted, M->d_cptr);
  SVG_HIP(hipGetLastError());
  // The gather's lane order: windows of rows in order and, inside a window,
  // slots sorted (stably) by their entry count, so that the lanes of a
  // wavefront loop over equally many entries (diagonal blocks have about four
  // times the entries of the others) while the element records a window
  // reads stay few enough for one XCD's L2.
  if ((rc = dev_alloc(&M->d_sorder, (size_t)nnz))) return rc;
  hipLaunchKernelGGL(k_slot_count, dim3(blocks_for((size_t)nnz, 256)), dim3(256), 0, s, nnz,
                     (const int*)M->d_cptr, (const int*)M->d_clist, reinterpret_cast<const int4*>(M->d_ien0), key,
                     val);
  SVG_HIP(hipGetLastError());
  const int endBit2 = bits_for((unsigned long long)(L->nNo / kGatherRows) * 256 + 255);
  size_t tb2 = 0;
  SVG_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb2, (const int*)key, sorted, (const int*)val, M->d_sorder,
                                             nnz, 0, endBit2, s));
  unsigned char* tmp2 = tmp;
  if (tb2 > tb && (rc = S.get(&tmp2, tb2))) return rc;
  SVG_HIP(hipcub::DeviceRadixSort::SortPairs(tmp2, tb2, (const int*)key, sorted, (const int*)val, M->d_sorder, nnz,
                                             0, endBit2, s));
  int herr = 0;
  SVG_HIP(hipMemcpyAsync(&herr, d_err, sizeof(int), hipMemcpyDeviceToHost, s));
  SVG_HIP(hipStreamSynchronize(s));
  if (herr & 4) return set_error(SVG_ERR_ARG, "slot_map: pattern does not contain an element pair");
  return SVG_OK;
}

}  // namespace svg

using namespace svg;

extern "C" {

int svg_lhs_create_dev(int nNo, int nEl, const int* d_IEN, int nFaces, svg_lhs** out, void* stream)
{
  hipStream_t s = (hipStream_t)stream;
  if (!out || !d_IEN || nNo <= 0 || nEl <= 0 || nFaces < 0)
    return set_error(SVG_ERR_ARG, "svg_lhs_create_dev: bad argument");
  if (reinterpret_cast<uintptr_t>(d_IEN) & 15)
    return set_error(SVG_ERR_ARG, "svg_lhs_create_dev: d_IEN must be 16-byte aligned (read as int4)");
  if ((unsigned long long)nNo * (unsigned long long)nNo >= (1ULL << 62))
    return set_error(SVG_ERR_ARG, "svg_lhs_create_dev: nNo too large");
  Scratch S;
  int rc;
  int* d_err = nullptr;
  if ((rc = S.get(&d_err, 1))) return rc;
  if ((rc = check_ien(4LL * nEl, d_IEN, nNo, d_err, s, "svg_lhs_create_dev"))) return rc;
  const long long K = 16LL * nEl + nNo;
  if (K > 0x7fffffffLL) return set_error(SVG_ERR_ARG, "svg_lhs_create_dev: too many node pairs");
  unsigned long long *keys = nullptr, *sorted = nullptr, *uniq = nullptr;
  int* d_nnz = nullptr;
  if ((rc = S.get(&keys, (size_t)K)) || (rc = S.get(&sorted, (size_t)K)) || (rc = S.get(&uniq, (size_t)K)) ||
      (rc = S.get(&d_nnz, 1)))
    return rc;
  hipLaunchKernelGGL(k_pair_keys, dim3((unsigned)((K + 255) / 256)), dim3(256), 0, s, nEl, nNo,
                     reinterpret_cast<const int4*>(d_IEN), keys);
  const int endBit = bits_for((unsigned long long)nNo * (unsigned long long)nNo - 1);
  size_t tb1 = 0, tb2 = 0;
  SVG_HIP(hipcub::DeviceRadixSort::SortKeys(nullptr, tb1, keys, sorted, (int)K, 0, endBit, s));
  SVG_HIP(hipcub::DeviceSelect::Unique(nullptr, tb2, sorted, uniq, d_nnz, (int)K, s));
  unsigned char* tmp = nullptr;
  if ((rc = S.get(&tmp, std::max(tb1, tb2)))) return rc;
  SVG_HIP(hipcub::DeviceRadixSort::SortKeys(tmp, tb1, keys, sorted, (int)K, 0, endBit, s));
  SVG_HIP(hipcub::DeviceSelect::Unique(tmp, tb2, sorted, uniq, d_nnz, (int)K, s));
  int nnz = 0;
  SVG_HIP(hipMemcpyAsync(&nnz, d_nnz, sizeof(int), hipMemcpyDeviceToHost, s));
  SVG_HIP(hipStreamSynchronize(s));
  if ((long long)nnz * 16 > 0x7fffffffLL) return set_error(SVG_ERR_ARG, "svg_lhs_create_dev: 16*nnz overflows int");
  auto* L = new svg_lhs();
  L->nNo = nNo;
  L->nnz = nnz;
  L->nFaces = nFaces;
  L->face.resize(nFaces);
  if ((rc = dev_alloc(&L->d_rowPtr, nNo + 1)) || (rc = dev_alloc(&L->d_colPtr, nnz)) ||
      (rc = dev_alloc(&L->d_diagPtr, nNo)) || (rc = dev_alloc(&L->d_mask, (size_t)4 * nNo))) {
    svg_lhs_free(L);
    return rc;
  }
  hipLaunchKernelGGL(k_csr_from_keys, dim3(blocks_for(nnz, 256)), dim3(256), 0, s, nnz, nNo,
                     (const unsigned long long*)uniq, L->d_rowPtr, L->d_colPtr, L->d_diagPtr);
  // host mirrors of the pattern (the lhs keeps them for its host-side users)
  L->rowPtr.resize(nNo + 1);
  L->colPtr.resize(nnz);
  L->diagPtr.resize(nNo);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess)
    e = hipMemcpyAsync(L->rowPtr.data(), L->d_rowPtr, sizeof(int) * (nNo + 1), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipMemcpyAsync(L->colPtr.data(), L->d_colPtr, sizeof(int) * nnz, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess)
    e = hipMemcpyAsync(L->diagPtr.data(), L->d_diagPtr, sizeof(int) * nNo, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    svg_lhs_free(L);
    return hip_check(e, "svg_lhs_create_dev");
  }
  *out = L;
  return SVG_OK;
}

int svg_lhs_nnz(const svg_lhs* L, int* nnz)
{
  if (!L || !nnz) return set_error(SVG_ERR_ARG, "svg_lhs_nnz: null argument");
  *nnz = L->nnz;
  return SVG_OK;
}

int svg_lhs_pattern(const svg_lhs* L, int* rowPtr, int* colPtr)
{
  if (!L || !rowPtr || !colPtr) return set_error(SVG_ERR_ARG, "svg_lhs_pattern: null argument");
  std::memcpy(rowPtr, L->rowPtr.data(), sizeof(int) * (L->nNo + 1));
  std::memcpy(colPtr, L->colPtr.data(), sizeof(int) * L->nnz);
  return SVG_OK;
}

int svg_msh_create_dev(const svg_lhs* L, int nEl, const int* d_IEN, const double* d_x, svg_msh** out,
                       void* stream)
{
  hipStream_t s = (hipStream_t)stream;
  if (!L || !d_IEN || !d_x || !out || nEl <= 0) return set_error(SVG_ERR_ARG, "svg_msh_create_dev: bad argument");
  if (reinterpret_cast<uintptr_t>(d_IEN) & 15)
    return set_error(SVG_ERR_ARG, "svg_msh_create_dev: d_IEN must be 16-byte aligned (read as int4)");
  const int nNo = L->nNo;
  Scratch S;
  int rc;
  int *d_err = nullptr, *color = nullptr, *d_cnt = nullptr, *order = nullptr, *ids = nullptr, *colSorted = nullptr;
  unsigned long long *best = nullptr, *used = nullptr;
  if ((rc = S.get(&d_err, 1))) return rc;
  if ((rc = check_ien(4LL * nEl, d_IEN, nNo, d_err, s, "svg_msh_create_dev"))) return rc;
  if ((rc = S.get(&color, nEl)) || (rc = S.get(&d_cnt, 2)) || (rc = S.get(&best, nNo)) ||
      (rc = S.get(&used, (size_t)CW * nNo)) || (rc = S.get(&order, nEl)) || (rc = S.get(&ids, nEl)) ||
      (rc = S.get(&colSorted, nEl)))
    return rc;
  const int4* ien4 = reinterpret_cast<const int4*>(d_IEN);
  SVG_HIP(hipMemsetAsync(color, 0xff, sizeof(int) * nEl, s));  // -1: uncoloured
  SVG_HIP(hipMemsetAsync(used, 0, sizeof(unsigned long long) * CW * nNo, s));
  SVG_HIP(hipMemsetAsync(d_cnt, 0, 2 * sizeof(int), s));     // [0] coloured elements, [1] largest colour
  const unsigned gE = blocks_for(nEl, 256);
  int h[2] = {0, 0};
  int rounds = 0;
  while (h[0] < nEl) {
    if (++rounds > 100000) return set_error(SVG_ERR_ARG, "svg_msh_create_dev: colouring did not finish");
    // four rounds per host check: a round after the last element is coloured does nothing
    for (int r = 0; r < 4; r++) {
      SVG_HIP(hipMemsetAsync(best, 0, sizeof(unsigned long long) * nNo, s));
      hipLaunchKernelGGL(k_jp_propose, dim3(gE), dim3(256), 0, s, nEl, ien4, (const int*)color, best);
      hipLaunchKernelGGL(k_jp_select, dim3(gE), dim3(256), 0, s, nEl, ien4, color,
                         (const unsigned long long*)best, used, d_cnt, d_cnt + 1, d_err);
    }
    SVG_HIP(hipMemcpyAsync(h, d_cnt, 2 * sizeof(int), hipMemcpyDeviceToHost, s));
    SVG_HIP(hipStreamSynchronize(s));
  }
  int herr = 0;
  SVG_HIP(hipMemcpy(&herr, d_err, sizeof(int), hipMemcpyDeviceToHost));
  if (herr & 2) return set_error(SVG_ERR_ARG, "svg_msh_create_dev: more than 256 colours needed");
  const int nColors = h[1] + 1;
  // colour order: stable radix sort of the element ids by colour
  hipLaunchKernelGGL(k_iota, dim3(gE), dim3(256), 0, s, nEl, ids);
  size_t tb = 0;
  SVG_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, (const int*)color, colSorted, (const int*)ids, order, nEl,
                                             0, 8, s));
  unsigned char* tmp = nullptr;
  if ((rc = S.get(&tmp, tb))) return rc;
  SVG_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, tb, (const int*)color, colSorted, (const int*)ids, order, nEl, 0,
                                             8, s));
  std::vector<int> hc(nEl);
  SVG_HIP(hipMemcpyAsync(hc.data(), colSorted, sizeof(int) * nEl, hipMemcpyDeviceToHost, s));
  auto* M = new svg_msh();
  M->lhs = L;
  M->nEl = nEl;
  M->nNo = nNo;
  M->nColors = nColors;
  M->nElPad = (nEl + 63) & ~63;
  if ((rc = dev_alloc(&M->d_x, (size_t)3 * nNo)) || (rc = dev_alloc(&M->d_ien, (size_t)4 * nEl)) ||
      (rc = dev_alloc(&M->d_slot, (size_t)16 * M->nElPad)) || (rc = dev_alloc(&M->d_err, 1))) {
    svg_msh_free(M);
    return rc;
  }
  hipError_t e = hipMemsetAsync(M->d_slot, 0, sizeof(int) * 16 * M->nElPad, s);
  if (e == hipSuccess) e = hipMemsetAsync(d_err, 0, sizeof(int), s);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_slots, dim3(gE), dim3(256), 0, s, nEl, M->nElPad, (const int*)order, ien4,
                       (const int*)L->d_rowPtr, (const int*)L->d_colPtr, reinterpret_cast<int4*>(M->d_ien),
                       M->d_slot, d_err);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(M->d_x, d_x, sizeof(double) * 3 * nNo, hipMemcpyDeviceToDevice, s);
  if (e == hipSuccess) e = hipMemsetAsync(M->d_err, 0, sizeof(int), s);
  M->ien.resize((size_t)4 * nEl);
  if (e == hipSuccess)
    e = hipMemcpyAsync(M->ien.data(), d_IEN, sizeof(int) * 4 * (size_t)nEl, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipMemcpyAsync(&herr, d_err, sizeof(int), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e == hipSuccess) e = hipHostMalloc((void**)&M->h_err, sizeof(int), hipHostMallocDefault);
  if (e != hipSuccess) {
    svg_msh_free(M);
    return hip_check(e, "svg_msh_create_dev");
  }
  if (herr & 4) {
    svg_msh_free(M);
    return set_error(SVG_ERR_ARG, "slot_map: pattern does not contain an element pair");
  }
  M->colorOff.assign(nColors + 1, 0);
  for (int q = 0; q < nEl; q++) M->colorOff[hc[q] + 1]++;
  for (int c = 0; c < nColors; c++) M->colorOff[c + 1] += M->colorOff[c];
  if ((rc = msh_gather_build(M, d_IEN, s))) {
    svg_msh_free(M);
    return rc;
  }
  *out = M;
  return SVG_OK;
}

int svg_msh_colors(const svg_msh* M, int* colorOff)
{
  if (!M || !colorOff) return set_error(SVG_ERR_ARG, "svg_msh_colors: null argument");
  std::memcpy(colorOff, M->colorOff.data(), sizeof(int) * (M->nColors + 1));
  return SVG_OK;
}

int svg_msh_sorted_ien(const svg_msh* M, int* ien)
{
  if (!M || !ien) return set_error(SVG_ERR_ARG, "svg_msh_sorted_ien: null argument");
  SVG_HIP(hipMemcpy(ien, M->d_ien, sizeof(int) * 4 * (size_t)M->nEl, hipMemcpyDeviceToHost));
  return SVG_OK;
}

}  // extern "C"
