// Device-side declarations shared by the HIP translation units: the opaque
// handle types behind the C ABI, error/launch helpers and event timing.
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "svg_internal.h"

namespace svg {

// Returns SVG_ERR_HIP with a message if `e` is not hipSuccess.
int hip_check(hipError_t e, const char* what);

#define SVG_HIP(call)                                   \
  do {                                                  \
    hipError_t _e = (call);                             \
    if (_e != hipSuccess) return ::svg::hip_check(_e, #call); \
  } while (0)

// Event timing of named kernels, with the algorithmic bytes of each launch
// (for achieved-bandwidth reporting). A timed launch goes through
// hipExtLaunchKernelGGL with the slot's start and stop events, which the
// dispatch records at the kernel's own start and end (the duration rocprof
// reports), not around the launch call with the gap before it.
int prof_begin(const char* name, hipStream_t s, double bytes, hipEvent_t* start, hipEvent_t* stop);
bool prof_set(bool on);  // returns the previous setting (event timing off under graph capture)
// Position in the list of timed launches not yet read back, and removal of the
// timed launches between two positions (launches that found their solve
// finished and exited at once: their durations are not the kernel's).
size_t prof_mark();
void prof_drop(size_t from, size_t to);

template <typename K, typename... Args>
inline void launch_b(const char* name, double bytes, K kernel, dim3 grid, dim3 block, size_t shmem, hipStream_t s,
                     Args... args)
{
  hipEvent_t a = nullptr, b = nullptr;
  if (prof_begin(name, s, bytes, &a, &b) >= 0)
    hipExtLaunchKernelGGL(kernel, grid, block, (uint32_t)shmem, s, a, b, 0u, args...);
  else
    hipLaunchKernelGGL(kernel, grid, block, shmem, s, args...);
}

template <typename K, typename... Args>
inline void launch(const char* name, K kernel, dim3 grid, dim3 block, size_t shmem, hipStream_t s, Args... args)
{
  launch_b(name, 0.0, kernel, grid, block, shmem, s, args...);
}

inline unsigned blocks_for(size_t n, unsigned bs) { return (unsigned)((n + bs - 1) / bs); }

// Workgroups are dispatched round-robin over the 8 XCDs; remap so that each
// XCD walks one contiguous band of rows and its L2 keeps the entries its rows
// gather (-1: a padding workgroup of a grid rounded up to a multiple of 8).
__device__ __forceinline__ int xcd_block(int b, int nb)
{
  const int per = (nb + 7) >> 3;
  const int xcd = b & 7, idx = b >> 3;
  const int r = xcd * per + idx;
  return r < nb ? r : -1;
}

__device__ __forceinline__ double wave_sum(double v)
{
  // Fixed butterfly over the 64 lanes of a wavefront.
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

template <typename T>
int dev_alloc(T** p, size_t count)
{
  *p = nullptr;
  if (count == 0) count = 1;
  hipError_t e = hipMalloc((void**)p, count * sizeof(T));
  if (e != hipSuccess) return hip_check(e, "hipMalloc");
  return SVG_OK;
}

template <typename T>
void dev_free(T*& p)
{
  if (p) (void)hipFree(p);
  p = nullptr;
}

// Device state of the GMRES cycle (one per solve workspace).
struct GmState {
  int done;     // |err(i+1)| < eps reached in the current cycle (first field: kernels skip on it)
  int last_i;   // last Arnoldi column of the cycle
  int itr;      // SpMVs + restarts counted in this call
  int suc;
  int sD;
  int divzero;  // err(0) == 0 at a restart (Array division by zero)
  int early;    // inner GMRES: |R| <= absTol on entry, X stays 0 (gmres.cpp:146-150)
  int tconv;    // mixed precision: the FP64 true residual at a restart is below eps
  int fallback; // learned preconditioner: no new residual minimum for the fallback window
  int nondec;   // iterations since the last new minimum of |err|
  int act;      // batched solve: the system still iterates (not converged before this cycle)
  int cyc;      // batched solve: restart cycles the system ran
  int fbitr;    // batched solve: itr after the cycle in which the system fell back (-1: none)
  int pad;
  double best;  // smallest |err(i+1)| of the solve
  double eps;
  double scale;  // err(0) after a restart, 1 / h(i+1,i) after an Arnoldi step
  double iNorm, fNorm;
  double fStart;  // batched solve: fNorm at the start of the system's last cycle
};

// Device state of the Schur-complement CG (cgrad.cpp:50-160). errO is double
// buffered by trip parity: trip i reads errO[i&1] and its last kernel writes
// errO[(i+1)&1], so no kernel writes a value its own workgroups read.
struct CgState {
  int stop;     // err < eps seen (or mItr reached): later trips are skipped
  int last_i;
  int suc;
  int pad;
  double eps, err, iNorm;
  double errO[2];
  double errO_last;  // errO of the last executed trip (for dB)
};

// Krylov workspace kept resident between solves on one lhs.
struct KrylovWork {
  int n = 0;        // dof * nNo
  int sD = 0;
  size_t ld = 0;    // stride between basis vectors (elements)
  double* u = nullptr;        // (sD+1) basis vectors
  double* X = nullptr;        // solution accumulator
  double* W = nullptr;        // preconditioner diagonal (dof,nNo)
  double* h = nullptr;        // (sD+1) x sD Hessenberg, column-major
  double* cs = nullptr;       // Givens cosines (sD)
  double* sn = nullptr;       // Givens sines (sD)
  double* err = nullptr;      // (sD+1)
  double* coef = nullptr;     // projection coefficients of the current column (sD+2)
  double* y = nullptr;        // (sD)
  double* partial = nullptr;  // reduction partials (nblk x (sD+2))
  int nblk = 0;
  GmState* st = nullptr;
  GmState* h_st = nullptr;    // pinned host mirror
  int* h_done = nullptr;      // mapped host flag written by the device tail
  int* d_done = nullptr;      // device view of h_done
  hipEvent_t ev[8] = {};      // per-step completion events (look-ahead ring)
  float* u32 = nullptr;       // mixed precision, flexible: the network's FP32 z_i (the SpMV input)
  float* v32 = nullptr;       // mixed precision: the Krylov basis in FP32, (sD+1) vectors at stride ld
  double* z = nullptr;        // flexible GMRES: preconditioned basis z_j = M(u_j), sD vectors
  float* z32 = nullptr;       // mixed precision, two-level preconditioner: z_j in FP32, sD vectors
  float* w32 = nullptr;       // mixed precision, two-level preconditioner: the FP32 SpMV's output A z_i
  double* r2 = nullptr;       // two-level with the network as smoother: its input B (u - A P zc)
  double* scal = nullptr;     // scratch scalars for standalone dots
};

// Workspace of the bipartitioned NS solver (ns_solver.cpp:167-499).
struct NsWork {
  int nNo = 0, cap = 0;     // capacity in outer iterations
  size_t ld3 = 0, ld1 = 0;  // strides of the (3,nNo) and (nNo) vectors
  double *mK = nullptr, *mG = nullptr, *mD = nullptr, *mL = nullptr, *Gt = nullptr;
  double *Rm = nullptr, *Rmi = nullptr, *Rc = nullptr, *Rci = nullptr;
  double *U = nullptr, *MU = nullptr, *P = nullptr, *MP = nullptr;  // cap, 2cap, cap, 2cap vectors
  double *cgX = nullptr, *cgP = nullptr, *cgSP = nullptr, *cgGP = nullptr;
  double* cgGP0 = nullptr;  // copy of G P read by a coupled face's op PRE (cgrad.cpp:104-107)
  CgState* cg = nullptr;
  CgState* h_cg = nullptr;
  double* coef = nullptr;   // xB on the device
  double* dots = nullptr;   // Gram entries
  double* h_dots = nullptr;
  double* partial = nullptr;
  int partialCap = 0;
};

// Two-level (aggregation coarse level + block Jacobi) preconditioner of an lhs.
struct CoarseLevel;
void coarse_free(CoarseLevel*& c);
// Builds the coarse level of L for aggregates of about aggSize nodes (once per
// lhs and size), then this system's operators from the scaled Val (synchronises s).
// keep: reuse the coarse inverse of the lhs's last set-up when there is one
// (A P and the block-Jacobi inverses are rebuilt from Val either way).
int coarse_prepare_dev(svg_lhs* L, const double* Val, int aggSize, hipStream_t s, bool keep = false);
// The same in three parts: the build (host; one dense coarse block per system of
// sysNodeOff, null: one system), the launches alone (capturable) and the pivot check.
int coarse_setup(svg_lhs* L, int aggSize, const int* sysNodeOff, int nsys);
void coarse_enqueue(svg_lhs* L, const double* Val, hipStream_t s, bool invert = true);
int coarse_check(const svg_lhs* L, hipStream_t s);
// z = M(v) in FP64 (z64) and/or FP32 (z32); skip: kernels return at once when *skip.
void coarse_apply_dev(const svg_lhs* L, const float* v, double* z64, float* z32, const int* skip, hipStream_t s);
void coarse_apply_dev(const svg_lhs* L, const double* v, double* z64, float* z32, const int* skip, hipStream_t s);
// The batched solve's form (one coarse block per system): the systems that do
// not iterate in the cycle (per st and sysOfNode) are skipped, and the
// systems that fell back to the diagonal preconditioner get z = v.
void coarse_apply_batch_dev(const svg_lhs* L, const float* v, double* z64, float* z32, const GmState* st,
                            const int* sysOfNode, hipStream_t s);
void coarse_apply_batch_dev(const svg_lhs* L, const double* v, double* z64, float* z32, const GmState* st,
                            const int* sysOfNode, hipStream_t s);
// With another smoother: r2 = v - A P zc (FP64; the coarse solution zc is kept),
// then, after the smoother wrote z from r2, z(3, i) += zc(agg(i)).
void coarse_residual_dev(const svg_lhs* L, const float* v, double* r2, const int* skip, hipStream_t s);
void coarse_residual_dev(const svg_lhs* L, const double* v, double* r2, const int* skip, hipStream_t s);
// y = B (v - A P zc) (block Jacobi on the coarse-corrected residual, without P zc).
void coarse_smooth_dev(const svg_lhs* L, const float* v, double* y, const int* skip, hipStream_t s);
void coarse_smooth_dev(const svg_lhs* L, const double* v, double* y, const int* skip, hipStream_t s);
void coarse_add_dev(const svg_lhs* L, double* z64, float* z32, const int* skip, hipStream_t s);
int coarse_num_aggregates(const svg_lhs* L);
unsigned long long coarse_generation(const svg_lhs* L);  // changes when the level is rebuilt

// Shared between the translation units.
void krylov_free(KrylovWork& k);   // every buffer, pinned mirror and event of a Krylov workspace
int lhs_ensure_trans(svg_lhs* L);  // block-transpose map (pattern must be symmetric)
int spmv_residual(const svg_lhs* L, const double* K, const double* U, const double* B, double* out, hipStream_t s);
int lpc_prepare_dev(svg_lhs* L, const double* Val, hipStream_t s);
// nsc: per-node (sc, 1/sc); sysOf/st: skip the rows of batched systems that stopped iterating.
int lpc_apply_dev(svg_lpc* N, const svg_lhs* L, const double* V, double* Z, float* Z32, const int* skip,
                  hipStream_t s, const double2* nsc = nullptr, const int* sysOf = nullptr,
                  const GmState* st = nullptr);
// Every buffer lpc_prepare_dev / lpc_apply_dev use on L, allocated ahead (graph capture).
int lpc_reserve(svg_lpc* N, svg_lhs* L, hipStream_t s);
// Differs between networks and changes when a network's activation buffers
// are reallocated (a graph holding its forward pass must then be recaptured).
unsigned long long lpc_generation(const svg_lpc* N);
// construct_fluid split in two: the launches alone (no host synchronisation, so
// they can be captured into a graph; the zero-Jacobian flag stays on the device)
// and the check of that flag, which synchronises s.
int construct_fluid_enqueue(svg_msh* M, const svg_fluid_params* p, const double* Ag, const double* Yg,
                            const double* Bf, double* R, double* Val, hipStream_t s);
int construct_fluid_check(svg_msh* M, hipStream_t s);

}  // namespace svg

struct svg_face_rec {
  int nNo = 0, dof = 0, bGrp = SVG_BC_DIR;
  bool created = false, incFlag = false, coupledFlag = false;
  double res = 0.0;
  double nS = 0.0;  // |valM|^2 (the NS solver's bc_pre, ns_solver.cpp:56-89)
  std::vector<int> glob;
  std::vector<double> val;
  // Neumann faces: device copies for add_bc_mul; valM = val W after precond_diag
  int* d_glob = nullptr;
  double* d_val = nullptr;
  double* d_valM = nullptr;
};

struct svg_lhs {
  int nNo = 0, nnz = 0, nFaces = 0;
  std::vector<int> rowPtr, colPtr, diagPtr;
  int* d_rowPtr = nullptr;   // nNo+1, exclusive end
  int* d_colPtr = nullptr;
  int* d_diagPtr = nullptr;
  int* d_trans = nullptr;    // block index of the transposed block (lazy, NS solver)
  std::vector<svg_face_rec> face;
  double* d_mask = nullptr;  // (4,nNo): product of Dirichlet face values per dof
  std::vector<int> maskKey;  // incFlag pattern the mask was built for
  bool maskBuilt = false;
  int spmvVariant = -1;  // 4x4 SpMV launch variant picked by the autotuner (-1: not tuned)
  int spmv32Variant = -1;  // FP32 4x4 SpMV launch variant picked by the autotuner (-1: not tuned)
  mutable int msgVariant = -1;  // network message-gather launch: 0 round-robin, 1 XCD bands (-1: not tuned)
  float* d_val32 = nullptr;  // mixed precision: FP32 copy of the scaled Val
  unsigned short* d_val16 = nullptr;   // learned preconditioner: BF16 copy of the scaled Val
  unsigned short* d_diag16 = nullptr;  // learned preconditioner: BF16 diagonal blocks (16, nNo)
  svg::CoarseLevel* coarse = nullptr;  // two-level preconditioner (built on first use)
  svg::KrylovWork kw;    // gmres_v on the 4x4 system
  svg::KrylovWork kw3;   // inner GMRES of the NS solver on the 3x3 momentum block
  svg::NsWork ns;
};

// Batched GMRES over a block-diagonal union of independent systems (one lhs
// over all of them): per-system Krylov spaces, reductions and convergence.
struct svg_batch {
  svg_lhs* L = nullptr;
  int nsys = 0, sD = 0, n = 0, nsp = 0;   // n = 4 * union nodes; nsp = reduction spans
  size_t ld = 0;
  std::vector<int> nodeOff;               // nsys + 1
  int *d_sysOfNode = nullptr, *d_spanStart = nullptr, *d_spanEnd = nullptr, *d_spanSys = nullptr;
  int* d_sysSpanOff = nullptr;            // nsys + 1
  double *u = nullptr, *X = nullptr, *h = nullptr, *cs = nullptr, *sn = nullptr, *err = nullptr, *coef = nullptr,
         *y = nullptr, *partial = nullptr;
  svg::GmState* st = nullptr;
  svg::GmState* h_st = nullptr;           // pinned mirror (nsys)
  float* u32 = nullptr;                   // mixed, learned: the network's FP32 z_i (the SpMV input)
  float* v32 = nullptr;                   // mixed precision: the Krylov basis in FP32 ((sD+1) vectors)
  double* z = nullptr;                    // learned preconditioner: flexible basis z_j (sD vectors)
  float* z32 = nullptr;                   // two-level preconditioner, mixed: the flexible basis in FP32
  double2* d_nsc = nullptr;               // per node: (sqrt(nNo_q), 1 / sqrt(nNo_q)) of its system q
  std::vector<int> fbItr;                 // per system: iteration count at the fallback (-1: none)
  hipStream_t cs_stream = nullptr;        // capture/replay stream
  // restart cycles: [FP64 / mixed][one cycle / kBatchCycles cycles per replay]
  hipGraphExec_t graph[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
  std::vector<unsigned char> gKey[2][2];  // the arguments each cycle graph was captured with
  long long replays = 0;                  // restart-cycle graph replays (host synchronisations) so far
  hipGraphExec_t pro = nullptr, epi = nullptr;   // Newton iteration: pici..initial norms / result..picc
  std::vector<unsigned char> newtonKey;          // the arguments pro and epi were captured with
  hipEvent_t ev = nullptr;
};

struct svg_msh {
  const svg_lhs* lhs = nullptr;
  int nEl = 0, nNo = 0, nColors = 0;
  std::vector<int> colorOff;   // nColors+1 offsets into the colour-sorted order
  double* d_x = nullptr;       // (3,nNo)
  int* d_ien = nullptr;        // (4, nEl) in colour-sorted order (int4 per element)
  int* d_slot = nullptr;       // 16 slots per element, SoA: slot[p*nElPad + e]
  int* d_err = nullptr;        // zero-Jacobian flag (element index + 1)
  int* h_err = nullptr;        // pinned mirror
  int nElPad = 0;
  std::vector<int> ien;        // (4, nEl) in the caller's element order (face -> parent lookups)
  // Element-order assembly (construct_fluid): per block slot s of the pattern,
  // the element pairs that add to it, clist[cptr[s] .. cptr[s+1]) = e*16 + 4a + b
  // in increasing element order e, and the per-element record the gather reads.
  int* d_ien0 = nullptr;       // (4, nEl) in the caller's element order (int4 per element)
  int* d_cptr = nullptr;       // nnz + 1
  int* d_clist = nullptr;      // 16 nEl
  int* d_sorder = nullptr;     // the nnz slots ordered by their entry count (the gather's lane order)
  double* d_rec = nullptr;     // kTet4Rec * nEl
};

// Doubles per element in the element-order assembly's record.
constexpr int kTet4Rec = 80;
// Builds d_ien0, d_cptr, d_clist and d_rec from the caller-order IEN on the
// device (d_IEN: 4 nEl ints, 16-byte aligned); synchronises s.
namespace svg {
int msh_gather_build(svg_msh* M, const int* d_IEN, hipStream_t s);
}

// A TRI3 Neumann face of a fluid mesh (faceType: IEN, gE), colour-sorted.
struct svg_face {
  const svg_msh* msh = nullptr;
  int nFe = 0, nColors = 0;
  std::vector<int> colorOff;
  int* d_ien = nullptr;    // (4, nFe): the 3 face nodes and the parent element's interior node
  int* d_slot = nullptr;   // 9 slots per face element, SoA: slot[p*nFePad + e]
  int nFePad = 0;
};
