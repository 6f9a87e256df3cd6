// Runtime of the C ABI: errors, event timing, the device-resident lhs and
// mesh handles, and the element-wise generalized-alpha kernels.
#include "svg_device.h"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>

namespace svg {

static thread_local std::string g_err;

int set_error(int code, const std::string& msg)
{
  g_err = msg;
  return code;
}

int hip_check(hipError_t e, const char* what)
{
  return set_error(SVG_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

// ---------------------------------------------------------------- timing --
struct ProfAcc {
  int count = 0;
  double ms = 0.0;
  double bytes = 0.0;
};
struct ProfPending {
  std::string name;
  hipEvent_t a, b;
  double bytes;
};
static std::mutex g_prof_mu;
static bool g_prof_on = false;
static std::vector<std::string> g_prof_only;  // empty: every named launch is timed
static int g_prof_every = 1;                  // time one launch in this many of each kernel
static std::map<std::string, long> g_prof_calls;

bool prof_set(bool on)
{
  const bool prev = g_prof_on;
  g_prof_on = on;
  return prev;
}
static std::map<std::string, ProfAcc> g_prof;
static std::vector<ProfPending> g_pending;
static std::vector<hipEvent_t> g_event_pool;

static hipEvent_t take_event()
{
  if (!g_event_pool.empty()) {
    hipEvent_t e = g_event_pool.back();
    g_event_pool.pop_back();
    return e;
  }
  hipEvent_t e;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

int prof_begin(const char* name, hipStream_t s, double bytes, hipEvent_t* start, hipEvent_t* stop)
{
  (void)s;
  if (!g_prof_on) return -1;
  std::lock_guard<std::mutex> lk(g_prof_mu);
  if (!g_prof_only.empty() &&
      std::find(g_prof_only.begin(), g_prof_only.end(), std::string(name)) == g_prof_only.end())
    return -1;
  if (g_prof_every > 1 && (g_prof_calls[name]++ % g_prof_every) != 0) return -1;
  ProfPending p;
  p.name = name;
  p.bytes = bytes;
  p.a = take_event();
  p.b = take_event();
  if (!p.a || !p.b) return -1;
  *start = p.a;
  *stop = p.b;
  g_pending.push_back(p);
  return (int)g_pending.size() - 1;
}

size_t prof_mark()
{
  std::lock_guard<std::mutex> lk(g_prof_mu);
  return g_pending.size();
}

void prof_drop(size_t from, size_t to)
{
  std::lock_guard<std::mutex> lk(g_prof_mu);
  if (to > g_pending.size()) to = g_pending.size();
  if (from >= to) return;
  for (size_t q = from; q < to; q++) {
    g_event_pool.push_back(g_pending[q].a);
    g_event_pool.push_back(g_pending[q].b);
  }
  g_pending.erase(g_pending.begin() + from, g_pending.begin() + to);
}

static void prof_flush()
{
  std::lock_guard<std::mutex> lk(g_prof_mu);
  for (auto& p : g_pending) {
    float ms = 0.f;
    if (hipEventSynchronize(p.b) == hipSuccess && hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
      auto& acc = g_prof[p.name];
      acc.count++;
      acc.ms += ms;
      acc.bytes += p.bytes;
    }
    g_event_pool.push_back(p.a);
    g_event_pool.push_back(p.b);
  }
  g_pending.clear();
}

// ------------------------------------------------------- element kernels --
__global__ void k_pici(int n, double c0, double c1, double c2, double c3, const double* __restrict__ Ao,
                       const double* __restrict__ An, const double* __restrict__ Yo, const double* __restrict__ Yn,
                       double* __restrict__ Ag, double* __restrict__ Yg)
{
#pragma clang fp contract(off)
  size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (size_t)n) return;
  Ag[k] = Ao[k] * c0 + An[k] * c1;
  Yg[k] = Yo[k] * c2 + Yn[k] * c3;
}

__global__ void k_picc(int n, double coef, const double* __restrict__ R, double* __restrict__ An,
                       double* __restrict__ Yn)
{
#pragma clang fp contract(off)
  size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (size_t)n) return;
  double r = R[k];
  An[k] = An[k] - r;
  Yn[k] = Yn[k] - r * coef;
}

// picp for the fluid rows (pic.cpp:676-693): An = Ao coef, Yn = Yo.
__global__ void k_picp(int n, double coef, const double* __restrict__ Ao, const double* __restrict__ Yo,
                       double* __restrict__ An, double* __restrict__ Yn)
{
#pragma clang fp contract(off)
  size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (size_t)n) return;
  An[k] = Ao[k] * coef;
  Yn[k] = Yo[k];
}

// set_bc_dir_l (set_bc.cpp:1005-1063, non-generic profile) with the scatter of
// set_bc_dir (set_bc.cpp:884-938, no impD): one lane per (face node a, k < lDof).
// Row of k: s + k, or, with the direction mask eDir, s + (the k-th set bit).
__global__ void k_set_bc_dir(int nNo, int lDof, int tDof, int s, unsigned eDir, const int* __restrict__ gN,
                             const double* __restrict__ gxn, double dirA, double dirY, double* __restrict__ A,
                             double* __restrict__ Y)
{
#pragma clang fp contract(off)
  size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)nNo * lDof) return;
  const int a = (int)(t / lDof), k = (int)(t % lDof);
  int row = s + k;
  if (eDir) {
    int c = 0;
    for (int i = 0; i < 32; i++)
      if ((eDir >> i) & 1u) {
        if (c == k) {
          row = s + i;
          break;
        }
        c++;
      }
  }
  const double g = gxn[(size_t)lDof * a + k];
  const size_t at = (size_t)tDof * gN[a] + row;
  A[at] = dirA * g;
  Y[at] = dirY * g;
}

__global__ void k_fill(size_t n, double v, double* __restrict__ d)
{
  size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; k < n; k += stride) d[k] = v;
}

}  // namespace svg

using namespace svg;

extern "C" {

const char* svg_last_error(void) { return svg::g_err.c_str(); }

int svg_prof_enable(int on)
{
  g_prof_on = (on != 0);
  return SVG_OK;
}

int svg_prof_only(const char* names)
{
  std::lock_guard<std::mutex> lk(g_prof_mu);
  g_prof_only.clear();
  if (!names) return SVG_OK;
  std::string cur;
  for (const char* c = names;; c++) {
    if (*c == ',' || *c == 0) {
      if (!cur.empty()) g_prof_only.push_back(cur);
      cur.clear();
      if (*c == 0) break;
    } else {
      cur += *c;
    }
  }
  return SVG_OK;
}

int svg_prof_sample(int every)
{
  if (every < 1) return set_error(SVG_ERR_ARG, "svg_prof_sample: every must be >= 1");
  std::lock_guard<std::mutex> lk(g_prof_mu);
  g_prof_every = every;
  g_prof_calls.clear();
  return SVG_OK;
}

int svg_prof_reset(void)
{
  prof_flush();
  std::lock_guard<std::mutex> lk(g_prof_mu);
  g_prof.clear();
  return SVG_OK;
}

int svg_prof_get(const char* name, int* count, double* total_ms)
{
  prof_flush();
  std::lock_guard<std::mutex> lk(g_prof_mu);
  auto it = g_prof.find(name ? name : "");
  *count = it == g_prof.end() ? 0 : it->second.count;
  *total_ms = it == g_prof.end() ? 0.0 : it->second.ms;
  return SVG_OK;
}

int svg_prof_get_bytes(const char* name, double* bytes)
{
  prof_flush();
  std::lock_guard<std::mutex> lk(g_prof_mu);
  auto it = g_prof.find(name ? name : "");
  *bytes = it == g_prof.end() ? 0.0 : it->second.bytes;
  return SVG_OK;
}

int svg_pici(int nNo, double am, double af, const double* Ao, const double* An, const double* Yo, const double* Yn,
             double* Ag, double* Yg, void* stream)
{
  if (nNo < 0) return set_error(SVG_ERR_ARG, "svg_pici: nNo < 0");
  if (nNo == 0) return SVG_OK;
  int n = 4 * nNo;
  launch("pici", k_pici, dim3(blocks_for(n, 256)), dim3(256), 0, (hipStream_t)stream, n, 1.0 - am, am, 1.0 - af,
         af, Ao, An, Yo, Yn, Ag, Yg);
  SVG_HIP(hipGetLastError());
  return SVG_OK;
}

int svg_picc_update(int nNo, double gam, double dt, const double* R, double* An, double* Yn, void* stream)
{
  if (nNo < 0) return set_error(SVG_ERR_ARG, "svg_picc_update: nNo < 0");
  if (nNo == 0) return SVG_OK;
  int n = 4 * nNo;
  launch("picc", k_picc, dim3(blocks_for(n, 256)), dim3(256), 0, (hipStream_t)stream, n, gam * dt, R, An, Yn);
  SVG_HIP(hipGetLastError());
  return SVG_OK;
}

int svg_picp(int nNo, double gam, const double* Ao, const double* Yo, double* An, double* Yn, void* stream)
{
  if (nNo < 0) return set_error(SVG_ERR_ARG, "svg_picp: nNo < 0");
  if (nNo == 0) return SVG_OK;
  if (!Ao || !Yo || !An || !Yn) return set_error(SVG_ERR_ARG, "svg_picp: null pointer");
  if (gam == 0.0) return set_error(SVG_ERR_ARG, "svg_picp: gam = 0");
  const int n = 4 * nNo;
  launch("picp", k_picp, dim3(blocks_for(n, 256)), dim3(256), 0, (hipStream_t)stream, n, (gam - 1.0) / gam, Ao, Yo,
         An, Yn);
  SVG_HIP(hipGetLastError());
  return SVG_OK;
}

int svg_set_bc_dir(int nNo, const int* gN, int tDof, int s, int lDof, unsigned eDir, const double* gxn, double dirA,
                   double dirY, double* A, double* Y, void* stream)
{
  if (nNo < 0 || lDof < 1 || tDof < 1 || s < 0) return set_error(SVG_ERR_ARG, "svg_set_bc_dir: bad sizes");
  if (nNo == 0) return SVG_OK;
  if (!gN || !gxn || !A || !Y) return set_error(SVG_ERR_ARG, "svg_set_bc_dir: null pointer");
  int top = s + lDof - 1;
  if (eDir) {
    if (__builtin_popcount(eDir) != lDof) return set_error(SVG_ERR_ARG, "svg_set_bc_dir: lDof != directions in eDir");
    top = s + 31 - __builtin_clz(eDir);
  }
  if (top >= tDof) return set_error(SVG_ERR_ARG, "svg_set_bc_dir: rows beyond tDof");
  const size_t n = (size_t)nNo * lDof;
  launch("set_bc_dir", k_set_bc_dir, dim3(blocks_for(n, 256)), dim3(256), 0, (hipStream_t)stream, nNo, lDof, tDof,
         s, eDir, gN, gxn, dirA, dirY, A, Y);
  SVG_HIP(hipGetLastError());
  return SVG_OK;
}

int svg_fill(double* d, size_t n, double v, void* stream)
{
  if (n == 0) return SVG_OK;
  if (!d) return set_error(SVG_ERR_ARG, "svg_fill: null pointer");
  unsigned nb = (unsigned)std::min<size_t>(blocks_for(n, 256), 8192);
  launch("fill", k_fill, dim3(nb), dim3(256), 0, (hipStream_t)stream, n, v, d);
  SVG_HIP(hipGetLastError());
  return SVG_OK;
}

// ------------------------------------------------------------------ lhs --
int svg_lhs_create(int nNo, int nnz, const int* rowPtr, const int* colPtr, int nFaces, svg_lhs** out)
{
  if (!out || !rowPtr || !colPtr || nNo <= 0 || nnz <= 0 || nFaces < 0)
    return set_error(SVG_ERR_ARG, "svg_lhs_create: bad argument");
  if ((long long)nnz * 16 > 0x7fffffffLL) return set_error(SVG_ERR_ARG, "svg_lhs_create: 16*nnz overflows int");
  if (rowPtr[0] != 0 || rowPtr[nNo] != nnz) return set_error(SVG_ERR_ARG, "svg_lhs_create: rowPtr does not span nnz");
  auto* L = new svg_lhs();
  L->nNo = nNo;
  L->nnz = nnz;
  L->nFaces = nFaces;
  L->rowPtr.assign(rowPtr, rowPtr + nNo + 1);
  L->colPtr.assign(colPtr, colPtr + nnz);
  L->diagPtr.assign(nNo, -1);
  for (int a = 0; a < nNo; a++) {
    if (rowPtr[a + 1] < rowPtr[a]) {
      delete L;
      return set_error(SVG_ERR_ARG, "svg_lhs_create: rowPtr not monotone");
    }
    for (int i = rowPtr[a]; i < rowPtr[a + 1]; i++) {
      if (colPtr[i] < 0 || colPtr[i] >= nNo) {
        delete L;
        return set_error(SVG_ERR_ARG, "svg_lhs_create: colPtr entry out of range");
      }
      if (i > rowPtr[a] && colPtr[i] <= colPtr[i - 1]) {
        delete L;
        return set_error(SVG_ERR_ARG, "svg_lhs_create: columns not strictly increasing within a row");
      }
      if (colPtr[i] == a) L->diagPtr[a] = i;
    }
    if (L->diagPtr[a] < 0) {
      delete L;
      return set_error(SVG_ERR_ARG, "svg_lhs_create: row without a diagonal block");
    }
  }
  L->face.resize(nFaces);
  int rc = SVG_OK;
  if ((rc = dev_alloc(&L->d_rowPtr, nNo + 1)) || (rc = dev_alloc(&L->d_colPtr, nnz)) ||
      (rc = dev_alloc(&L->d_diagPtr, nNo)) || (rc = dev_alloc(&L->d_mask, (size_t)4 * nNo))) {
    svg_lhs_free(L);
    return rc;
  }
  hipError_t e = hipMemcpy(L->d_rowPtr, rowPtr, sizeof(int) * (nNo + 1), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(L->d_colPtr, colPtr, sizeof(int) * nnz, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(L->d_diagPtr, L->diagPtr.data(), sizeof(int) * nNo, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    svg_lhs_free(L);
    return hip_check(e, "svg_lhs_create: upload");
  }
  *out = L;
  return SVG_OK;
}

int svg_lhs_bc_create(svg_lhs* L, int faIn, int nNo, int dof, int bGrp, const int* gNodes, const double* val)
{
  if (!L) return set_error(SVG_ERR_ARG, "svg_lhs_bc_create: null lhs");
  if (faIn >= L->nFaces)
    return set_error(SVG_ERR_ARG, "FSILS: faIn is exceeding lhs structure maximum number of faces (" +
                                      std::to_string(L->nFaces) + ") is less than " + std::to_string(faIn) + ".");
  if (faIn <= -1) return set_error(SVG_ERR_ARG, "FSILS: faIn is smaller than zero");
  if (nNo < 0 || dof <= 0 || (nNo > 0 && !gNodes)) return set_error(SVG_ERR_ARG, "svg_lhs_bc_create: bad face");
  auto& f = L->face[faIn];
  f.created = true;
  f.nNo = nNo;
  f.dof = dof;
  f.bGrp = bGrp;
  f.glob.assign(gNodes, gNodes + nNo);
  for (int a = 0; a < nNo; a++)
    if (f.glob[a] < 0 || f.glob[a] >= L->nNo) return set_error(SVG_ERR_ARG, "svg_lhs_bc_create: node out of range");
  f.val.assign((size_t)dof * nNo, 0.0);
  if (val) std::memcpy(f.val.data(), val, sizeof(double) * dof * nNo);
  L->maskBuilt = false;
  dev_free(f.d_glob);
  dev_free(f.d_val);
  dev_free(f.d_valM);
  if (bGrp == SVG_BC_NEU && nNo > 0) {
    // a Neumann face may be coupled (fsils_solve's res): add_bc_mul needs it on the device
    int rc;
    if ((rc = dev_alloc(&f.d_glob, (size_t)nNo)) || (rc = dev_alloc(&f.d_val, (size_t)dof * nNo)) ||
        (rc = dev_alloc(&f.d_valM, (size_t)dof * nNo)))
      return rc;
    SVG_HIP(hipMemcpy(f.d_glob, f.glob.data(), sizeof(int) * nNo, hipMemcpyHostToDevice));
    SVG_HIP(hipMemcpy(f.d_val, f.val.data(), sizeof(double) * dof * nNo, hipMemcpyHostToDevice));
    SVG_HIP(hipMemset(f.d_valM, 0, sizeof(double) * dof * nNo));  // valM(dof, nNo) starts at 0 (bc.cpp:80)
  }
  return SVG_OK;
}

int svg_lhs_free(svg_lhs* L)
{
  if (!L) return SVG_OK;
  dev_free(L->d_rowPtr);
  dev_free(L->d_colPtr);
  dev_free(L->d_diagPtr);
  dev_free(L->d_trans);
  dev_free(L->d_mask);
  svg::krylov_free(L->kw);
  svg::krylov_free(L->kw3);
  svg::coarse_free(L->coarse);
  dev_free(L->d_val32);
  dev_free(L->d_val16);
  dev_free(L->d_diag16);
  for (auto& f : L->face) {
    dev_free(f.d_glob);
    dev_free(f.d_val);
    dev_free(f.d_valM);
  }
  auto& w = L->ns;
  for (double* p : {w.mK, w.mG, w.mD, w.mL, w.Gt, w.Rm, w.Rmi, w.Rc, w.Rci, w.U, w.MU, w.P, w.MP, w.cgX, w.cgP,
                    w.cgSP, w.cgGP, w.cgGP0, w.coef, w.dots, w.partial})
    if (p) (void)hipFree(p);
  if (w.cg) (void)hipFree(w.cg);
  if (w.h_cg) (void)hipHostFree(w.h_cg);
  if (w.h_dots) (void)hipHostFree(w.h_dots);
  delete L;
  return SVG_OK;
}

// ----------------------------------------------------------------- mesh --
int svg_msh_create(const svg_lhs* L, int nEl, const int* IEN, int nNo, const double* x, svg_msh** out)
{
  if (!L || !IEN || !x || !out || nEl <= 0) return set_error(SVG_ERR_ARG, "svg_msh_create: bad argument");
  if (nNo != L->nNo) return set_error(SVG_ERR_ARG, "svg_msh_create: nNo differs from the lhs");
  std::vector<int> color, slots;
  int nColors = 0;
  int rc = color_elements(nNo, nEl, 4, IEN, color, &nColors);
  if (rc) return rc;
  rc = slot_map(nEl, 4, IEN, L->rowPtr.data(), L->colPtr.data(), slots);
  if (rc) return rc;
  // Colour-sorted order (stable: element order inside each colour).
  std::vector<int> off(nColors + 1, 0), order(nEl);
  for (int e = 0; e < nEl; e++) off[color[e] + 1]++;
  for (int c = 0; c < nColors; c++) off[c + 1] += off[c];
  {
    std::vector<int> fill(off.begin(), off.end() - 1);
    for (int e = 0; e < nEl; e++) order[fill[color[e]]++] = e;
  }
  auto* M = new svg_msh();
  M->lhs = L;
  M->nEl = nEl;
  M->nNo = nNo;
  M->nColors = nColors;
  M->colorOff = off;
  M->nElPad = (nEl + 63) & ~63;
  M->ien.assign(IEN, IEN + (size_t)4 * nEl);
  std::vector<int> ien((size_t)4 * nEl), sl((size_t)16 * M->nElPad, 0);
  for (int q = 0; q < nEl; q++) {
    int e = order[q];
    for (int a = 0; a < 4; a++) ien[4 * (size_t)q + a] = IEN[4 * (size_t)e + a];
    for (int p = 0; p < 16; p++) sl[(size_t)p * M->nElPad + q] = slots[16 * (size_t)e + p];
  }
  if ((rc = dev_alloc(&M->d_x, (size_t)3 * nNo)) || (rc = dev_alloc(&M->d_ien, (size_t)4 * nEl)) ||
      (rc = dev_alloc(&M->d_slot, (size_t)16 * M->nElPad)) || (rc = dev_alloc(&M->d_err, 1))) {
    svg_msh_free(M);
    return rc;
  }
  hipError_t e = hipMemcpy(M->d_x, x, sizeof(double) * 3 * nNo, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(M->d_ien, ien.data(), sizeof(int) * ien.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(M->d_slot, sl.data(), sizeof(int) * sl.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(M->d_err, 0, sizeof(int));
  if (e == hipSuccess) e = hipHostMalloc((void**)&M->h_err, sizeof(int), hipHostMallocDefault);
  if (e != hipSuccess) {
    svg_msh_free(M);
    return hip_check(e, "svg_msh_create: upload");
  }
  // the element-order assembly map, built on the device from the caller-order IEN
  int* d_IEN = nullptr;
  if ((rc = dev_alloc(&d_IEN, (size_t)4 * nEl))) {
    svg_msh_free(M);
    return rc;
  }
  e = hipMemcpy(d_IEN, IEN, sizeof(int) * 4 * (size_t)nEl, hipMemcpyHostToDevice);
  rc = e == hipSuccess ? msh_gather_build(M, d_IEN, nullptr) : hip_check(e, "svg_msh_create: upload");
  dev_free(d_IEN);
  if (rc) {
    svg_msh_free(M);
    return rc;
  }
  *out = M;
  return SVG_OK;
}

int svg_msh_free(svg_msh* M)
{
  if (!M) return SVG_OK;
  dev_free(M->d_x);
  dev_free(M->d_ien);
  dev_free(M->d_slot);
  dev_free(M->d_err);
  dev_free(M->d_ien0);
  dev_free(M->d_cptr);
  dev_free(M->d_clist);
  dev_free(M->d_sorder);
  dev_free(M->d_rec);
  if (M->h_err) (void)hipHostFree(M->h_err);
  delete M;
  return SVG_OK;
}

int svg_msh_num_colors(const svg_msh* M, int* nColors)
{
  if (!M || !nColors) return set_error(SVG_ERR_ARG, "svg_msh_num_colors: null argument");
  *nColors = M->nColors;
  return SVG_OK;
}

}  // extern "C"
