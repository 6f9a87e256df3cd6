// C++ host side of the fluid Newton step, above the C ABI of libsvgpu.so.
//
// Mirrors the reference's host interface for this path with the same names,
// argument meaning and error behaviour (std::runtime_error carrying the
// library's message where the reference throws):
//   fsi_linear_solver::fsils_lhs_create / fsils_bc_create / fsils_ls_create /
//   fsils_solve / fsils_lhs_free        (Code/Source/liner_solver/fsils_api.hpp:37-61)
//   fluid::construct_fluid              (Code/Source/solver/fluid.h:51, fluid.cpp:464)
//   pic::pici / the picc update / picp   (Code/Source/solver/pic.cpp:486-574, 155-167, 591-715)
//   set_bc::set_bc_dir                   (Code/Source/solver/set_bc.cpp:794-1063)
// Arrays keep the reference layouts; the solver arrays (Ri, Val) and the
// nodal state live in device memory (DeviceArray), the pattern and face
// lists are host vectors as in the reference.
#pragma once

#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "svfsi_gpu.h"

namespace svg_host {

// fils_struct.hpp:60-77, consts.h:426-440
enum class BcType { BC_TYPE_Dir = SVG_BC_DIR, BC_TYPE_Neu = SVG_BC_NEU };
enum class LinearSolverType {
  LS_TYPE_BICGS = SVG_LS_BICGS,
  LS_TYPE_NS = SVG_LS_NS,
  LS_TYPE_GMRES = SVG_LS_GMRES,
  LS_TYPE_CG = SVG_LS_CG
};
enum class PreconditionerType { PREC_NONE = SVG_PREC_NONE, PREC_FSILS = SVG_PREC_FSILS };

// Throws std::runtime_error with the library's message when rc != SVG_OK.
void check(int rc);

// An owning device buffer.
template <typename T>
class DeviceArray {
 public:
  DeviceArray() = default;
  explicit DeviceArray(size_t n) { resize(n); }
  DeviceArray(const DeviceArray&) = delete;
  DeviceArray& operator=(const DeviceArray&) = delete;
  DeviceArray(DeviceArray&& o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr, o.n_ = 0; }
  ~DeviceArray() { release(); }
  void resize(size_t n)
  {
    release();
    if (n && hipMalloc((void**)&p_, n * sizeof(T)) != hipSuccess) throw std::runtime_error("hipMalloc failed");
    n_ = n;
  }
  void upload(const std::vector<T>& h)
  {
    if (h.size() != n_) resize(h.size());
    if (n_ && hipMemcpy(p_, h.data(), n_ * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
      throw std::runtime_error("hipMemcpy failed");
  }
  std::vector<T> download() const
  {
    std::vector<T> h(n_);
    if (n_ && hipMemcpy(h.data(), p_, n_ * sizeof(T), hipMemcpyDeviceToHost) != hipSuccess)
      throw std::runtime_error("hipMemcpy failed");
    return h;
  }
  void copy_from(const DeviceArray& o)
  {
    if (o.n_ != n_) resize(o.n_);
    if (n_ && hipMemcpy(p_, o.p_, n_ * sizeof(T), hipMemcpyDeviceToDevice) != hipSuccess)
      throw std::runtime_error("hipMemcpy failed");
  }
  T* data() { return p_; }
  const T* data() const { return p_; }
  size_t size() const { return n_; }

 private:
  void release()
  {
    if (p_) (void)hipFree(p_);
    p_ = nullptr;
    n_ = 0;
  }
  T* p_ = nullptr;
  size_t n_ = 0;
};

// FSILS_lhsType (fils_struct.hpp:171-213) on one rank: the handle of the
// device-resident pattern, faces and Krylov workspace.
struct FSILS_lhsType {
  svg_lhs* h = nullptr;
  int nNo = 0, nnz = 0, nFaces = 0;
  FSILS_lhsType() = default;
  FSILS_lhsType(const FSILS_lhsType&) = delete;
  FSILS_lhsType& operator=(const FSILS_lhsType&) = delete;
  ~FSILS_lhsType();
};
using FSILS_subLsType = svg_subls;  // fils_struct.hpp:216-263
using FSILS_lsType = svg_ls;        // fils_struct.hpp:265-283

// lhs.cpp:57-120 with the exclusive-end rowPtr(nNo+1) of com_mod.rowPtr.
void fsils_lhs_create(FSILS_lhsType& lhs, int nNo, int nnz, const std::vector<int>& rowPtr,
                      const std::vector<int>& colPtr, int nFaces);
// bc.cpp:45-130; Val(dof, nNo) defaults to zero (baf_ini.cpp:725, 742).
void fsils_bc_create(FSILS_lhsType& lhs, int faIn, int nNo, int dof, BcType BC_type, const std::vector<int>& gNodes,
                     const std::vector<double>& Val = {});
// ls.cpp:40-120 defaults.
void fsils_ls_create(FSILS_lsType& ls, LinearSolverType LS_type);
// solve.cpp:50-193: Ri(dof,nNo) in device memory (residual in, increment out),
// Val(dof*dof,nnz) in device memory (scaled in place).
void fsils_solve(FSILS_lhsType& lhs, FSILS_lsType& ls, int dof, double* Ri, double* Val, PreconditionerType prec,
                 const std::vector<int>& incL, const std::vector<double>& res, hipStream_t stream = nullptr);
void fsils_lhs_free(FSILS_lhsType& lhs);

// The element data of one TET4 fluid mesh (mshType's nEl, IEN, and com_mod.x).
struct FluidMesh {
  svg_msh* h = nullptr;
  int nEl = 0, nColors = 0;
  FluidMesh() = default;
  FluidMesh(const FluidMesh&) = delete;
  FluidMesh& operator=(const FluidMesh&) = delete;
  ~FluidMesh();
};
void fluid_mesh_create(FluidMesh& msh, const FSILS_lhsType& lhs, int nEl, const std::vector<int>& IEN,
                       const std::vector<double>& x);

// A TRI3 Neumann face of the fluid mesh (faceType IEN(3,nFe), gE(nFe)).
struct FluidFace {
  svg_face* h = nullptr;
  int nFe = 0;
  FluidFace() = default;
  FluidFace(const FluidFace&) = delete;
  FluidFace& operator=(const FluidFace&) = delete;
  ~FluidFace();
};
void fluid_face_create(FluidFace& fa, const FluidMesh& msh, const std::vector<int>& faceIEN,
                       const std::vector<int>& gE);
// b_assem_neu_bc (eq_assem.cpp:58-172) with b_fluid (fluid.cpp:46-133): the
// Neumann value hg (device, per node; null: zero traction) and the backflow
// stabilisation beta.
void b_assem_neu_bc(FluidFace& fa, const svg_fluid_params& p, double beta, double qmTRI3, const double* hg,
                    const double* Yg, double* R, double* Val, hipStream_t stream = nullptr);

// construct_fluid (fluid.cpp:464-708): R(4,nNo) += residuals, Val(16,nnz) +=
// tangents; every array in device memory; Bf may be null.
void construct_fluid(FluidMesh& msh, const svg_fluid_params& p, const double* Ag, const double* Yg, const double* Bf,
                     double* R, double* Val, hipStream_t stream = nullptr);

// pici (pic.cpp:486-574) and the picc update (pic.cpp:155-167).
void pici(int nNo, double am, double af, const double* Ao, const double* An, const double* Yo, const double* Yn,
          double* Ag, double* Yg, hipStream_t stream = nullptr);
void picc_update(int nNo, double gam, double dt, const double* R, double* An, double* Yn, hipStream_t stream = nullptr);
// picp (pic.cpp:591-715) for the fluid equation: An = Ao (gam-1)/gam, Yn = Yo.
void picp(int nNo, double gam, const double* Ao, const double* Yo, double* An, double* Yn,
          hipStream_t stream = nullptr);

// One strong Dirichlet face for set_bc_dir (set_bc.cpp:794-1063): its nodes
// (fa.gN) and gx nV (lDof = nsd) or gx per node, (lDof, nNo), on the device.
struct DirichletFace {
  DeviceArray<int> gN;
  DeviceArray<double> gxn;
  int nNo = 0, lDof = 0;
  unsigned eDir = 0;  // bc.eDrn as a direction mask (0: all lDof rows)
};
void dirichlet_face_create(DirichletFace& f, const std::vector<int>& gN, int lDof, const std::vector<double>& gxn,
                           unsigned eDir = 0);
// A(s.., gN) = dirA gxn, Y(s.., gN) = dirY gxn for A, Y (tDof, tnNo); steady
// values: dirA = 0, dirY = bc.g.
void set_bc_dir(const DirichletFace& f, int tDof, int s, double dirA, double dirY, double* A, double* Y,
                hipStream_t stream = nullptr);

// write_restart (output.cpp:202-360, header output.cpp:365-376) for one rank
// and the fluid equation: stamp {1, 1, 1, nNo, 0, 4, 0}, cTS, time, CPU time,
// eq.iNorm, Yn, An (4 x nNo each, dof fastest, copied from the device), into
// <stFileName>_<cTS>.bin (cTS as %03d below 1000) with the hard link
// <stFileName>_last.bin, or into <stFileName>_last.bin alone with replace.
// Returns the file name.
std::string write_restart(const std::string& stFileName, int cTS, double time, double cpuTime, double iNorm,
                          int nNo, const double* d_Yn, const double* d_An, bool replace = false);

}  // namespace svg_host
