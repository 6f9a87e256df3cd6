// GpuFsilsLinearAlgebra: the reference's LinearAlgebra interface
// (LinearAlgebra.h:40-63) over libsvgpu.so. What each method stands in for:
//
//   alloc            FsilsLinearAlgebra::alloc (FsilsLinearAlgebra.cpp:52-62),
//                    called by ls_alloc (ls.cpp:50-60) before every assembly
//   assemble         FsilsLinearAlgebra::assemble: lhsa_ns::do_assem on the host
//   solve            FsilsLinearAlgebra::solve (FsilsLinearAlgebra.cpp:139-148):
//                    fsils_solve(lhs, lEq.FSILS, dof, R, Val, prec, incL, res)
//   construct_fluid  fluid::construct_fluid (fluid.cpp:464-708)
//   b_assem_neu_bc   eq_assem::b_assem_neu_bc + fluid::b_fluid
//                    (eq_assem.cpp:58-172, fluid.cpp:46-133)
#include "GpuFsilsLinearAlgebra.h"

#include "lhsa.h"

#include <hip/hip_runtime_api.h>

#include <cstring>
#include <stdexcept>
#include <string>

namespace {

void ok(int rc, const char* what)
{
  if (rc != SVG_OK) throw std::runtime_error(std::string("[GpuFsilsLinearAlgebra] ") + what + ": " + svg_last_error());
}

void hip_ok(hipError_t e, const char* what)
{
  if (e != hipSuccess)
    throw std::runtime_error(std::string("[GpuFsilsLinearAlgebra] ") + what + ": " + hipGetErrorString(e));
}

template <typename T>
void dev_alloc(T** p, size_t n, const char* what)
{
  hip_ok(hipMalloc((void**)p, sizeof(T) * (n ? n : 1)), what);
}

bool fsils_preconditioner(consts::PreconditionerType p)
{
  return p == consts::PreconditionerType::PREC_FSILS || p == consts::PreconditionerType::PREC_NONE;
}

// copies FSILS_subLsType <-> svg_subls (the same fields, fils_struct.hpp:211-260)
void sub_in(svg_subls& d, const fsi_linear_solver::FSILS_subLsType& s)
{
  d.suc = s.suc ? 1 : 0;
  d.mItr = s.mItr;
  d.sD = s.sD;
  d.itr = s.itr;
  d.cM = s.cM;
  d.cN = s.cN;
  d.cD = s.cD;
  d.absTol = s.absTol;
  d.relTol = s.relTol;
  d.iNorm = s.iNorm;
  d.fNorm = s.fNorm;
  d.dB = s.dB;
  d.callD = s.callD;
}

void sub_out(fsi_linear_solver::FSILS_subLsType& d, const svg_subls& s)
{
  d.suc = s.suc != 0;
  d.itr = s.itr;
  d.cM = s.cM;
  d.cN = s.cN;
  d.cD = s.cD;
  d.iNorm = s.iNorm;
  d.fNorm = s.fNorm;
  d.dB = s.dB;
  d.callD = s.callD;
}

}  // namespace

GpuFsilsLinearAlgebra::GpuFsilsLinearAlgebra(consts::LinearAlgebraType type)
{
  interface_type = type;
  assembly_type = type;
  preconditioner_type = consts::PreconditionerType::PREC_FSILS;
  opts_.precision = SVG_PRECISION_FP64;
  opts_.method = SVG_METHOD_PLAIN;
}

GpuFsilsLinearAlgebra::~GpuFsilsLinearAlgebra()
{
  for (auto& f : faces_) svg_face_free(f.second);
  if (msh_) svg_msh_free(msh_);
  if (lhs_) svg_lhs_free(lhs_);
  for (double* p : {dR_, dVal_, dAg_, dYg_, dHg_, dBf_})
    if (p) (void)hipFree(p);
  if (hStage_) (void)hipHostFree(hStage_);
}

void GpuFsilsLinearAlgebra::alloc(ComMod& com_mod, eqType& lEq)
{
  // a new Newton iteration: the device R / Val start from zero, as ls_alloc's
  // fresh com_mod.R / com_mod.Val do; the host Val is only needed while the
  // host assembles
  zeroPending_ = true;
  device_assembly_ = false;
  if (!device_mode_) com_mod.Val.resize(com_mod.dof * com_mod.dof, com_mod.lhs.nnz);
}

void GpuFsilsLinearAlgebra::assemble(ComMod& com_mod, const int num_elem_nodes, const Vector<int>& eqN,
                                     const Array3<double>& lK, const Array<double>& lR)
{
  if (device_mode_)
    throw std::runtime_error("[GpuFsilsLinearAlgebra] host assembly after device assembly in one run is not supported");
  lhsa_ns::do_assem(com_mod, num_elem_nodes, eqN, lK, lR);
}

void GpuFsilsLinearAlgebra::check_options(const consts::PreconditionerType prec_cond_type,
                                          const consts::LinearAlgebraType assembly_type)
{
  if (!fsils_preconditioner(prec_cond_type))
    throw std::runtime_error("[GpuFsilsLinearAlgebra] only the fsils (diagonal) preconditioner or none is built");
  if (assembly_type != consts::LinearAlgebraType::none && assembly_type != consts::LinearAlgebraType::fsils &&
      assembly_type != interface_type)
    throw std::runtime_error("[GpuFsilsLinearAlgebra] assembly must be fsils or none");
}

void GpuFsilsLinearAlgebra::initialize(ComMod& com_mod, eqType& lEq)
{
  // The device lhs is created at the first assembly or solve: this is called
  // (main.cpp:72) before fsils_lhs_create has built com_mod.lhs.
}

void GpuFsilsLinearAlgebra::set_assembly(consts::LinearAlgebraType atype)
{
  if (atype == consts::LinearAlgebraType::none) return;
  if (atype != consts::LinearAlgebraType::fsils && atype != interface_type)
    throw std::runtime_error("[GpuFsilsLinearAlgebra] assembly must be fsils or none");
  assembly_type = atype;
}

void GpuFsilsLinearAlgebra::set_preconditioner(consts::PreconditionerType prec_type)
{
  if (!fsils_preconditioner(prec_type))
    throw std::runtime_error("[GpuFsilsLinearAlgebra] only the fsils (diagonal) preconditioner or none is built");
  preconditioner_type = prec_type;
}

void GpuFsilsLinearAlgebra::ensure_device(ComMod& com_mod)
{
  if (lhs_) return;
  auto& lhs = com_mod.lhs;
  if (com_mod.dof != 4) throw std::runtime_error("[GpuFsilsLinearAlgebra] only the fluid equation (dof 4) is built");
  nNo_ = com_mod.tnNo;
  nnz_ = lhs.nnz;
  if (com_mod.rowPtr.size() != nNo_ + 1 || com_mod.colPtr.size() != nnz_)
    throw std::runtime_error("[GpuFsilsLinearAlgebra] com_mod.rowPtr / colPtr do not match lhs");
  // one rank: fsils_lhs_create's map is the identity (lhs.cpp:100-118), so
  // the faces' glob are node ids of com_mod.rowPtr's numbering
  if (lhs.commu.nTasks > 1) throw std::runtime_error("[GpuFsilsLinearAlgebra] one rank per GPU: nTasks must be 1");
  ok(svg_lhs_create(nNo_, nnz_, com_mod.rowPtr.data(), com_mod.colPtr.data(), lhs.nFaces, &lhs_), "lhs_create");
  for (int f = 0; f < lhs.nFaces; f++) {
    const auto& F = lhs.face[f];
    if (F.nNo == 0) continue;  // a face fsils_bc_create never filled
    ok(svg_lhs_bc_create(lhs_, f, F.nNo, F.dof,
                         F.bGrp == fsi_linear_solver::BcType::BC_TYPE_Neu ? SVG_BC_NEU : SVG_BC_DIR, F.glob.data(),
                         F.val.data()),
       "bc_create");
  }
  dev_alloc(&dR_, (size_t)4 * nNo_, "R");
  dev_alloc(&dVal_, (size_t)16 * nnz_, "Val");
  dev_alloc(&dAg_, (size_t)4 * nNo_, "Ag");
  dev_alloc(&dYg_, (size_t)4 * nNo_, "Yg");
  hip_ok(hipHostMalloc((void**)&hStage_, sizeof(double) * 4 * nNo_, hipHostMallocDefault), "staging");
  zeroPending_ = true;
}

void GpuFsilsLinearAlgebra::zero_if_pending(ComMod& com_mod)
{
  if (!zeroPending_) return;
  ok(svg_fill(dR_, (size_t)4 * nNo_, 0.0, nullptr), "fill R");
  ok(svg_fill(dVal_, (size_t)16 * nnz_, 0.0, nullptr), "fill Val");
  zeroPending_ = false;
}

// Ag / Yg are (tDof, tnNo); the fluid equation's four dofs are rows eq.s .. eq.s+3.
void GpuFsilsLinearAlgebra::upload_state(ComMod& com_mod, const Array<double>& Ag, const Array<double>& Yg,
                                         bool withAg)
{
  const int s = com_mod.eq[com_mod.cEq].s;
  const int rows = Yg.nrows();
  if (Yg.ncols() != nNo_ || (withAg && (Ag.ncols() != nNo_ || Ag.nrows() != rows)) || s < 0 || s + 4 > rows)
    throw std::runtime_error("[GpuFsilsLinearAlgebra] Ag / Yg do not hold the fluid dofs of every node");
  const size_t bytes = sizeof(double) * 4 * nNo_;
  auto up = [&](const Array<double>& A, double* d) {
    if (rows == 4) {
      hip_ok(hipMemcpy(d, A.data(), bytes, hipMemcpyHostToDevice), "upload");
      return;
    }
    const double* src = A.data();
    for (int a = 0; a < nNo_; a++)
      for (int i = 0; i < 4; i++) hStage_[i + 4 * a] = src[s + i + (size_t)rows * a];
    hip_ok(hipMemcpy(d, hStage_, bytes, hipMemcpyHostToDevice), "upload");
  };
  if (withAg) up(Ag, dAg_);
  up(Yg, dYg_);
}

static svg_fluid_params fluid_params(const ComMod& com_mod)
{
  using namespace consts;
  const auto& eq = com_mod.eq[com_mod.cEq];
  if (eq.nDmn != 1 || eq.dmn[0].phys != EquationType::phys_fluid)
    throw std::runtime_error("[GpuFsilsLinearAlgebra] one fluid domain is built");
  const auto& dmn = eq.dmn[0];
  if (dmn.fluid_visc.viscType != FluidViscosityModelType::viscType_Const)
    throw std::runtime_error("[GpuFsilsLinearAlgebra] constant viscosity is built");
  auto prop = [&](PhysicalProperyType k) {
    auto it = dmn.prop.find(k);
    return it == dmn.prop.end() ? 0.0 : it->second;
  };
  svg_fluid_params p{};
  p.rho = prop(PhysicalProperyType::fluid_density);
  p.mu = dmn.fluid_visc.mu_i;
  p.f[0] = prop(PhysicalProperyType::f_x);
  p.f[1] = prop(PhysicalProperyType::f_y);
  p.f[2] = prop(PhysicalProperyType::f_z);
  p.K_darcy = prop(PhysicalProperyType::inverse_darcy_permeability);
  p.af = eq.af;
  p.am = eq.am;
  p.gam = eq.gam;
  p.dt = com_mod.dt;
  return p;
}

void GpuFsilsLinearAlgebra::construct_fluid(ComMod& com_mod, const mshType& lM, const Array<double>& Ag,
                                            const Array<double>& Yg)
{
  if (com_mod.mvMsh) throw std::runtime_error("[GpuFsilsLinearAlgebra] a moving mesh is not built");
  if (lM.eNoN != 4 || lM.eType != consts::ElementType::TET4)
    throw std::runtime_error("[GpuFsilsLinearAlgebra] TET4 P1-P1 elements are built");
  ensure_device(com_mod);
  if (!msh_) {
    ok(svg_msh_create(lhs_, lM.nEl, lM.IEN.data(), nNo_, com_mod.x.data(), &msh_), "msh_create");
    mshOf_ = &lM;
    bool bf = false;
    for (int i = 0; i < com_mod.Bf.size() && !bf; i++) bf = com_mod.Bf.data()[i] != 0.0;
    if (bf) {
      if (com_mod.Bf.size() != 3 * nNo_) throw std::runtime_error("[GpuFsilsLinearAlgebra] Bf is not (3, tnNo)");
      dev_alloc(&dBf_, (size_t)3 * nNo_, "Bf");
      hip_ok(hipMemcpy(dBf_, com_mod.Bf.data(), sizeof(double) * 3 * nNo_, hipMemcpyHostToDevice), "Bf");
    }
  } else if (mshOf_ != &lM) {
    throw std::runtime_error("[GpuFsilsLinearAlgebra] one fluid mesh is built");
  }
  svg_fluid_params p = fluid_params(com_mod);
  p.qmTET4 = lM.qmTET4;
  zero_if_pending(com_mod);
  upload_state(com_mod, Ag, Yg, true);
  ok(svg_construct_fluid(msh_, &p, dAg_, dYg_, dBf_, dR_, dVal_, nullptr), "construct_fluid");
  device_assembly_ = true;
  device_mode_ = true;
}

void GpuFsilsLinearAlgebra::b_assem_neu_bc(ComMod& com_mod, const faceType& lFa, const Vector<double>& hg,
                                           const Array<double>& Yg)
{
  if (!msh_) throw std::runtime_error("[GpuFsilsLinearAlgebra] b_assem_neu_bc needs construct_fluid first");
  if (lFa.eNoN != 3) throw std::runtime_error("[GpuFsilsLinearAlgebra] TRI3 faces are built");
  auto it = faces_.find(&lFa);
  if (it == faces_.end()) {
    svg_face* f = nullptr;
    ok(svg_face_create(msh_, lFa.nEl, lFa.IEN.data(), lFa.gE.data(), &f), "face_create");
    it = faces_.emplace(&lFa, f).first;
  }
  svg_fluid_params p = fluid_params(com_mod);
  const auto& dmn = com_mod.eq[com_mod.cEq].dmn[0];
  auto bs = dmn.prop.find(consts::PhysicalProperyType::backflow_stab);
  const double beta = bs == dmn.prop.end() ? 0.0 : bs->second;
  const double* dHg = nullptr;
  bool any = false;
  for (int a = 0; a < hg.size() && !any; a++) any = hg(a) != 0.0;
  if (any) {
    if (hg.size() != nNo_) throw std::runtime_error("[GpuFsilsLinearAlgebra] hg must hold one value per node");
    if (!dHg_) dev_alloc(&dHg_, (size_t)nNo_, "hg");
    hip_ok(hipMemcpy(dHg_, hg.data(), sizeof(double) * nNo_, hipMemcpyHostToDevice), "hg");
    dHg = dHg_;
  }
  zero_if_pending(com_mod);
  upload_state(com_mod, Yg, Yg, false);
  ok(svg_b_assem_neu_bc(it->second, &p, beta, lFa.qmTRI3, dHg, dYg_, dR_, dVal_, nullptr), "b_assem_neu_bc");
  device_assembly_ = true;
  device_mode_ = true;
}

void GpuFsilsLinearAlgebra::solve(ComMod& com_mod, eqType& lEq, const Vector<int>& incL, const Vector<double>& res)
{
  ensure_device(com_mod);
  auto& R = com_mod.R;
  if (R.nrows() != 4 || R.ncols() != nNo_) throw std::runtime_error("[GpuFsilsLinearAlgebra] R is not (4, tnNo)");
  const size_t rBytes = sizeof(double) * 4 * nNo_;
  if (device_mode_ && !device_assembly_)
    throw std::runtime_error("[GpuFsilsLinearAlgebra] solve without an assembly in this Newton iteration");
  if (device_assembly_) {
    const double* h = R.data();
    for (size_t i = 0; i < (size_t)4 * nNo_; i++)
      if (h[i] != 0.0)
        throw std::runtime_error(
            "[GpuFsilsLinearAlgebra] com_mod.R was changed on the host while the assembly ran on the device");
  } else {
    // solver only: the host assembled R and Val (FsilsLinearAlgebra's contract)
    if (com_mod.Val.nrows() != 16 || com_mod.Val.ncols() != nnz_)
      throw std::runtime_error("[GpuFsilsLinearAlgebra] Val is not (16, nnz)");
    hip_ok(hipMemcpy(dR_, R.data(), rBytes, hipMemcpyHostToDevice), "upload R");
    hip_ok(hipMemcpy(dVal_, com_mod.Val.data(), sizeof(double) * 16 * nnz_, hipMemcpyHostToDevice), "upload Val");
    zeroPending_ = true;
  }
  auto& F = lEq.FSILS;
  svg_ls ls{};
  ls.LS_type = static_cast<int>(F.LS_type);
  ls.Resm = F.Resm;
  ls.Resc = F.Resc;
  sub_in(ls.GM, F.GM);
  sub_in(ls.CG, F.CG);
  sub_in(ls.RI, F.RI);
  svg_solve_opts opts = opts_;
  const bool ex = opts.precision != SVG_PRECISION_FP64 || opts.method != SVG_METHOD_PLAIN || opts.coarse > 0;
  const int prec = static_cast<int>(lEq.linear_algebra_preconditioner);
  ok(svg_fsils_solve_ex(lhs_, &ls, 4, dR_, dVal_, prec, incL.size(), incL.size() ? incL.data() : nullptr,
                        res.size() ? res.data() : nullptr, ex ? &opts : nullptr, nullptr),
     "fsils_solve");
  opts_.fallback_itr = opts.fallback_itr;
  sub_out(F.GM, ls.GM);
  sub_out(F.CG, ls.CG);
  sub_out(F.RI, ls.RI);
  F.Resm = ls.Resm;
  F.Resc = ls.Resc;
  // picc reads the increment from com_mod.R (pic.cpp:155-167); Val is left
  // unscaled on the host (ls_alloc clears it before the next assembly)
  hip_ok(hipMemcpy(R.data(), dR_, rBytes, hipMemcpyDeviceToHost), "download R");
  zeroPending_ = true;
}
