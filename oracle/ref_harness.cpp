// Reference harness (test infrastructure only, never shipped with the
// product): drives the reference's own construct_fluid and fsils_solve,
// compiled from the sources under /root/reference by oracle/ref.mk, through
// a plain C interface so that tests can compare the oracle restatement (and,
// via committed golden fixtures, the HIP path) with the real reference.
//
// Objects are built the way the reference's own driver builds them for a
// single-rank run:
//   mesh shape functions   nn::select_ele + fs::init_fs_msh (read_msh.cpp)
//   equation coefficients  initialize.cpp:424-463 (values passed in)
//   FSILS lhs / faces / ls fsils_lhs_create, fsils_bc_create, fsils_ls_create
//                          then the <LS> overrides of read_files.cpp:2056-2090
#include "ComMod.h"
#include "FsilsLinearAlgebra.h"
#include "Simulation.h"
#include "eq_assem.h"
#include "fluid.h"
#include "fs.h"
#include "fsils_api.hpp"
#include "lhs.h"
#include "nn.h"
#include "output.h"
#include "utils.h"

#include <mpi.h>

#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>

static std::string g_ref_err;

extern "C" const char* ref_last_error(void) { return g_ref_err.c_str(); }

// prm: rho, mu, fx, fy, fz, K_darcy, af, am, gam, dt
extern "C" int ref_construct_fluid(int nEl, const int* IEN, int nNo, const double* x, const double* Bf,
                                   const double* Ag, const double* Yg, const double* prm, int nnz,
                                   const int* rowPtr, const int* colPtr, double* R, double* Val)
{
  using namespace consts;
  try {
    ComMod com_mod;
    com_mod.nsd = 3;
    com_mod.nsymd = 6;
    com_mod.dof = 4;
    com_mod.tDof = 4;
    com_mod.tnNo = nNo;
    com_mod.cEq = 0;
    com_mod.mvMsh = false;
    com_mod.dt = prm[9];
    com_mod.x.resize(3, nNo);
    com_mod.Bf.resize(3, nNo);
    for (int a = 0; a < nNo; a++)
      for (int i = 0; i < 3; i++) {
        com_mod.x(i, a) = x[i + 3 * a];
        com_mod.Bf(i, a) = Bf ? Bf[i + 3 * a] : 0.0;
      }
    com_mod.rowPtr.resize(nNo + 1);
    com_mod.colPtr.resize(nnz);
    for (int i = 0; i <= nNo; i++) com_mod.rowPtr(i) = rowPtr[i];
    for (int i = 0; i < nnz; i++) com_mod.colPtr(i) = colPtr[i];
    com_mod.R.resize(4, nNo);
    com_mod.Val.resize(16, nnz);
    std::memcpy(com_mod.R.data(), R, sizeof(double) * 4 * nNo);
    std::memcpy(com_mod.Val.data(), Val, sizeof(double) * 16 * (size_t)nnz);

    com_mod.eq.resize(1);
    auto& eq = com_mod.eq[0];
    eq.af = prm[6];
    eq.am = prm[7];
    eq.gam = prm[8];
    eq.dof = 4;
    eq.nDmn = 1;
    eq.dmn.resize(1);
    auto& dmn = eq.dmn[0];
    dmn.Id = -1;
    dmn.phys = EquationType::phys_fluid;
    dmn.prop[PhysicalProperyType::fluid_density] = prm[0];
    dmn.prop[PhysicalProperyType::f_x] = prm[2];
    dmn.prop[PhysicalProperyType::f_y] = prm[3];
    dmn.prop[PhysicalProperyType::f_z] = prm[4];
    dmn.prop[PhysicalProperyType::inverse_darcy_permeability] = prm[5];
    dmn.fluid_visc.viscType = FluidViscosityModelType::viscType_Const;
    dmn.fluid_visc.mu_i = prm[1];
    FsilsLinearAlgebra la;
    eq.linear_algebra = &la;

    mshType msh;
    msh.nEl = nEl;
    msh.gnEl = nEl;
    msh.eNoN = 4;
    msh.nFs = 1;
    msh.IEN.resize(4, nEl);
    for (int e = 0; e < nEl; e++)
      for (int a = 0; a < 4; a++) msh.IEN(a, e) = IEN[a + 4 * e];
    nn::select_ele(com_mod, msh);
    fs::init_fs_msh(com_mod, msh);

    fluid::construct_fluid(com_mod, msh, Array<double>(4, nNo, const_cast<double*>(Ag)),
                           Array<double>(4, nNo, const_cast<double*>(Yg)));
    std::memcpy(R, com_mod.R.data(), sizeof(double) * 4 * nNo);
    std::memcpy(Val, com_mod.Val.data(), sizeof(double) * 16 * (size_t)nnz);
    eq.linear_algebra = nullptr;
  } catch (const std::exception& ex) {
    g_ref_err = ex.what();
    return 1;
  }
  return 0;
}

// ls_prm: RI.relTol, RI.absTol, RI.mItr, RI.sD, GM.relTol, GM.absTol, GM.mItr, GM.sD,
//         CG.relTol, CG.absTol, CG.mItr
// out:    RI.itr, RI.iNorm, RI.fNorm, RI.suc, RI.dB, GM.itr, CG.itr, Resm, Resc
// faceBGrp (null: every face Dirichlet) gives each face's group (0 Dir, 1 Neu)
// and res (null: zeros) the resistance passed to fsils_solve; faces are
// registered with dof 3 and val(3, nNo) = faceVal.
extern "C" int ref_fsils_solve_bc(int nNo, int nnz, const int* rowPtr, const int* colPtr, int nFaces,
                                  const int* faceNNo, const int* faceGlob, const double* faceVal,
                                  const int* faceBGrp, const double* faceRes, int LS_type, const double* ls_prm,
                                  int dof, double* Ri, double* Val, int prec, double* out)
{
  using namespace fsi_linear_solver;
  try {
    FSILS_commuType commu;
    commu.foC = true;
    commu.masF = true;
    commu.master = 0;
    commu.task = 0;
    commu.tF = 1;
    commu.nTasks = 1;
    FSILS_lhsType lhs;
    Vector<int> gNodes(nNo), vRow(nNo + 1), vCol(nnz);
    for (int a = 0; a < nNo; a++) gNodes(a) = a;
    for (int a = 0; a <= nNo; a++) vRow(a) = rowPtr[a];
    for (int i = 0; i < nnz; i++) vCol(i) = colPtr[i];
    fsils_lhs_create(lhs, commu, nNo, nNo, nnz, gNodes, vRow, vCol, nFaces);
    int off = 0;
    for (int f = 0; f < nFaces; f++) {
      int n = faceNNo[f];
      Vector<int> g(n);
      Array<double> v(3, n);
      for (int a = 0; a < n; a++) {
        g(a) = faceGlob[off + a];
        for (int i = 0; i < 3; i++) v(i, a) = faceVal[3 * (off + a) + i];
      }
      const BcType bt = (faceBGrp && faceBGrp[f] == 1) ? BcType::BC_TYPE_Neu : BcType::BC_TYPE_Dir;
      fsils_bc_create(lhs, f, n, 3, bt, g, v);
      off += n;
    }
    FSILS_lsType ls;
    fsils_ls_create(ls, static_cast<LinearSolverType>(LS_type));
    ls.RI.relTol = ls_prm[0];
    ls.RI.absTol = ls_prm[1];
    ls.RI.mItr = (int)ls_prm[2];
    ls.RI.sD = (int)ls_prm[3];
    ls.GM.relTol = ls_prm[4];
    ls.GM.absTol = ls_prm[5];
    ls.GM.mItr = (int)ls_prm[6];
    ls.GM.sD = (int)ls_prm[7];
    ls.CG.relTol = ls_prm[8];
    ls.CG.absTol = ls_prm[9];
    ls.CG.mItr = (int)ls_prm[10];

    Array<double> aRi(dof, nNo), aVal(dof * dof, nnz);
    std::memcpy(aRi.data(), Ri, sizeof(double) * dof * nNo);
    std::memcpy(aVal.data(), Val, sizeof(double) * dof * dof * (size_t)nnz);
    Vector<int> incL(nFaces);
    Vector<double> res(nFaces);
    for (int f = 0; f < nFaces; f++) {
      incL(f) = 1;
      res(f) = faceRes ? faceRes[f] : 0.0;
    }
    fsils_solve(lhs, ls, dof, aRi, aVal,
                prec == 701 ? consts::PreconditionerType::PREC_FSILS : consts::PreconditionerType::PREC_NONE,
                incL, res);
    std::memcpy(Ri, aRi.data(), sizeof(double) * dof * nNo);
    std::memcpy(Val, aVal.data(), sizeof(double) * dof * dof * (size_t)nnz);
    out[0] = ls.RI.itr;
    out[1] = ls.RI.iNorm;
    out[2] = ls.RI.fNorm;
    out[3] = ls.RI.suc ? 1.0 : 0.0;
    out[4] = ls.RI.dB;
    out[5] = ls.GM.itr;
    out[6] = ls.CG.itr;
    out[7] = ls.Resm;
    out[8] = ls.Resc;
  } catch (const std::exception& ex) {
    g_ref_err = ex.what();
    return 1;
  }
  return 0;
}

extern "C" int ref_fsils_solve(int nNo, int nnz, const int* rowPtr, const int* colPtr, int nFaces,
                               const int* faceNNo, const int* faceGlob, const double* faceVal, int LS_type,
                               const double* ls_prm, int dof, double* Ri, double* Val, int prec, double* out)
{
  return ref_fsils_solve_bc(nNo, nnz, rowPtr, colPtr, nFaces, faceNNo, faceGlob, faceVal, nullptr, nullptr, LS_type,
                            ls_prm, dof, Ri, Val, prec, out);
}

// fsils_ls_create (ls.cpp:40-120) with no optional argument: the fields it sets
// for LS_type. Each field starts at the sentinel -7777 (integers) or -7777.0,
// so a field the reference leaves unset comes back as the sentinel.
// out (16 doubles): RI.{relTol, absTol, mItr, sD}, GM.{...}, CG.{...},
//                   then LS_type, foC
extern "C" int ref_ls_create(int LS_type, double* out)
{
  using namespace fsi_linear_solver;
  try {
    FSILS_lsType ls;
    for (FSILS_subLsType* s : {&ls.RI, &ls.GM, &ls.CG}) {
      s->relTol = -7777.0;
      s->absTol = -7777.0;
      s->mItr = -7777;
      s->sD = -7777;
    }
    fsils_ls_create(ls, static_cast<LinearSolverType>(LS_type));
    int k = 0;
    for (const FSILS_subLsType* s : {&ls.RI, &ls.GM, &ls.CG}) {
      out[k++] = s->relTol;
      out[k++] = s->absTol;
      out[k++] = s->mItr;
      out[k++] = s->sD;
    }
    out[k++] = static_cast<int>(ls.LS_type);
    out[k++] = ls.foC ? 1.0 : 0.0;
  } catch (const std::exception& ex) {
    g_ref_err = ex.what();
    return 1;
  }
  return 0;
}

// The fluid equation's Neumann-face assembly through the reference's own
// eq_assem::b_assem_neu_bc (eq_assem.cpp:58-172: nn::gnnb for the outward
// normal, the TRI3 face rule of nn_elem_gip.h, fluid::b_fluid fluid.cpp:46-133
// and assembly through FsilsLinearAlgebra / do_assem) on one face of TRI3
// elements of a TET4 mesh. faceIEN(3, nFe) node ids, gE(nFe) the parent
// element of each face element (index into IEN), hg(nNo) the Neumann value
// per node (null: zero), Yg(4, nNo). prm as for ref_construct_fluid; beta the
// backflow stabilisation coefficient, qmTRI3 the face quadrature modifier.
// R(4, nNo) and Val(16, nnz) are accumulated into.
extern "C" int ref_b_assem_neu_bc(int nEl, const int* IEN, int nNo, const double* x, int nFe, const int* faceIEN,
                                  const int* gE, const double* hg, const double* Yg, const double* prm, double beta,
                                  double qmTRI3, int nnz, const int* rowPtr, const int* colPtr, double* R,
                                  double* Val)
{
  using namespace consts;
  try {
    ComMod com_mod;
    com_mod.nsd = 3;
    com_mod.nsymd = 6;
    com_mod.dof = 4;
    com_mod.tDof = 4;
    com_mod.tnNo = nNo;
    com_mod.cEq = 0;
    com_mod.mvMsh = false;
    com_mod.dt = prm[9];
    com_mod.x.resize(3, nNo);
    for (int a = 0; a < nNo; a++)
      for (int i = 0; i < 3; i++) com_mod.x(i, a) = x[i + 3 * a];
    com_mod.rowPtr.resize(nNo + 1);
    com_mod.colPtr.resize(nnz);
    for (int i = 0; i <= nNo; i++) com_mod.rowPtr(i) = rowPtr[i];
    for (int i = 0; i < nnz; i++) com_mod.colPtr(i) = colPtr[i];
    com_mod.R.resize(4, nNo);
    com_mod.Val.resize(16, nnz);
    std::memcpy(com_mod.R.data(), R, sizeof(double) * 4 * nNo);
    std::memcpy(com_mod.Val.data(), Val, sizeof(double) * 16 * (size_t)nnz);

    com_mod.eq.resize(1);
    auto& eq = com_mod.eq[0];
    eq.af = prm[6];
    eq.am = prm[7];
    eq.gam = prm[8];
    eq.dof = 4;
    eq.nDmn = 1;
    eq.dmn.resize(1);
    auto& dmn = eq.dmn[0];
    dmn.Id = -1;
    dmn.phys = EquationType::phys_fluid;
    dmn.prop[PhysicalProperyType::fluid_density] = prm[0];
    dmn.prop[PhysicalProperyType::backflow_stab] = beta;
    dmn.fluid_visc.viscType = FluidViscosityModelType::viscType_Const;
    dmn.fluid_visc.mu_i = prm[1];
    FsilsLinearAlgebra la;
    eq.linear_algebra = &la;

    com_mod.nMsh = 1;
    com_mod.msh.resize(1);
    mshType& msh = com_mod.msh[0];
    msh.name = "volume";
    msh.nEl = nEl;
    msh.gnEl = nEl;
    msh.eNoN = 4;
    msh.nFs = 1;
    msh.IEN.resize(4, nEl);
    for (int e = 0; e < nEl; e++)
      for (int a = 0; a < 4; a++) msh.IEN(a, e) = IEN[a + 4 * e];
    nn::select_ele(com_mod, msh);
    fs::init_fs_msh(com_mod, msh);

    // the face as select_eleb sets it up for a 3-noded face of a 3D mesh
    // (nn_elem_props.h: TRI3, nG = 3); get_gip/get_gnn of a face use only the face
    faceType lFa;
    lFa.name = "outlet";
    lFa.iM = 0;
    lFa.nEl = nFe;
    lFa.gnEl = nFe;
    lFa.eNoN = 3;
    lFa.eType = ElementType::TRI3;
    lFa.nG = 3;
    lFa.qmTRI3 = qmTRI3;
    lFa.IEN.resize(3, nFe);
    lFa.gE.resize(nFe);
    for (int e = 0; e < nFe; e++) {
      for (int a = 0; a < 3; a++) lFa.IEN(a, e) = faceIEN[a + 3 * e];
      lFa.gE(e) = gE[e];
    }
    lFa.w = Vector<double>(lFa.nG);
    lFa.xi = Array<double>(2, lFa.nG);
    nn::get_gip(nullptr, lFa);
    lFa.N = Array<double>(lFa.eNoN, lFa.nG);
    lFa.Nx = Array3<double>(2, lFa.eNoN, lFa.nG);
    for (int g = 0; g < lFa.nG; g++) nn::get_gnn(nullptr, g, lFa);

    Vector<double> vhg(nNo);
    for (int a = 0; a < nNo; a++) vhg(a) = hg ? hg[a] : 0.0;
    eq_assem::b_assem_neu_bc(com_mod, lFa, vhg, Array<double>(4, nNo, const_cast<double*>(Yg)));
    std::memcpy(R, com_mod.R.data(), sizeof(double) * 4 * nNo);
    std::memcpy(Val, com_mod.Val.data(), sizeof(double) * 16 * (size_t)nnz);
    eq.linear_algebra = nullptr;
  } catch (const std::exception& ex) {
    g_ref_err = ex.what();
    return 1;
  }
  return 0;
}

// The reference's own write_restart (output.cpp:202-360, with
// write_restart_header :361-376) for the fluid equation on one rank: a
// Simulation whose ComMod holds the given stamp, time step, time, per-equation
// iNorm, cplBC.xn and Yn/An (tDof, tnNo), no moving mesh, no structure,
// prestress, CEP or immersed-body fields. Writes <stFileName>_<cTS>.bin and its
// _last.bin link (or _last.bin alone with stFileRepl), as the reference does.
// The header's CPU-time field is utils::cput() - timeP[0] at the write.
extern "C" int ref_write_restart(const char* stFileName, int stFileRepl, int cTS, double time, const int* stamp,
                                 int nEq, const double* iNorm, int nX, const double* xn, int tnNo, int tDof,
                                 const double* Yn, const double* An, int recLn)
{
  try {
    int inited = 0;
    MPI_Initialized(&inited);
    if (!inited) MPI_Init(nullptr, nullptr);
    Simulation sim;
    ComMod& com = sim.com_mod;
    com.stFileName = stFileName;
    com.stFileRepl = stFileRepl != 0;
    com.cTS = cTS;
    com.time = time;
    for (int i = 0; i < 7; i++) com.stamp[i] = stamp[i];
    com.recLn = recLn;
    com.eq.resize(nEq);
    for (int i = 0; i < nEq; i++) com.eq[i].iNorm = iNorm[i];
    com.cplBC.xn.resize(nX);
    for (int i = 0; i < nX; i++) com.cplBC.xn(i) = xn[i];
    com.Yn.resize(tDof, tnNo);
    com.An.resize(tDof, tnNo);
    for (int a = 0; a < tnNo; a++)
      for (int i = 0; i < tDof; i++) {
        com.Yn(i, a) = Yn[i + tDof * a];
        com.An(i, a) = An[i + tDof * a];
      }
    com.ibFlag = false;
    com.dFlag = false;
    com.sstEq = false;
    com.pstEq = false;
    sim.cep_mod.cepEq = false;
    std::array<double, 3> timeP{utils::cput(), 0.0, 0.0};
    output::write_restart(&sim, timeP);
  } catch (const std::exception& ex) {
    g_ref_err = ex.what();
    return 1;
  }
  return 0;
}
