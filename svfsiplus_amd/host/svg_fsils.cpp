// C++ host API of the fluid Newton step over the C ABI (see svg_fsils.hpp).
#include "svg_fsils.hpp"

#include <unistd.h>

#include <cstdio>

namespace svg_host {

void check(int rc)
{
  if (rc != SVG_OK) throw std::runtime_error(std::string("[svFSIplus-gpu] ") + svg_last_error());
}

FSILS_lhsType::~FSILS_lhsType()
{
  if (h) svg_lhs_free(h);
}

void fsils_lhs_create(FSILS_lhsType& lhs, int nNo, int nnz, const std::vector<int>& rowPtr,
                      const std::vector<int>& colPtr, int nFaces)
{
  if ((int)rowPtr.size() != nNo + 1 || (int)colPtr.size() != nnz)
    throw std::runtime_error("fsils_lhs_create: rowPtr/colPtr sizes do not match nNo/nnz");
  if (lhs.h) fsils_lhs_free(lhs);
  check(svg_lhs_create(nNo, nnz, rowPtr.data(), colPtr.data(), nFaces, &lhs.h));
  lhs.nNo = nNo;
  lhs.nnz = nnz;
  lhs.nFaces = nFaces;
}

void fsils_bc_create(FSILS_lhsType& lhs, int faIn, int nNo, int dof, BcType BC_type, const std::vector<int>& gNodes,
                     const std::vector<double>& Val)
{
  if ((int)gNodes.size() != nNo) throw std::runtime_error("fsils_bc_create: gNodes size differs from nNo");
  std::vector<double> v = Val;
  if (v.empty()) v.assign((size_t)dof * nNo, 0.0);
  if (v.size() != (size_t)dof * nNo) throw std::runtime_error("fsils_bc_create: Val must be (dof, nNo)");
  check(svg_lhs_bc_create(lhs.h, faIn, nNo, dof, (int)BC_type, gNodes.data(), v.data()));
}

void fsils_ls_create(FSILS_lsType& ls, LinearSolverType LS_type) { check(svg_ls_create(&ls, (int)LS_type)); }

void fsils_solve(FSILS_lhsType& lhs, FSILS_lsType& ls, int dof, double* Ri, double* Val, PreconditionerType prec,
                 const std::vector<int>& incL, const std::vector<double>& res, hipStream_t stream)
{
  check(svg_fsils_solve(lhs.h, &ls, dof, Ri, Val, (int)prec, (int)incL.size(), incL.empty() ? nullptr : incL.data(),
                        res.empty() ? nullptr : res.data(), stream));
}

void fsils_lhs_free(FSILS_lhsType& lhs)
{
  if (lhs.h) svg_lhs_free(lhs.h);
  lhs.h = nullptr;
  lhs.nNo = lhs.nnz = lhs.nFaces = 0;
}

FluidMesh::~FluidMesh()
{
  if (h) svg_msh_free(h);
}

void fluid_mesh_create(FluidMesh& msh, const FSILS_lhsType& lhs, int nEl, const std::vector<int>& IEN,
                       const std::vector<double>& x)
{
  if (IEN.size() != (size_t)4 * nEl || x.size() != (size_t)3 * lhs.nNo)
    throw std::runtime_error("fluid_mesh_create: IEN must be (4,nEl) and x (3,nNo)");
  if (msh.h) svg_msh_free(msh.h);
  check(svg_msh_create(lhs.h, nEl, IEN.data(), lhs.nNo, x.data(), &msh.h));
  msh.nEl = nEl;
  check(svg_msh_num_colors(msh.h, &msh.nColors));
}

FluidFace::~FluidFace()
{
  if (h) svg_face_free(h);
}

void fluid_face_create(FluidFace& fa, const FluidMesh& msh, const std::vector<int>& faceIEN,
                       const std::vector<int>& gE)
{
  if (faceIEN.size() != 3 * gE.size()) throw std::runtime_error("fluid_face_create: faceIEN must be (3,nFe)");
  if (fa.h) svg_face_free(fa.h);
  fa.h = nullptr;
  check(svg_face_create(msh.h, (int)gE.size(), faceIEN.data(), gE.data(), &fa.h));
  fa.nFe = (int)gE.size();
}

void b_assem_neu_bc(FluidFace& fa, const svg_fluid_params& p, double beta, double qmTRI3, const double* hg,
                    const double* Yg, double* R, double* Val, hipStream_t stream)
{
  check(svg_b_assem_neu_bc(fa.h, &p, beta, qmTRI3, hg, Yg, R, Val, stream));
}

void construct_fluid(FluidMesh& msh, const svg_fluid_params& p, const double* Ag, const double* Yg, const double* Bf,
                     double* R, double* Val, hipStream_t stream)
{
  check(svg_construct_fluid(msh.h, &p, Ag, Yg, Bf, R, Val, stream));
}

void pici(int nNo, double am, double af, const double* Ao, const double* An, const double* Yo, const double* Yn,
          double* Ag, double* Yg, hipStream_t stream)
{
  check(svg_pici(nNo, am, af, Ao, An, Yo, Yn, Ag, Yg, stream));
}

void picc_update(int nNo, double gam, double dt, const double* R, double* An, double* Yn, hipStream_t stream)
{
  check(svg_picc_update(nNo, gam, dt, R, An, Yn, stream));
}

void picp(int nNo, double gam, const double* Ao, const double* Yo, double* An, double* Yn, hipStream_t stream)
{
  check(svg_picp(nNo, gam, Ao, Yo, An, Yn, stream));
}

void dirichlet_face_create(DirichletFace& f, const std::vector<int>& gN, int lDof, const std::vector<double>& gxn,
                           unsigned eDir)
{
  if (lDof < 1 || gxn.size() != (size_t)lDof * gN.size())
    throw std::runtime_error("dirichlet_face_create: gxn must be (lDof, nNo)");
  f.gN.upload(gN);
  f.gxn.upload(gxn);
  f.nNo = (int)gN.size();
  f.lDof = lDof;
  f.eDir = eDir;
}

void set_bc_dir(const DirichletFace& f, int tDof, int s, double dirA, double dirY, double* A, double* Y,
                hipStream_t stream)
{
  check(svg_set_bc_dir(f.nNo, f.gN.data(), tDof, s, f.lDof, f.eDir, f.gxn.data(), dirA, dirY, A, Y, stream));
}

std::string write_restart(const std::string& stFileName, int cTS, double time, double cpuTime, double iNorm,
                          int nNo, const double* d_Yn, const double* d_An, bool replace)
{
  const int stamp[7] = {1, 1, 1, nNo, 0, 4, 0};  // ranks, nEq, nMsh, tnNo, nX, tDof, dFlag (initialize.cpp:497)
  std::vector<double> Y((size_t)4 * nNo), A((size_t)4 * nNo);
  if (hipMemcpy(Y.data(), d_Yn, sizeof(double) * Y.size(), hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(A.data(), d_An, sizeof(double) * A.size(), hipMemcpyDeviceToHost) != hipSuccess)
    throw std::runtime_error("write_restart: device copy failed");
  char num[32];
  std::snprintf(num, sizeof num, cTS >= 1000 ? "%d" : "%03d", cTS);
  const std::string last = stFileName + "_last.bin";
  const std::string fName = replace ? last : stFileName + "_" + num + ".bin";
  std::FILE* f = std::fopen(fName.c_str(), "wb");
  if (!f) throw std::runtime_error("write_restart: cannot open " + fName);
  const double td[2] = {time, cpuTime};
  bool ok = std::fwrite(stamp, sizeof(int), 7, f) == 7 && std::fwrite(&cTS, sizeof(int), 1, f) == 1 &&
            std::fwrite(td, sizeof(double), 2, f) == 2 && std::fwrite(&iNorm, sizeof(double), 1, f) == 1 &&
            std::fwrite(Y.data(), sizeof(double), Y.size(), f) == Y.size() &&
            std::fwrite(A.data(), sizeof(double), A.size(), f) == A.size();
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) throw std::runtime_error("write_restart: short write to " + fName);
  if (!replace) {  // output.cpp:353-358: ln -f <file> <stFileName>_last.bin
    std::remove(last.c_str());
    if (link(fName.c_str(), last.c_str()) != 0) throw std::runtime_error("write_restart: cannot link " + last);
  }
  return fName;
}

}  // namespace svg_host
