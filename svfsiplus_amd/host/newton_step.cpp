// Time steps of the fluid equation on a seeded pipe, driven from C++ over the
// host API (svg_fsils.hpp): the reference's time loop for one fluid equation
// (Code/Source/solver/main.cpp:300-760) —
//   picp -> set_bc_dir (inlet, wall) -> Newton loop { pici -> ls_alloc (R = 0,
//   Val = 0) -> construct_fluid -> b_assem_neu_bc on the outlet (zero traction
//   + backflow stabilisation) -> fsils_solve -> picc } -> Ao = An, Yo = Yn
// with picc's convergence test (pic.cpp:236-281) on the host; the state stays
// on the device throughout.
//
//   svg_newton_step <n> [relTol] [krylov] [maxItr] [steps] [stFileName]
//
// prints one JSON record per Newton iteration (step, itr, ls_itr, iNorm, fNorm, suc).
// The workload: pipe radius 0.3, length 3, n x n cross-section, 5n axial,
// jitter 0.1 h, seed 20261202; rho 1.06, mu 0.04, parabolic inflow peak 30,
// no-slip walls, zero-traction outlet, dt 1e-3, rho_inf 0.5; GMRES with the
// diagonal preconditioner, absolute tolerance 1e-30.
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "svg_fsils.hpp"

using namespace svg_host;

namespace {

bool is_zero(double v)  // utils::is_zero, utils.cpp:170-189, with value2 = 0
{
  const double a = std::fabs(v), eps = DBL_EPSILON;
  return a / std::fmax(a, eps) < 10.0 * eps;
}

}  // namespace

int main(int argc, char** argv)
{
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s <n> [relTol] [krylov] [maxItr]\n", argv[0]);
    return 2;
  }
  const int n = std::atoi(argv[1]);
  const double relTol = argc > 2 ? std::atof(argv[2]) : 1e-6;
  const int sD = argc > 3 ? std::atoi(argv[3]) : 50;
  const int mItr = argc > 4 ? std::atoi(argv[4]) : 1000;
  const int nSteps = argc > 5 ? std::atoi(argv[5]) : 1;
  // restart files after every time step (Increment_in_saving_restart_files = 1), when a name is given
  const std::string stFileName = argc > 6 ? argv[6] : "";
  if (n < 2 || n % 2 || nSteps < 1) {
    std::fprintf(stderr, "usage: %s <n> [relTol] [krylov] [maxItr] [steps]  (n even >= 2, steps >= 1)\n", argv[0]);
    return 2;
  }
  try {
    // mesh and pattern (setup)
    svg_pipe_spec sp{n, 5 * n, 0.3, 3.0, 0.1, 20261202ULL, 0.0, 0.0, 0.0};
    int nNo = 0, nEl = 0, nnz = 0;
    check(svg_pipe_counts(&sp, &nNo, &nEl));
    std::vector<double> x((size_t)3 * nNo);
    std::vector<int> IEN((size_t)4 * nEl), flags(nNo);
    check(svg_pipe_generate(&sp, x.data(), IEN.data(), flags.data()));
    check(svg_csr_pattern_count(nNo, nEl, 4, IEN.data(), &nnz));
    std::vector<int> rowPtr(nNo + 1), colPtr(nnz);
    check(svg_csr_pattern(nNo, nEl, 4, IEN.data(), rowPtr.data(), colPtr.data()));
    std::vector<int> inlet, wall;
    for (int a = 0; a < nNo; a++) {
      if (flags[a] & SVG_NODE_INLET) inlet.push_back(a);
      if (flags[a] & SVG_NODE_WALL) wall.push_back(a);
    }
    FSILS_lhsType lhs;
    fsils_lhs_create(lhs, nNo, nnz, rowPtr, colPtr, 2);
    fsils_bc_create(lhs, 0, (int)inlet.size(), 3, BcType::BC_TYPE_Dir, inlet);
    fsils_bc_create(lhs, 1, (int)wall.size(), 3, BcType::BC_TYPE_Dir, wall);
    FluidMesh msh;
    fluid_mesh_create(msh, lhs, nEl, IEN, x);
    // outlet: every tet face whose three nodes carry the outlet flag
    std::vector<int> fIEN, gE;
    for (int e = 0; e < nEl; e++)
      for (int k = 0; k < 4; k++) {
        int tri[3], m = 0;
        bool out = true;
        for (int a = 0; a < 4; a++)
          if (a != k) {
            tri[m++] = IEN[4 * e + a];
            out = out && (flags[IEN[4 * e + a]] & SVG_NODE_OUTLET);
          }
        if (out) {
          fIEN.insert(fIEN.end(), tri, tri + 3);
          gE.push_back(e);
        }
      }
    FluidFace outlet;
    fluid_face_create(outlet, msh, fIEN, gE);

    // physics and generalized-alpha (initialize.cpp:430-463)
    const double ro = 0.5, peak = 30.0, R0 = 0.3;
    svg_fluid_params p{};
    p.rho = 1.06;
    p.mu = 0.04;
    p.am = 0.5 * (3.0 - ro) / (1.0 + ro);
    p.af = 1.0 / (1.0 + ro);
    p.gam = 0.5 + p.am - p.af;
    p.dt = 1e-3;
    p.qmTET4 = (5.0 + 3.0 * std::sqrt(5.0)) / 20.0;

    // strong Dirichlet faces (steady: dirA = 0, dirY = 1 with gx nV holding the
    // values): the inlet parabola u_z = peak (1 - r^2/R^2), then the wall (u = 0)
    std::vector<double> gIn((size_t)3 * inlet.size(), 0.0), gWall((size_t)3 * wall.size(), 0.0);
    for (size_t k = 0; k < inlet.size(); k++) {
      const int a = inlet[k];
      const double r2 = (x[3 * a] * x[3 * a] + x[3 * a + 1] * x[3 * a + 1]) / (R0 * R0);
      gIn[3 * k + 2] = peak * (1.0 - r2);
    }
    DirichletFace dirIn, dirWall;
    dirichlet_face_create(dirIn, inlet, 3, gIn);
    dirichlet_face_create(dirWall, wall, 3, gWall);

    // initial state Ao = Yo = 0
    const std::vector<double> zero((size_t)4 * nNo, 0.0);
    DeviceArray<double> dAo, dYo, dAn((size_t)4 * nNo), dYn((size_t)4 * nNo), dAg((size_t)4 * nNo),
        dYg((size_t)4 * nNo), dR((size_t)4 * nNo), dVal((size_t)16 * nnz);
    dAo.upload(zero);
    dYo.upload(zero);

    FSILS_lsType ls;
    fsils_ls_create(ls, LinearSolverType::LS_TYPE_GMRES);
    ls.RI.relTol = relTol;
    ls.RI.absTol = 1e-30;
    ls.RI.mItr = mItr;
    ls.RI.sD = sD;

    // Newton loop with picc's convergence test (minItr = maxItr = 3, tol 0.5)
    const int minIt = 3, maxIt = 3;
    const double tol = 0.5;
    double eqINorm = 0.0, pNorm = DBL_MAX;
    const std::vector<int> incL(2, 1);
    const std::vector<double> res(2, 0.0);
    std::printf("{\"nNo\": %d, \"nEl\": %d, \"nnz\": %d, \"colours\": %d}\n", nNo, nEl, nnz, msh.nColors);
    for (int step = 1; step <= nSteps; step++) {
      // predictor and strong Dirichlet values (main.cpp:346-360)
      picp(nNo, p.gam, dAo.data(), dYo.data(), dAn.data(), dYn.data());
      set_bc_dir(dirIn, 4, 0, 0.0, 1.0, dAn.data(), dYn.data());
      set_bc_dir(dirWall, 4, 0, 0.0, 1.0, dAn.data(), dYn.data());
      for (int itr = 1;; itr++) {
        pici(nNo, p.am, p.af, dAo.data(), dAn.data(), dYo.data(), dYn.data(), dAg.data(), dYg.data());
        check(svg_fill(dR.data(), dR.size(), 0.0, nullptr));
        check(svg_fill(dVal.data(), dVal.size(), 0.0, nullptr));
        construct_fluid(msh, p, dAg.data(), dYg.data(), nullptr, dR.data(), dVal.data());
        b_assem_neu_bc(outlet, p, 0.2, 2.0 / 3.0, nullptr, dYg.data(), dR.data(), dVal.data());
        fsils_solve(lhs, ls, 4, dR.data(), dVal.data(), PreconditionerType::PREC_FSILS, incL, res);
        picc_update(nNo, p.gam, p.dt, dR.data(), dAn.data(), dYn.data());
        if (is_zero(ls.RI.iNorm)) ls.RI.iNorm = DBL_EPSILON;
        if (is_zero(eqINorm)) eqINorm = ls.RI.iNorm;
        if (itr == 1) pNorm = ls.RI.iNorm / eqINorm;
        const double r1 = ls.RI.iNorm / eqINorm;
        const bool ok = itr >= maxIt || ((r1 <= tol || r1 <= tol * pNorm) && itr >= minIt);
        std::printf(
            "{\"step\": %d, \"itr\": %d, \"ls_itr\": %d, \"iNorm\": %.17g, \"fNorm\": %.17g, \"suc\": %d, "
            "\"r1\": %.17g}\n",
            step, itr, ls.RI.itr, ls.RI.iNorm, ls.RI.fNorm, ls.RI.suc, r1);
        std::fflush(stdout);
        if (ok) break;
      }
      if (!stFileName.empty())  // main.cpp:695-708
        write_restart(stFileName, step, step * p.dt, 0.0, eqINorm, nNo, dYn.data(), dAn.data());
      // Ao = An, Yo = Yn (main.cpp:754-755)
      dAo.copy_from(dAn);
      dYo.copy_from(dYn);
    }
  } catch (const std::exception& e) {
    std::fprintf(stderr, "%s\n", e.what());
    return 1;
  }
  return 0;
}
