/*
 * CPU oracle for the svFSIplus fluid Newton step — TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C, single-threaded restatement of the reference algorithm
 * (SimVascular svFSIplus / svMultiPhysics) for the path
 *   construct_fluid -> fluid_3d_m / fluid_3d_c -> do_assem
 *   fsils_solve -> precond_diag -> gmres_v | ns_solver(gmres, schur, ge)
 *   pici / picc.
 * Loop order and summation order follow the reference line by line, so the
 * oracle is the checker for the HIP path and the single-core CPU baseline.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it; the product library never links it.
 *
 * Layouts are the reference's (column-major Array<T>, Code/Source/solver/Array.h:374-388):
 *   x(3,nNo)        x[i + 3*a]
 *   IEN(4,nEl)      IEN[a + 4*e]
 *   Ag/Yg/R(4,nNo)  R[i + 4*a]
 *   Val(16,nnz)     Val[k + 16*j], k = 4*row + col inside the block
 *   rowPtr(nNo+1)   exclusive-end CSR (com_mod.rowPtr); the FSILS view is
 *                   rowPtr2(2,nNo) inclusive [s,e] (lhs.cpp:113-114).
 *
 * Pinning (how each function is checked; see tests/test_oracle_*.py):
 *   orc_construct_fluid : against the reference's own construct_fluid
 *                         compiled from its sources (oracle/_ref) on seeded
 *                         pipes, committed as tests/golden fixtures; plus
 *                         closed forms (sum w = 1/6, Jac = 6 vol, sum Nx = 0,
 *                         u=0 tauM) and a finite-difference Stokes-limit test.
 *   orc_precond_diag    : fixtures + identity checks (unit diagonal off
 *                         Dirichlet rows, zero Dirichlet rows/cols).
 *   orc_spar_mul_*      : dense numpy mat-vec on the expanded matrix.
 *   orc_dot/norm        : numpy with long-double accumulation.
 *   orc_gmres_v         : fixtures (solution + iteration count) + dense LU.
 *   orc_ns_solver       : fixtures (solution, RI.itr, Resm/Resc) + dense LU.
 *   orc_ge              : numpy.linalg.solve.
 *   orc_pici/picc/picp  : closed-form generalized-alpha algebra.
 *   orc_set_bc_dir      : the first time step's state (Ao = Yo = 0) equals
 *                         the closed-form Dirichlet state of the workload;
 *                         direction masks against hand-placed values.
 */
#ifndef SVFSI_ORACLE_H
#define SVFSI_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Physical and time-integration parameters for one fluid domain. */
typedef struct {
  double rho;        /* fluid_density */
  double mu;         /* constant viscosity (viscType_Const) */
  double f[3];       /* body force f_x, f_y, f_z */
  double K_darcy;    /* inverse_darcy_permeability */
  double af, am, gam;/* generalized-alpha coefficients (initialize.cpp:430-463) */
  double dt;
  double qmTET4;     /* TET4 quadrature modifier (ComMod.h:1011) */
} orc_fluid_params;

/* FSILS sub-solver record (fils_struct.hpp:216-263). */
typedef struct {
  int suc, mItr, sD, itr, cM, cN, cD;
  double absTol, relTol, iNorm, fNorm, dB, callD;
} orc_subls;

enum { ORC_LS_NS = 796, ORC_LS_GMRES = 797, ORC_LS_CG = 798 };

typedef struct {
  int LS_type;
  int Resm, Resc;
  orc_subls GM, CG, RI;
} orc_ls;

/* A face registered with FSILS (fsils_bc_create, bc.cpp:45-130). */
typedef struct {
  int nNo, dof, bGrp;  /* bGrp: 0 = Dir, 1 = Neu */
  int incFlag, coupledFlag;
  const int* glob;     /* nNo node ids */
  const double* val;   /* (dof, nNo) */
  double res;          /* resistance of a coupled Neumann face (set by fsils_solve) */
  double* valM;        /* (dof, nNo): val scaled by the preconditioner (caller-owned, zero at creation) */
  double nS;           /* |valM|^2 over the face (the NS solver's bc_pre, ns_solver.cpp:56-89) */
} orc_face;

typedef struct {
  int nNo, nnz;
  const int* rowPtr2;  /* (2,nNo) inclusive [s,e] */
  const int* colPtr;
  const int* diagPtr;
  int nFaces;
  orc_face* face;
} orc_lhs;

/* Error reporting: functions return 0 on success, non-zero on the same
 * conditions on which the reference throws std::runtime_error. */
const char* orc_last_error(void);

/* ---- element routines (TET4, nsd = 3) ---- */
void orc_tet4_shape(double qmTET4, double w[4], double N[4][4], double Nxi[3][4]);
int  orc_gnn_tet4(const double Nxi[3][4], const double xl[3][4], double Nx[3][4],
                  double* Jac, double ks[3][3]);

/* construct_fluid (fluid.cpp:464-708): accumulates into R(4,nNo), Val(16,nnz). */
int orc_construct_fluid(int nEl, const int* IEN, const double* x, const double* Bf,
                        const double* Ag, const double* Yg, const orc_fluid_params* p,
                        const int* rowPtr, const int* colPtr, double* R, double* Val);

/* Neumann-face assembly with backflow stabilisation on TRI3 faces
 * (b_assem_neu_bc eq_assem.cpp:58-172 + gnnb nn.cpp:552-760 + b_fluid
 * fluid.cpp:46-133): faceIEN(3,nFe), inner(nFe) = parent tet node off the
 * face, hg(nNo) or NULL for zero traction, beta = backflow_stab. */
int orc_b_assem_neu_fluid(int nFe, const int* faceIEN, const int* inner, const double* x, const double* hg,
                          const double* Yg, const orc_fluid_params* p, double beta, double qmTRI3,
                          const int* rowPtr, const int* colPtr, double* R, double* Val);

/* Element-local residual/tangent of one element (for unit tests).
 * lR(4,4): lR[i + 4a]; lK(16,4,4): lK[k + 16*(a + 4b)]. */
int orc_fluid_element(const double xl[3][4], const double al[4][4], const double yl[4][4],
                      const double bfl[3][4], const orc_fluid_params* p,
                      double lR[16], double lK[256]);

/* ---- FSILS pieces ---- */
void orc_spar_mul_vv(const orc_lhs* lhs, int dof, const double* K, const double* U, double* KU);
/* add_bc_mul (add_bc_mul.cpp:53-121), op ADD: Y += res valM (valM^T X) for each coupled face. */
void orc_add_bc_mul(const orc_lhs* lhs, int dof, const double* X, double* Y);
/* add_bc_mul op PRE (add_bc_mul.cpp:60-64): Y += coef valM (valM^T X) with
 * coef = -res / (1 + res nS), face after face; X may be Y (in place). */
void orc_add_bc_pre(const orc_lhs* lhs, int dof, const double* X, double* Y);
void orc_spar_mul_sv(const orc_lhs* lhs, int dof, const double* K, const double* U, double* KU);
void orc_spar_mul_vs(const orc_lhs* lhs, int dof, const double* K, const double* U, double* KU);
void orc_spar_mul_ss(const orc_lhs* lhs, const double* K, const double* U, double* KU);
double orc_dot_v(int dof, int nNo, const double* U, const double* V);
double orc_dot_s(int nNo, const double* U, const double* V);
double orc_norm_v(int dof, int nNo, const double* U);
double orc_norm_s(int nNo, const double* U);
/* 0: the reference's serial sums (default); 1: the HIP library's reduction order. */
void orc_set_dot_order(int order);
int orc_get_dot_order(void);
void orc_precond_diag(const orc_lhs* lhs, int dof, double* Val, double* R, double* W);
int  orc_ge(int N, int ldA, const double* A, double* B);
void orc_gmres_v(const orc_lhs* lhs, orc_subls* ls, int dof, const double* Val, double* R);
void orc_gmres(const orc_lhs* lhs, orc_subls* ls, int dof, const double* Val,
               const double* R, double* X);
void orc_schur(const orc_lhs* lhs, orc_subls* ls, int dof, const double* D, const double* G,
               const double* L, double* R);
int  orc_ns_solver(const orc_lhs* lhs, orc_ls* ls, int dof, const double* Val, double* Ri);
void orc_ls_create(orc_ls* ls, int LS_type, double relTol, double absTol, int maxItr,
                   int dimKry, const double relTolIn[2], const double absTolIn[2],
                   const int maxItrIn[2]);
/* fsils_solve (solve.cpp:50-193), prec: 701 = PREC_FSILS (diagonal), 0 = none. */
int orc_fsils_solve(orc_lhs* lhs, orc_ls* ls, int dof, double* Ri, double* Val, int prec,
                    int nIncL, const int* incL, const double* res);

/* ---- time integration (pic.cpp) ---- */
void orc_pici(int nNo, double am, double af, const double* Ao, const double* An,
              const double* Yo, const double* Yn, double* Ag, double* Yg);
void orc_picc_update(int nNo, double gam, double dt, const double* R, double* An, double* Yn);
void orc_picp(int nNo, double gam, const double* Ao, const double* Yo, double* An, double* Yn);
void orc_set_bc_dir(int nNo, const int* gN, int tDof, int s, int lDof, unsigned eDir, const double* gxn,
                    double dirA, double dirY, double* A, double* Y);

#ifdef __cplusplus
}
#endif
#endif
