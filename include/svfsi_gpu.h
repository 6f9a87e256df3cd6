/*
 * svfsi_gpu — MI355X-native fluid Newton step for svFSIplus (C ABI).
 *
 * The two reference entry points this library replaces on the device:
 *
 *   void fluid::construct_fluid(ComMod& com_mod, const mshType& lM,
 *                               const Array<double>& Ag, const Array<double>& Yg)
 *        Code/Source/solver/fluid.cpp:464  (declared in fluid.h)
 *   void fsi_linear_solver::fsils_solve(FSILS_lhsType& lhs, FSILS_lsType& ls,
 *        const int dof, Array<double>& Ri, Array<double>& Val,
 *        const consts::PreconditionerType prec, const Vector<int>& incL,
 *        const Vector<double>& res)
 *        Code/Source/liner_solver/fsils_api.hpp:60-61, solve.cpp:50
 *
 * All arrays keep the reference layouts (column-major Array<T>,
 * Code/Source/solver/Array.h:374-388):
 *   x(3,nNo)  IEN(4,nEl)  Ag/Yg/R(4,nNo)  Val(16,nnz) with block j at
 *   Val[16j .. 16j+15], row-major inside the 4x4 block (spar_mul.cpp:232-235).
 *   rowPtr(nNo+1) is the exclusive-end CSR of com_mod.rowPtr; colPtr(nnz)
 *   holds sorted column node ids per row (lhsa.cpp:366-379).
 *
 * Pointers named d_* are device (HBM) pointers; the others are host pointers.
 * Every function returns SVG_OK (0) or an SVG_ERR_* code; svg_last_error()
 * gives the message (the reference throws std::runtime_error on the same
 * conditions). A stream argument may be NULL for the default stream.
 * Calls on one svg_lhs / svg_msh are not re-entrant (as in the reference).
 */
#ifndef SVFSI_GPU_H
#define SVFSI_GPU_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SVG_OK 0
#define SVG_ERR_ARG 1          /* bad argument / size                                  */
#define SVG_ERR_HIP 2          /* HIP runtime error (allocation, launch)               */
#define SVG_ERR_JACOBIAN 3     /* fluid.cpp:612-621  zero element Jacobian             */
#define SVG_ERR_SINGULAR 4     /* ns_solver.cpp:386-391 "Singular matrix detected"     */
#define SVG_ERR_ILLCOND 5      /* ns_solver.cpp:464-472 ill-conditioned LHS            */
#define SVG_ERR_LSTYPE 6       /* solve.cpp:179 "LS_type not defined"                  */
#define SVG_ERR_NEU_RES 7      /* solve.cpp:97 "res is required for Neu surfaces"      */
#define SVG_ERR_UNSUPPORTED 8  /* a reference feature this library does not provide   */

/* Linear solver types (fils_struct.hpp:71-77) and preconditioners (consts.h:426-440). */
#define SVG_LS_BICGS 795
#define SVG_LS_NS 796
#define SVG_LS_GMRES 797
#define SVG_LS_CG 798
#define SVG_PREC_NONE 700
#define SVG_PREC_FSILS 701

/* FSILS face groups (fils_struct.hpp:60-64). */
#define SVG_BC_DIR 0
#define SVG_BC_NEU 1

/* Node flags written by the pipe generator. */
#define SVG_NODE_INLET 1
#define SVG_NODE_WALL 2
#define SVG_NODE_OUTLET 4

const char* svg_last_error(void);

/* ------------------------------------------------------------------------ */
/* Synthetic workload generator (setup, not timed).                          */
/* Straight pipe of radius `radius`, length `length` along z: an n x n hex   */
/* cross-section (n even) mapped onto the disc by the elliptical map         */
/* (s,t) -> (s sqrt(1-t^2/2), t sqrt(1-s^2/2)), nz hexes axially, 6 tets per */
/* hex (Kuhn split mirrored per quadrant), interior nodes jittered by        */
/* U(-jitter, jitter) * h per coordinate from a splitmix64 stream seeded by  */
/* `seed`. Optional stenosis: radius scaled by 1 - depth*(1-cos(2pi(z-z0)/len))/2 */
/* on z0 <= z <= z0+len.                                                      */
typedef struct {
  int n, nz;
  double radius, length, jitter;
  unsigned long long seed;
  double sten_depth, sten_z0, sten_len;
} svg_pipe_spec;

int svg_pipe_counts(const svg_pipe_spec* sp, int* nNo, int* nEl);
int svg_pipe_generate(const svg_pipe_spec* sp, double* x, int* IEN, int* flags);

/* Bifurcating pipe (the batched workload): a parent pipe (n x n cross-section
 * on the disc of radius `radius`, nzp layers over length_parent along z) that
 * splits into two daughters of nzd layers over length_daughter. Daughter 0
 * takes the parent's end half x <= 0 and turns by angle0 (radians) towards -x,
 * daughter 1 the half x >= 0 and turns by angle1 towards +x, over the first
 * max(3, nzd/3) layers; each keeps its D-shaped half section, scaled to its
 * Murray radius (r0^3 = split r^3, r1^3 = (1 - split) r^3). n is a multiple of
 * 4; interior nodes are jittered as in the pipe. Node flags: inlet (z = 0),
 * wall (outer surfaces and the carina line), outlet (both daughter ends). */
typedef struct {
  int n, nzp, nzd;
  double radius, length_parent, length_daughter;
  double angle0, angle1, split, jitter;
  unsigned long long seed;
} svg_bif_spec;

int svg_bif_counts(const svg_bif_spec* sp, int* nNo, int* nEl);
int svg_bif_generate(const svg_bif_spec* sp, double* x, int* IEN, int* flags);

/* Block-CSR pattern of a mesh (what lhsa builds, lhsa.cpp:153-380). */
int svg_csr_pattern_count(int nNo, int nEl, int eNoN, const int* IEN, int* nnz);
int svg_csr_pattern(int nNo, int nEl, int eNoN, const int* IEN, int* rowPtr, int* colPtr);
/* Greedy element colouring (no two elements of one colour share a node). */
int svg_color_elements(int nNo, int nEl, int eNoN, const int* IEN, int* color, int* nColors);

/* ------------------------------------------------------------------------ */
/* Device-resident FSILS lhs.                                                 */
/* Replaces fsils_lhs_create (liner_solver/lhs.cpp:57-120, one rank: map is   */
/* the identity, rowPtr pairs are [rowPtr(a), rowPtr(a+1)-1], diagPtr found   */
/* by search) and fsils_bc_create (liner_solver/bc.cpp:45-130).               */
typedef struct svg_lhs svg_lhs;

int svg_lhs_create(int nNo, int nnz, const int* rowPtr, const int* colPtr, int nFaces, svg_lhs** lhs);
int svg_lhs_bc_create(svg_lhs* lhs, int faIn, int nNo, int dof, int bGrp, const int* gNodes, const double* val);
int svg_lhs_free(svg_lhs* lhs);
/* The lhs of a TET4 mesh whose connectivity d_IEN (4, nEl) is already in
 * device memory: the block-CSR pattern lhsa builds (lhsa.cpp:153-380: the
 * node pairs of every element and every node's diagonal, rows ascending,
 * columns sorted and unique) computed on the device (one radix sort of the
 * pairs and a unique pass), then fsils_lhs_create's structures as
 * svg_lhs_create makes them. Synchronises `stream`. */
int svg_lhs_create_dev(int nNo, int nEl, const int* d_IEN, int nFaces, svg_lhs** lhs, void* stream);
int svg_lhs_nnz(const svg_lhs* lhs, int* nnz);
int svg_lhs_pattern(const svg_lhs* lhs, int* rowPtr, int* colPtr); /* host copies: nNo+1 and nnz ints */

/* ------------------------------------------------------------------------ */
/* Element assembly: construct_fluid (fluid.cpp:464-708) for TET4 P1-P1 VMS  */
/* with constant viscosity, one fluid domain, fixed mesh (mvMsh false).       */
typedef struct svg_msh svg_msh;

/* Uploads x and IEN, colours the elements, builds the element -> block slot
 * map over `lhs`'s pattern (the binary search of do_assem, lhsa.cpp:123-134)
 * and, for svg_construct_fluid, its inverse: per block slot the element
 * entries (e, a, b) that add to it, in element order. */
int svg_msh_create(const svg_lhs* lhs, int nEl, const int* IEN, int nNo, const double* x, svg_msh** msh);
/* The same from device-resident IEN (4, nEl) and x (3, nNo), every step on
 * the device: Jones-Plassmann colouring (the sequential greedy colouring in
 * a fixed pseudo-random element order; no two elements of one colour share a
 * node; at most 256 colours), the stable colour order and the slot map.
 * Synchronises `stream`. */
int svg_msh_create_dev(const svg_lhs* lhs, int nEl, const int* d_IEN, const double* d_x, svg_msh** msh,
                       void* stream);
int svg_msh_free(svg_msh* msh);
int svg_msh_num_colors(const svg_msh* msh, int* nColors);
int svg_msh_colors(const svg_msh* msh, int* colorOff);  /* nColors+1 offsets into the colour order */
int svg_msh_sorted_ien(const svg_msh* msh, int* ien);   /* (4, nEl) in the colour order */

typedef struct {
  double rho;      /* fluid_density                                   */
  double mu;       /* viscosity (viscType_Const, fluid.cpp:2158-2162)  */
  double f[3];     /* body force f_x, f_y, f_z                         */
  double K_darcy;  /* inverse_darcy_permeability                       */
  double af, am, gam; /* generalized-alpha (initialize.cpp:430-463)    */
  double dt;
  double qmTET4;   /* quadrature modifier, default (5+3 sqrt5)/20      */
} svg_fluid_params;

/* R(4,nNo) += element residuals, Val(16,nnz) += element tangents (the
 * caller zeroes them first, as ls_alloc does). d_Bf may be NULL (zero).
 * Each block and residual row receives its element contributions in element
 * order with every operation rounded separately, as the reference's element
 * loop adds them: the result equals the reference's bit for bit. */
int svg_construct_fluid(svg_msh* msh, const svg_fluid_params* p, const double* d_Ag, const double* d_Yg,
                        const double* d_Bf, double* d_R, double* d_Val, void* stream);

/* Neumann faces of the fluid mesh (faceType, ComMod.h: IEN, gE) as TRI3
 * elements: faceIEN(3,nFe) node ids, gE(nFe) the parent element (index into
 * the IEN given to svg_msh_create), which orients the outward normal as
 * gnnb does (nn.cpp:552-760). */
typedef struct svg_face svg_face;
int svg_face_create(const svg_msh* msh, int nFe, const int* faceIEN, const int* gE, svg_face** face);
int svg_face_free(svg_face* face);
/* b_assem_neu_bc (eq_assem.cpp:58-172) for the fluid equation: b_fluid
 * (fluid.cpp:46-133) on every face element — R -= the Neumann value d_hg (per
 * node; NULL: zero traction) and the backflow stabilisation with coefficient
 * beta (Backflow_stabilization_coefficient, default 0.2); Val gets the
 * backflow tangent. qmTRI3: face quadrature modifier (default 2/3).
 * Asynchronous on the stream. */
int svg_b_assem_neu_bc(svg_face* face, const svg_fluid_params* p, double beta, double qmTRI3, const double* d_hg,
                       const double* d_Yg, double* d_R, double* d_Val, void* stream);

/* ------------------------------------------------------------------------ */
/* FSILS solve (solve.cpp:50-193). Mirrors FSILS_subLsType / FSILS_lsType    */
/* (fils_struct.hpp:216-283).                                                */
typedef struct {
  int suc, mItr, sD, itr, cM, cN, cD;
  double absTol, relTol, iNorm, fNorm, dB, callD;
} svg_subls;

typedef struct {
  int LS_type;
  int Resm, Resc;
  svg_subls GM, CG, RI;
} svg_ls;

/* fsils_ls_create defaults (ls.cpp:40-120) without the optional arguments. */
int svg_ls_create(svg_ls* ls, int LS_type);

/* Ri(dof,nNo): residual in, Newton increment out. Val(dof*dof,nnz) is scaled
 * in place by the diagonal preconditioner, as in the reference. dof = 4.
 * incL(nIncL)/res(nIncL) have the reference meaning (res may be NULL when no
 * Neumann face is registered). A Neumann face with res != 0 is coupled
 * (solve.cpp:96-111): with LS type GMRES every block product adds
 * res valM (valM^T x) (add_bc_mul.cpp:53-121, valM = val W from
 * precond_diag, precond.cpp:246-256); the NS solver adds the same term to
 * K U (ns_solver.cpp:334) and to its inner GMRES's products, and applies op
 * PRE, Y += -res/(1 + res nS) valM (valM^T X) with nS = |valM|^2, to each
 * inner Krylov vector (gmres.cpp:137-199) and to G P in the Schur CG
 * (cgrad.cpp:101-108). Synchronous: returns when Ri holds the result
 * and ls holds the iteration record. */
int svg_fsils_solve(svg_lhs* lhs, svg_ls* ls, int dof, double* d_Ri, double* d_Val, int prec, int nIncL,
                    const int* incL, const double* res, void* stream);

/* ------------------------------------------------------------------------ */
/* Learned preconditioner (no counterpart in the reference; it stands beside */
/* precond_diag, precond.cpp:122-257, and is used only through the options   */
/* of svg_fsils_solve_ex). A 3-layer message-passing network over the mesh   */
/* nodes, hidden width 64 (16 channels x 4 dofs), BF16 weights, activations   */
/* and matrix copy with FP32 accumulation on MFMA:                            */
/*   x = v sqrt(nNo); h0 = W0 [x, D_i (diagonal block, 16), 1, 0...]          */
/*   3x: m_i = sum_j A_ij (x) h_j;  h_i += relu(W_l [h_i; m_i] + b_l)          */
/*   z = v + (Wd h) / sqrt(nNo)          (Wd = 0: the diagonal preconditioner) */
/* on the diagonally scaled system. Parameters are one FP32 vector:          */
/*   W0 [64][32] | 3 x (W_l [64][128], b_l [64]) | Wd [4][64].                */
typedef struct svg_lpc svg_lpc;

int svg_lpc_num_params(int* count);
int svg_lpc_create(svg_lpc** net); /* all parameters zero */
int svg_lpc_free(svg_lpc* net);
int svg_lpc_set_params(svg_lpc* net, const float* params); /* host FP32; resets the Adam state */
int svg_lpc_get_params(const svg_lpc* net, float* params);
/* BF16 copies of the blocks and diagonal blocks of the scaled d_Val (after
 * svg_precond_diag) for the network on this lhs. */
int svg_lpc_prepare(svg_lhs* lhs, const double* d_Val, void* stream);
/* d_Z = M(d_V) for nvec stacked vectors (nvec, nNo, 4) (blocking). */
int svg_lpc_apply(svg_lpc* net, const svg_lhs* lhs, int nvec, const double* d_V, double* d_Z, void* stream);
/* Training loss L = (1/nvec) sum_b |R_b - A M(R_b)|^2 over nvec stacked
 * residuals d_R on the scaled FP64 matrix d_Val, and its gradient (host FP32,
 * svg_lpc_num_params entries; may be NULL). */
int svg_lpc_loss_grad(svg_lpc* net, svg_lhs* lhs, const double* d_Val, int nvec, const double* d_R, double* loss,
                      float* grad, void* stream);
/* One Adam step (beta 0.9/0.999) with learning rate lr on that loss. */
int svg_lpc_train_step(svg_lpc* net, svg_lhs* lhs, const double* d_Val, int nvec, const double* d_R, float lr,
                       double* loss, void* stream);
/* The same step on the loss (1/nvec) sum_b |T_b - A M(V_b)|^2 with network
 * input d_V and target d_T (d_T NULL: T = V, svg_lpc_train_step). The network
 * as the smoother of the two-level preconditioner is trained with V = the
 * block-Jacobi output and T = the coarse-corrected residual. */
int svg_lpc_train_step_ex(svg_lpc* net, svg_lhs* lhs, const double* d_Val, int nvec, const double* d_V,
                          const double* d_T, float lr, double* loss, void* stream);

/* Solve options beyond the reference (no counterpart in fsils_solve):
 *   precision SVG_PRECISION_MIXED (LS type GMRES): Arnoldi cycles use an FP32
 *   copy of the scaled Val and FP32 copies of the SpMV input; the
 *   Gram-Schmidt dots, Givens rotations, least squares and the solution stay
 *   FP64, and every restart recomputes the FP64 true residual R - A X, which
 *   both refines the solution and decides convergence (|R - A X| < eps).
 *   method SVG_METHOD_LEARNED (LS type GMRES): flexible GMRES with
 *   z_j = M(u_j) of the network `lpc` on the diagonally scaled system (prec
 *   SVG_PREC_FSILS scales it; SVG_PREC_NONE means the caller already did),
 *   restarts on the FP64 true residual; after fallback_window (0: 20)
 *   consecutive non-decreasing iterations (none lowering the smallest
 *   Givens residual estimate so far by 0.01 % or more) the solve continues
 *   from the current X with the plain diagonal-preconditioned GMRES, and
 *   fallback_itr reports the iteration count at that point (-1: none). */
#define SVG_PRECISION_FP64 0
#define SVG_PRECISION_MIXED 1
#define SVG_METHOD_PLAIN 0
#define SVG_METHOD_LEARNED 1
typedef struct {
  int precision;
  int method;
  int fallback_window;
  int fallback_itr; /* out */
  int net_period;   /* learned: the network preconditions Arnoldi steps i with i % net_period == 0, the
                       others use the diagonal scaling alone (0 or 1: every step) */
  int coarse;       /* > 0: the two-level preconditioner on every Arnoldi step (flexible GMRES, LS type
                       GMRES), with node aggregates of about `coarse` nodes (0: none); method PLAIN
                       smooths with block Jacobi, method LEARNED with the network `lpc` on its steps
                       (lpc may then be NULL: block Jacobi) */
  int coarse_keep;  /* with coarse > 0: 1 keeps the coarse inverse of this lhs's last two-level set-up
                       (only A P and the block-Jacobi inverses are rebuilt from Val), e.g. for the Newton
                       iterations after the first of a time step; 0 rebuilds it */
  int reserved;
  svg_lpc* lpc;
} svg_solve_opts;

int svg_fsils_solve_ex(svg_lhs* lhs, svg_ls* ls, int dof, double* d_Ri, double* d_Val, int prec, int nIncL,
                       const int* incL, const double* res, svg_solve_opts* opts, void* stream);

/* ------------------------------------------------------------------------ */
/* Two-level preconditioner (no counterpart in the reference; it stands      */
/* beside precond_diag, precond.cpp:122-257, and is used by the flexible     */
/* GMRES through svg_solve_opts.coarse). On the diagonally scaled system A:  */
/*   M(v) = P zc + B (v - A P zc),  zc = Ac^-1 P^T v,  Ac = P^T A P           */
/* P: piecewise-constant prolongation of the pressure (dof 3) over node      */
/* aggregates grown breadth-first over the lhs's block pattern (about aggSize */
/* nodes each; aggregates under aggSize/4 nodes join the neighbour they share */
/* most edges with); Ac: dense (at most 8192 aggregates), inverted in FP64 by */
/* blocked Gauss-Jordan elimination and applied in FP32; B: the inverses of  */
/* the 4x4 diagonal blocks (FP32).                                            */
/* svg_coarse_prepare builds the aggregates (once per lhs and aggSize) and    */
/* this system's operators from the scaled d_Val (synchronises stream).       */
/* svg_coarse_apply on nvec stacked vectors (nvec, nNo, 4): mode              */
/* SVG_COARSE_APPLY gives Z = M(V); SVG_COARSE_RESIDUAL gives the coarse-     */
/* corrected residual Z = V - A P zc (what a smoother other than B receives); */
/* SVG_COARSE_SMOOTH gives Z = B (V - A P zc).                                */
/* svg_coarse_info: the number of aggregates and (agg != NULL) each node's.   */
#define SVG_COARSE_APPLY 0
#define SVG_COARSE_RESIDUAL 1
#define SVG_COARSE_SMOOTH 2 /* Z = B (V - A P zc): the block-Jacobi smoothing without P zc */
int svg_coarse_prepare(svg_lhs* lhs, int aggSize, const double* d_Val, void* stream);
int svg_coarse_apply(const svg_lhs* lhs, int mode, int nvec, const double* d_V, double* d_Z, void* stream);
int svg_coarse_info(const svg_lhs* lhs, int* nAgg, int* agg);

/* ------------------------------------------------------------------------ */
/* Batched GMRES (no counterpart in the reference): many independent systems  */
/* assembled into one block-diagonal lhs (system q owns nodes                 */
/* [nodeOff[q], nodeOff[q+1])) are solved together by gmres_v's algorithm    */
/* (gmres.cpp:450-629) with per-system Krylov spaces, dot products, Givens    */
/* rotations and convergence; each system's arithmetic is that of a solve of  */
/* the system alone. The restart cycle of all systems is captured into two   */
/* hipGraphs, one cycle and four cycles per replay (systems that stop inside  */
/* a replay are skipped by every kernel, and mItr is enforced per system on   */
/* the device); the four-cycle graph is replayed while more than an eighth  */
/* of the systems iterate, the one-cycle graph after that, until every system */
/* has converged or ran mItr cycles. The host reads the systems' states once  */
/* per replay (svg_batch_replays counts the replays).                         */
/* d_R, d_Val: the union's (already diagonally scaled) right-hand side and    */
/* matrix; d_R receives the per-system solutions; out[nsys] the records.      */
/* precision SVG_PRECISION_MIXED: FP32 SpMV, FP64 true-residual restarts.     */
typedef struct svg_batch svg_batch;
int svg_batch_create(svg_lhs* lhs, int nsys, const int* nodeOff, int sD, svg_batch** batch);
int svg_batch_free(svg_batch* batch);
int svg_batch_gmres(svg_batch* batch, const svg_subls* proto, double* d_R, const double* d_Val, int precision,
                    svg_subls* out, void* stream);
/* fsils_solve (solve.cpp:50-193) for the batch: the diagonal preconditioner
 * over the union (every face included), svg_batch_gmres, R = W R. */
int svg_batch_fsils_solve(svg_batch* batch, const svg_subls* proto, double* d_R, double* d_Val, int precision,
                          svg_subls* out, void* stream);
/* One Newton iteration of every mesh of the batch, device-resident from the
 * initiator to the corrector (the loop body of main.cpp:371-617 for the fluid
 * equation): pici (pic.cpp:486-574), R = Val = 0 (ls_alloc), construct_fluid,
 * b_assem_neu_bc on face (null: none; Neumann value zero, backflow beta),
 * svg_batch_fsils_solve and the picc update (pic.cpp:155-167). msh and face
 * are over the batch's lhs. The work before the restart loop (assembly
 * included) is one captured hipGraph, the restart cycle a second and the
 * solution/scaling/corrector a third; all are captured at the first call and
 * replayed while the pointers and parameters stay the same. construct_fluid's
 * zero-Jacobian error is reported after the iteration (SVG_ERR_JACOBIAN). */
int svg_batch_newton_iteration(svg_batch* batch, svg_msh* msh, svg_face* face, const svg_fluid_params* p,
                               double beta, double qmTRI3, const double* d_Ao, double* d_An, const double* d_Yo,
                               double* d_Yn, double* d_Ag, double* d_Yg, double* d_R, double* d_Val,
                               const svg_subls* proto, const svg_solve_opts* opts, svg_subls* out, void* stream);
/* The batched solves with the options of svg_fsils_solve_ex (opts null: FP64
 * plain): precision, and method SVG_METHOD_LEARNED with opts->lpc — flexible
 * GMRES with the learned preconditioner for every system (each normalised by
 * its own size), with the fallback to the diagonal preconditioner decided per
 * system; each system's arithmetic is that of svg_fsils_solve_ex on the system
 * alone. opts->fallback_itr is not written: svg_batch_fallback_itr gives the
 * per-system iteration count at the fallback of the last solve (-1: none). */
int svg_batch_gmres_ex(svg_batch* batch, const svg_subls* proto, double* d_R, const double* d_Val,
                       const svg_solve_opts* opts, svg_subls* out, void* stream);
int svg_batch_fsils_solve_ex(svg_batch* batch, const svg_subls* proto, double* d_R, double* d_Val,
                             const svg_solve_opts* opts, svg_subls* out, void* stream);
int svg_batch_fallback_itr(const svg_batch* batch, int* itr);
/* Restart-cycle graph replays (one host synchronisation each) since the batch was created. */
int svg_batch_replays(const svg_batch* batch, long long* replays);

/* ------------------------------------------------------------------------ */
/* Kernels of the solve, exposed for unit parity tests.                      */
/* fsils_spar_mul_vv (spar_mul.cpp:191-258), dof in {1..4}.                  */
int svg_spar_mul_vv(const svg_lhs* lhs, int dof, const double* d_K, const double* d_U, double* d_KU, void* stream);
/* Autotuner of the 4x4 block SpMV (no counterpart in the reference): all
 * launch variants (unroll depth, non-temporal matrix streaming, per-XCD row
 * bands) compute bit-identical products; svg_spmv_tune times each on d_K over
 * lhs's pattern and keeps the fastest for this lhs (run automatically at the
 * first GMRES solve). */
int svg_spmv_num_variants(int* count);
int svg_spmv_time_variant(const svg_lhs* lhs, int variant, const double* d_K, const double* d_U, double* d_KU,
                          int launches, double* avg_us, void* stream);
int svg_spmv_tune(svg_lhs* lhs, const double* d_K, int* best_variant, void* stream);
/* The FP32 block SpMV of the mixed-precision solve (no counterpart in the
 * reference): KU = K32 U32 with FP32 blocks and vector, FP32 fused
 * multiply-adds in the block order of spar_mul.cpp:228-237, KU in FP64.
 * Asynchronous on the stream. */
int svg_spmv32(const svg_lhs* lhs, const float* d_K32, const float* d_U32, double* d_KU, void* stream);
/* Autotuner of the FP32 SpMV (no counterpart in the reference): times its
 * launch variants (unroll depth, per-XCD row bands; identical results) on
 * d_K32 and keeps the fastest for this lhs (run automatically at the first
 * mixed-precision solve). d_U32 is zeroed and d_KU overwritten (scratch). */
int svg_spmv32_tune(svg_lhs* lhs, const float* d_K32, float* d_U32, double* d_KU, int* best_variant, void* stream);
/* The launch variants the autotuners picked for lhs (-1: not tuned yet). */
int svg_spmv_variants(const svg_lhs* lhs, int* fp64_variant, int* fp32_variant);
/* precond_diag (precond.cpp:122-257) with every registered face included. */
int svg_precond_diag(svg_lhs* lhs, int dof, double* d_Val, double* d_R, double* d_W, void* stream);
/* fsils_dot_v / fsi_ls_normv (dot.cpp:69-117, norm.cpp:67-112); blocking. */
int svg_dot_v(int dof, int nNo, const double* d_U, const double* d_V, double* out, void* stream);
int svg_norm_v(int dof, int nNo, const double* d_U, double* out, void* stream);
/* svg_dot_v without any allocation. No counterpart in the reference: it is
 * for callers that cannot allocate device memory (tight solver or serving
 * loops, harnesses that forbid hipMalloc). Same kernels, same order and a
 * bit-identical result as svg_dot_v, but the reduction's partial sums live in
 * a 2 MiB static device array that each GPU holds until the process ends, so
 * dof*nNo is limited to 2^28 entries (larger sizes return SVG_ERR_ARG). Every
 * call on one GPU shares that array: calls that overlap on one GPU (other
 * host threads, or other streams) race and must be serialised by the caller.
 * Blocking, like svg_dot_v. */
int svg_dot_v_noalloc(int dof, int nNo, const double* d_U, const double* d_V, double* out, void* stream);

/* ------------------------------------------------------------------------ */
/* Generalized-alpha pieces of a Newton iteration (pic.cpp).                 */
/* pici (pic.cpp:486-574): Ag = (1-am) Ao + am An, Yg = (1-af) Yo + af Yn.   */
int svg_pici(int nNo, double am, double af, const double* d_Ao, const double* d_An, const double* d_Yo,
             const double* d_Yn, double* d_Ag, double* d_Yg, void* stream);
/* picc update (pic.cpp:155-167): An -= R, Yn -= gam dt R. */
int svg_picc_update(int nNo, double gam, double dt, const double* d_R, double* d_An, double* d_Yn, void* stream);
/* picp (pic.cpp:591-715) for a fluid equation (no displacement unknowns):
 * An = Ao (gam-1)/gam, Yn = Yo over (4, nNo). */
int svg_picp(int nNo, double gam, const double* d_Ao, const double* d_Yo, double* d_An, double* d_Yn, void* stream);
/* set_bc_dir (set_bc.cpp:794-1063) for one strong Dirichlet face with a
 * steady or Fourier-interpolated value (set_bc_dir_l's non-generic branch,
 * no impD): for face node a (global node d_gN[a]) and k < lDof,
 *   A(row, gN[a]) = dirA * gxn(k, a),  Y(row, gN[a]) = dirY * gxn(k, a),
 * where gxn(lDof, nNo) holds gx(a) nV(k, a) when lDof = nsd and gx(a)
 * otherwise, A and Y are (tDof, tnNo), and row = s + k (eDir = 0) or s plus
 * the k-th direction set in the mask eDir (bc.eDrn). The host evaluates dirA
 * and dirY (bc.g, or ifft of bc.gt at the current time). */
int svg_set_bc_dir(int nNo, const int* d_gN, int tDof, int s, int lDof, unsigned eDir, const double* d_gxn,
                   double dirA, double dirY, double* d_A, double* d_Y, void* stream);
/* d[0..n) = v (a fill kernel, safe under graph capture). */
int svg_fill(double* d, size_t n, double v, void* stream);

/* ------------------------------------------------------------------------ */
/* Kernel timing with HIP events on the launching stream.                    */
/* When enabled, every launch of the named kernels is bracketed by events;   */
/* svg_prof_get returns the launch count and summed duration (ms).           */
int svg_prof_enable(int on);
/* Restricts the event timing to the kernels named in the comma-separated
 * list (NULL or "": every named kernel). Events cost time between launches,
 * so a timed run brackets only the kernels it reports on. */
int svg_prof_only(const char* names);
/* Brackets only one launch in `every` of each timed kernel (1: all); the
 * count and summed duration then cover the sampled launches. */
int svg_prof_sample(int every);
int svg_prof_reset(void);
int svg_prof_get(const char* name, int* count, double* total_ms);
/* Summed algorithmic bytes of the named kernel's timed launches (where the
 * library states them: the block SpMV kernels). */
int svg_prof_get_bytes(const char* name, double* bytes);

#ifdef __cplusplus
}
#endif
#endif
