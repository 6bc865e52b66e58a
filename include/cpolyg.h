/*
 * cpolyg.h -- C ABI of the MI355X implicit DG library (libcpolyg_hip.so).
 *
 * Plain pointers and sizes only.  Vectors passed to compute calls are DEVICE
 * pointers (HBM); host pointers appear only in create/upload/download calls.
 * Every call returns an int status (CPG_OK = 0); cpg_last_error() gives the
 * message of the last failure on the calling thread.  No exception crosses the
 * ABI.  `stream` is a hipStream_t (NULL = the default stream); calls that take
 * a stream are asynchronous unless stated otherwise.
 *
 * Layouts (same as the reference):
 *   vector   u[i*bl + c*bs + k]     (MeshFn cube, src/MeshFn.cpp:14-21)
 *   block    bl x bl column-major   (arma::mat)
 *   BSR      rowptr[n+1], colind[nnzb], blocks[nnzb][bl][bl]
 *            (BlockMatrix, src/BlockMatrix.h:10-77)
 */
#ifndef CPOLYG_H
#define CPOLYG_H

#ifdef __cplusplus
extern "C" {
#endif

enum cpg_status {
    CPG_OK = 0,
    CPG_ERR_ARG = 1,           /* bad argument / shape */
    CPG_ERR_HIP = 2,           /* HIP runtime error */
    CPG_ERR_OOM = 3,           /* device allocation failed */
    CPG_ERR_SINGULAR = 4,      /* zero pivot in a block LU / inverse */
    CPG_ERR_NONFINITE = 5,     /* NaN/Inf in a residual or Krylov quantity */
    CPG_ERR_NOT_CONVERGED = 6, /* Newton or GMRES hit its iteration limit */
    CPG_ERR_UNSUPPORTED = 7    /* size/degree combination not compiled in */
};

typedef struct cpg_bsr_s *cpg_bsr;
typedef struct cpg_prec_s *cpg_prec;
typedef struct cpg_krylov_s *cpg_krylov;
typedef struct cpg_mesh_s *cpg_mesh;
typedef struct cpg_equation_s *cpg_equation;
typedef struct cpg_stepper_s *cpg_stepper;

const char *cpg_status_string(int status);
const char *cpg_last_error(void);
int cpg_version(void);

/* ---- device memory helpers (for callers without their own allocator) ---- */
int cpg_device_alloc(void **ptr, unsigned long long bytes);
int cpg_device_free(void *ptr);
int cpg_memcpy_h2d(void *dst, const void *src, unsigned long long bytes, void *stream);
int cpg_memcpy_d2h(void *dst, const void *src, unsigned long long bytes, void *stream);   /* synchronous */
int cpg_memcpy_d2d(void *dst, const void *src, unsigned long long bytes, void *stream);
/* out = a x + b y over n doubles (y null: out = a x; out may alias x or y): MeshFn's arithmetic
 * (operator+, -, *, /, +=, -=; src/MeshFn.h:59-84, src/MeshFn.cpp:209-236) */
int cpg_vec_axpby(unsigned long long n, double a, const double *x, double b, const double *y, double *out,
                  void *stream);
int cpg_stream_sync(void *stream);

/* ---- block sparse row matrix ------------------------------------------------
 * Replaces BlockMatrix (src/BlockMatrix.h:10-77): storage, matvec
 * (src/BlockMatrix.cpp:28-49), operator*= (:297-308).  rowptr/colind are host
 * arrays copied at create; the block array lives in HBM. */
int cpg_bsr_create(int n_rows, int bl, int nnzb, const int *rowptr, const int *colind, cpg_bsr *out);
int cpg_bsr_destroy(cpg_bsr A);
int cpg_bsr_info(cpg_bsr A, int *n_rows, int *bl, int *nnzb);
int cpg_bsr_upload(cpg_bsr A, const double *host_blocks, void *stream);
int cpg_bsr_download(cpg_bsr A, double *host_blocks, void *stream);   /* synchronous */
int cpg_bsr_device_blocks(cpg_bsr A, double **dev_blocks);
int cpg_bsr_scale(cpg_bsr A, double s, void *stream);
/* y = A x; x and y are device vectors of n_rows*bl doubles and must not alias (null when n_rows = 0) */
int cpg_bsr_matvec(cpg_bsr A, const double *x, double *y, void *stream);

/* ---- launch shapes of the block-row streams ----------------------------------
 * The matvec, DILU sweeps and block-Jacobi apply run one workgroup of 1, 2 or 4
 * wavefronts per block row.  No reference counterpart (the reference calls BLAS per
 * block).  cpg_stream_autotune times A's row stream in `precision` for each shape (reps
 * launches each; ms[3] receives the average ms for 1, 2, 4 waves; synchronous) and keeps
 * the fastest for (precision, bl) in a process-wide table that later launches read. */
int cpg_stream_set_wpr(int precision, int bl, int wpr);
int cpg_stream_get_wpr(int precision, int bl, int *wpr);
int cpg_stream_autotune(cpg_bsr A, int precision, int reps, int *best_wpr, float *ms, void *stream);

/* ---- preconditioners --------------------------------------------------------
 * Replace Preconditioner::solve (src/Preconditioners.h:10-15), BlockJacobi
 * (src/Preconditioners.cpp:8-47) and BlockILU0 = block DILU
 * (src/Preconditioners.cpp:49-158).  The DILU sweeps are level-scheduled from
 * the matrix pattern (a colour-ordered matrix gives one level per colour).
 * Factors are stored in `precision` (CPG_FP64, CPG_FP32, CPG_BF16 with one
 * power-of-two scale per block), packed in the order the sweeps read them; FP32
 * factors are applied with FP64 accumulation, BF16 factors with FP32 per-lane
 * partial sums and an FP64 total.  x is FP64 in and out. */
enum { CPG_PC_NONE = 0, CPG_PC_BJACOBI = 1, CPG_PC_DILU = 2, CPG_PC_INNER_GMRES = 3 };
enum { CPG_FP64 = 0, CPG_FP32 = 1, CPG_BF16 = 2, CPG_BF16_D32 = 3 };
/* CPG_BF16_D32 (block DILU of a colour-ordered matrix only): the L and DU blocks in BF16 with the
 * per-block scale, the diagonal inverses D_i^-1 in FP32.  The inverses carry the conditioning of
 * the diagonal blocks: with them in BF16 the p = 4 Euler Newton matrices at CFL 20 leave flexible
 * GMRES stalled, in FP32 it converges as with FP32 factors.  No reference counterpart. */
/* CPG_PC_INNER_GMRES: a variable preconditioner for CPG_GMRES_FLEXIBLE -- a fixed number
 * of FP32 GMRES iterations (cpg_prec_set_inner, default 20) on an FP32 copy of A, right-
 * preconditioned by the DILU factors in `precision` (FP32 or BF16); FP64 reductions,
 * device-side Givens, no host synchronisation.  No reference counterpart: it replaces the
 * preconditioner apply of BlockMatrix::gmres (src/BlockMatrix.cpp:122) in the mixed-
 * precision mode.  A must stay alive between cpg_prec_factor and cpg_prec_apply. */
int cpg_prec_create(cpg_bsr A, int kind, int precision, cpg_prec *out);
int cpg_prec_set_inner(cpg_prec P, int iterations);
int cpg_prec_destroy(cpg_prec P);
int cpg_prec_factor(cpg_prec P, cpg_bsr A, void *stream);
int cpg_prec_apply(cpg_prec P, double *x, void *stream);   /* x <- M^-1 x, in place */
/* out <- M^-1 in (distinct vectors; CPG_PC_NONE copies): the form flexible GMRES applies,
 * z_i = M^-1 v_i.  The reference's Preconditioner::solve (src/Preconditioners.h:10-15) works
 * in place only. */
int cpg_prec_apply_to(cpg_prec P, const double *in, double *out, void *stream);
/* Block DILU of the Radau IIA stage-coupled Newton matrix in cell-interleaved order (3 stages
 * x bl = 40: 120-wide blocks), with the coupled blocks formed inside the factor: block (s, t)
 * of slot q is coef[3 s + t] * J_s[q], plus E's mass block on the diagonal slots when s == t.
 * Replaces BlockMatrix::blockBlockMatrix (src/BlockMatrix.cpp:228-258) followed by the
 * BlockILU0 constructor (src/Preconditioners.cpp:49-123) for IRK::advance
 * (src/TimeIntegration.cpp:233-245), on the interleaved ordering (a different DILU from the
 * stage-major one) and without storing the coupled matrix.  P: kind CPG_PC_DILU created over a
 * BSR of block size 120 with the stage Jacobians' pattern; J0..J2: that pattern at block size
 * 40; coef: 9 host doubles (-dt A for the Newton matrix M - dt (A x J)); synchronous.  FP32 and
 * BF16 factors form their Schur and DU products on FP32 MFMA (FP32 factor blocks within about
 * 3e-6 of the FP64 factor's); the environment variable CPOLYG_COUPLED_FP64_PRODUCTS=1, read at
 * each call, selects FP64 products (within 1e-6).  The inverses are FP64 in every precision. */
int cpg_prec_factor_coupled(cpg_prec P, cpg_equation E, cpg_bsr J0, cpg_bsr J1, cpg_bsr J2, const double *coef,
                            void *stream);
/* factors converted to FP64 in A's slot layout: diag = D^-1, lower = L, upper = DU
 * (DILU) or n inverse diagonal blocks (block Jacobi); synchronous */
int cpg_prec_download(cpg_prec P, double *host_blocks);
int cpg_prec_levels(cpg_prec P, int *n_levels);
/* Block DILU apply on FP64 vectors in one launch: the rows of every level, forward then backward,
 * each waiting on device-side completion counters of the rows it reads, where the level launches
 * order them by launch boundaries (the same arithmetic per row: bitwise the same result).  Used by
 * default where it applies (CPG_PC_DILU, block sizes 3, 6, 10, 12, 15, 24, 40 and 60, at most 32
 * blocks per row; CPG_BF16_D32 factors take one launch per sweep, the backward one forming
 * h = D^-1 y from the FP32 inverses and x = h - sum DU x in each row's workgroup).  enable: 1 on,
 * 0 one launch per level, < 0 unchanged.
 * spin_limit >= 0 bounds each wait (polls; default 2^22; 0 gives up at once, for tests of the
 * failure path), < 0 unchanged.  A wait that runs out makes that apply's result invalid: the next
 * apply or the GMRES solve that used it returns CPG_ERR_HIP, and the level launches are used from
 * then on.  No reference counterpart (BlockILU0::solve, src/Preconditioners.cpp:125-158, is a
 * sequential loop). */
int cpg_prec_set_flow(cpg_prec P, int enable, int spin_limit);
int cpg_prec_flow_state(cpg_prec P, int *in_use);

/* ---- GMRES ------------------------------------------------------------------
 * Replaces BlockMatrix::gmres (src/BlockMatrix.cpp:79-163) and Jacobian::solve
 * (src/Jacobian.cpp:68-97).  mode CPG_GMRES_LEFT_REF is the reference algorithm
 * (left preconditioning, classical Gram-Schmidt, Givens, tolerance relative to
 * ||P^-1 b||); CPG_GMRES_FLEXIBLE is right-preconditioned flexible GMRES with
 * the tolerance relative to ||b||.  tol and maxit are in/out exactly as in the
 * reference; x holds the initial guess on entry.  Synchronous on return. */
enum { CPG_GMRES_LEFT_REF = 0, CPG_GMRES_FLEXIBLE = 1 };
int cpg_krylov_create(int n_rows, int bl, int restart, int mode, cpg_krylov *out);
int cpg_krylov_destroy(cpg_krylov K);
int cpg_gmres(cpg_krylov K, cpg_bsr A, cpg_prec P, const double *b, double *x, double *tol, int *maxit,
              void *stream);
/* Replaces BlockMatrix::jacobi (src/BlockMatrix.cpp:51-77): x <- (P^-1 b + x) - P^-1 A x (omega = 1)
 * until ||A x - b||_2 < tol (absolute), at most *maxit iterations.  On return *tol is ||A x - b||_2
 * of the last check and *maxit the iterations taken, as in the reference (0 and 0 when *maxit is
 * 0).  P is any preconditioner of A (the reference passes its BlockJacobi) or null for P = I; x
 * holds the initial guess.  Uses K's vectors (any mode); synchronous on return. */
int cpg_jacobi(cpg_krylov K, cpg_bsr A, cpg_prec P, const double *b, double *x, double *tol, int *maxit,
               void *stream);

/* ---- mesh and equations -----------------------------------------------------
 * Mesh arrays are PolyMesh's public fields (src/PolyMesh.h:22-28) flattened:
 * v[nv*2], polygon CSR pptr/pv (vertex lists), p2p (neighbour per edge, <0 is
 * boundary id), bc[nbc*7] rows (periodic,p1,p2,a1,b1,a2,b2) for ids -1..-nbc,
 * sub-triangles tptr/txy[ntri*6], bounding boxes bb[np*4]. */
int cpg_mesh_create(int np, int nv, const double *v, const int *pptr, const int *pv, const int *p2p, int nbc,
                    const int *bc, const int *tptr, const double *txy, const double *bb, cpg_mesh *out);
int cpg_mesh_destroy(cpg_mesh M);

enum { CPG_ADVECTION = 0, CPG_EULER = 1 };
enum { CPG_BC_ZERO = 0, CPG_BC_CONSTANT = 1, CPG_BC_ISENTROPIC_VORTEX = 2, CPG_BC_REF_VORTEX = 3 };
typedef struct {
    int kind;               /* CPG_ADVECTION (nc=1) or CPG_EULER (nc=4) */
    double beta[6];         /* advection: bx = b0+b1 x+b2 y, by = b3+b4 x+b5 y */
    double gamma;           /* Euler ratio of specific heats */
    int bc_kind;            /* Dirichlet functor for non-periodic boundary ids */
    double bc_par[8];
    int periodic_trace;     /* 0: reference trace (src/BoundaryConditions.cpp:18), 1: exact */
    int jacobian;           /* Euler Jacobian form.  0: the reference's (off-diagonal blocks with the
                               row cell's state, src/Equation.cpp:84,109; its A_x energy-row entry
                               H-(gamma-1)|u|^2, eqns/Euler.cpp:157); 1: exact derivative */
} cpg_eq_params;

/* Equation (src/Equation.h:8-67) bound to a mesh and polynomial degree 1..4.
 * Creation precomputes face geometry and the mass matrix (src/MassMatrix.cpp:30-94). */
int cpg_equation_create(cpg_mesh M, const cpg_eq_params *params, int deg, cpg_equation *out);
int cpg_equation_destroy(cpg_equation E);
int cpg_equation_info(cpg_equation E, int *n_cells, int *bl, int *nnzb);
/* Jacobian pattern (Jacobian::Jacobian, src/Jacobian.cpp:7-47) into host arrays */
int cpg_jacobian_pattern(cpg_equation E, int *rowptr, int *colind);
int cpg_jacobian_create(cpg_equation E, cpg_bsr *out);
/* b = assemble(u, t)  (Equation::assemble, src/Equation.cpp:152-191) */
int cpg_assemble_residual(cpg_equation E, const double *u, double t, double *b, void *stream);
/* J = scale * jacobian(u, t) (+ mass matrix on the diagonal blocks if add_mass)
 * (Equation::jacobian src/Equation.cpp:193-279, operator*= and J += M) */
int cpg_assemble_jacobian(cpg_equation E, const double *u, double t, double scale, int add_mass, cpg_bsr J,
                          void *stream);
/* y = M v + scale * J(u, t) v: the product of v with the Newton matrix that
 * cpg_assemble_jacobian(E, u, t, scale, add_mass = 1, ...) writes (Jacobian::operator*= and
 * J += M, src/Jacobian.cpp:99-110, then BlockMatrix::matvec, src/BlockMatrix.cpp:28-49, as in
 * DIRK::advance, src/TimeIntegration.cpp:127-140), formed at the quadrature points without
 * assembling the matrix.  The per-point flux Jacobians are those of the assembled form selected
 * by the equation's `jacobian` option, so the two agree to rounding.  y must not alias u or v. */
int cpg_newton_apply(cpg_equation E, const double *u, double t, double scale, const double *v, double *y,
                     void *stream);
/* cpg_newton_apply in two parts, for many products with one Newton matrix (no reference
 * counterpart; the reference assembles the matrix once and calls BlockMatrix::matvec per GMRES
 * iteration): cpg_newton_prepare evaluates what depends on (u, t) alone at every quadrature point
 * and keeps it in the equation; cpg_newton_apply_prepared then forms y = M v + scale J(u, t) v for
 * the (u, t) of the last prepare on E (v and y distinct).  Calls on one equation are ordered by the
 * stream, as the equation's other calls are. */
int cpg_newton_prepare(cpg_equation E, const double *u, double t, void *stream);
int cpg_newton_apply_prepared(cpg_equation E, double scale, const double *v, double *y, void *stream);
int cpg_mass_matvec(cpg_equation E, const double *x, double *y, void *stream);
int cpg_mass_download(cpg_equation E, double *host_blocks);   /* n blocks bl x bl, synchronous */
/* MeshFn::L2Norm (src/MeshFn.cpp:299-312) of u, nc values to host, synchronous */
int cpg_l2norm(cpg_equation E, const double *u, double *out, void *stream);
/* cpg_l2norm that allocates nothing.  No counterpart in the reference: it is for callers that
 * cannot allocate device memory (tight serving loops, code that forbids hipMalloc).  Its partial
 * sums live in a fixed 32.8 KB device array of the library (1024 partial sums for each of at most
 * 4 components, then the sums), held on every GPU the library runs on until the process ends;
 * any cell count is accepted, an equation with more than 4 components is rejected.  Calls that
 * overlap on one GPU share the array and race.  Same result as cpg_l2norm up to the order of
 * the sums; call with E's device current; synchronous. */
int cpg_l2norm_noalloc(cpg_equation E, const double *u, double *out, void *stream);

/* ---- analytic fields: projection and L2 error on the device ----------------
 * MeshFn::interp (src/MeshFn.cpp:56-145): per cell the least-squares fit of the basis to the
 * field at the (p+1)^2 tensor Gauss-Legendre points of the cell's bounding box.
 * MeshFn::L2Error (src/MeshFn.cpp:277-296): per component sqrt(sum_i int_i (u_h - f)^2) with the
 * cell's degree-2p polygon rule.  The reference takes a host functor; here the field is one of
 * a fixed set evaluated in the kernel, or values the caller computed at the points.
 *   CPG_FIELD_GAUSSIAN (nc 1): par = cx, cy, a, vx, vy, Lx, Ly:
 *     exp(-a |x - (c + v t)|^2), nearest periodic image when Lx, Ly > 0
 *   CPG_FIELD_VORTEX (nc 4): par = beta, x0, y0, uinf, vinf, gamma, Lx, Ly: the isentropic vortex
 *     (conserved variables) centred at (x0, y0) + (uinf, vinf) t, nearest image when Lx, Ly > 0
 *   CPG_FIELD_SINE_DENSITY (nc 4): par = ux, uy, p, gamma, amp: density
 *     1 + amp sin 2pi(x - ux t) sin 2pi(y - uy t), velocity (ux, uy), pressure p
 * u, vals, xy: device arrays; the field's nc must be E's. */
enum { CPG_FIELD_GAUSSIAN = 0, CPG_FIELD_VORTEX = 1, CPG_FIELD_SINE_DENSITY = 2 };
int cpg_interp_field(cpg_equation E, int field, const double *par, double t, double *u, void *stream);
/* the n (p+1)^2 fitting points, (x, y) pairs in cell-major order */
int cpg_interp_points(cpg_equation E, double *xy, void *stream);
/* the fit of caller values at those points (vals: n (p+1)^2 nc, component fastest) */
int cpg_interp_values(cpg_equation E, const double *vals, double *u, void *stream);
/* err: nc host doubles; synchronous */
int cpg_l2_error_field(cpg_equation E, const double *u, int field, const double *par, double t, double *err,
                       void *stream);

/* ---- time stepping ----------------------------------------------------------
 * TimeStepper::advance (src/TimeIntegration.h:9-18) for DIRK (1 = backward
 * Euler, 2 = SDIRK2, 3, 5 stages; src/TimeIntegration.cpp:31-151) with the
 * Newton/GMRES loop on the device. */
enum { CPG_STEP_DIRK = 0, CPG_STEP_IRK = 1, CPG_STEP_RK4 = 2, CPG_STEP_IRK_W = 3, CPG_STEP_RK2 = 4,
       CPG_STEP_FORWARD_EULER = 5 };
/* Explicit steps, k_q = dt M^-1 b(u + c_q k_{q-1}, t + c_q dt), u_{n+1} = u + sum_q w_q k_q:
 * CPG_STEP_RK4 (RK4::advance, src/TimeIntegration.cpp:14-22), CPG_STEP_RK2 (RK2::advance, :24-29:
 * c = (0, 1/2), w = (1/2, 1/2), as the reference computes it), CPG_STEP_FORWARD_EULER
 * (ForwardEuler::advance, :10-12).  M^-1 is applied per component (the reference's
 * MassMatrix::solve indexes past its output for nc > 1, src/MassMatrix.cpp:108-109); nstages
 * and the solver parameters are not used. */
/* CPG_STEP_IRK_W: Radau IIA in the form of IRK::newAdvance (src/TimeIntegration.cpp:283-395):
 * unknowns w_s = sum_t A[s][t] k_t (stage states u + w_s), residuals M (A^-1 w)_s - dt b(u + w_s),
 * Newton matrix blocks A^-1[s][j] M (+ -dt J_s on the diagonal), preconditioned by the
 * stage-block-diagonal factors of A^-1[s][s] M - dt J_s (pc_kind DILU or block Jacobi, any
 * precision); the coupled matrix is applied matrix-free, so gmres_mode must be flexible. */
enum { CPG_DIRK3_REFERENCE = 0, CPG_DIRK3_L_STABLE = 1 };
typedef struct {
    int gmres_mode;         /* CPG_GMRES_* */
    int restart;
    double rtol;
    int maxit;
    int pc_kind;            /* CPG_PC_* */
    int precision;          /* factor storage CPG_FP64/FP32/BF16 */
    double newton_tol;      /* absolute, on max_c L2Norm of the Newton residual */
    int newton_maxit;
    int inner_iters;        /* CPG_PC_INNER_GMRES: FP32 iterations per outer iteration (0: 20) */
    int newton_guess;       /* 0: each stage starts from k_s = 0 (reference, src/TimeIntegration.cpp:176);
                               1: from the previous stage's k */
    int newton_backtrack;   /* 0: none (reference); 1: halve a Newton update (up to 10 times)
                               while the residual at the new iterate is non-finite */
    int newton_operator;    /* operator of the GMRES solves (DIRK/backward Euler): 0 the BSR matvec of the
                               assembled Newton matrix (reference, src/TimeIntegration.cpp:127-140); 1 the
                               same matrix applied at the quadrature points (cpg_newton_apply), which agrees
                               with it to rounding; the assembled matrix still feeds the preconditioner */
    int pc_refresh;         /* DIRK/backward Euler: 0 assemble the Newton matrix and factor the preconditioner
                               at every Newton iteration (reference, src/TimeIntegration.cpp:127-136); with
                               newton_operator = 1 only: 1 at the first Newton iteration of each stage, 2 at
                               the first Newton iteration of the step, 3 keep the factor across steps and
                               renew it when a step needs more than 1.5 x (+2) the GMRES iterations of the
                               first step after the last renewal, or when dt changes (the Newton iteration
                               itself is unchanged: only the preconditioner lags).  Radau IIA: 0 every
                               Newton iteration, 1 and 2 once per step, 3 as for DIRK */
    int dirk_tableau;       /* DIRK with 3 stages: CPG_DIRK3_REFERENCE the reference's tableau
                               (src/TimeIntegration.cpp:71-77: second order, not A-stable, R(-inf) = -1.78);
                               CPG_DIRK3_L_STABLE Alexander's L-stable SDIRK3 (third order, R(-inf) = 0;
                               no reference counterpart) */
    int pc_level_launches;  /* block DILU apply on FP64 vectors: 0 one launch per apply where it applies
                               (cpg_prec_set_flow), 1 one launch per colour level */
} cpg_solver_params;

typedef struct {
    int newton_iters;
    int gmres_iters;
    int linear_solves;
    int status;
    double last_resid;
    double ms_residual, ms_jacobian, ms_factor, ms_gmres, ms_total;
    int backtracks;         /* Newton update halvings (newton_backtrack) */
} cpg_step_stats;

int cpg_stepper_create(cpg_equation E, int method, int nstages, const cpg_solver_params *params, cpg_stepper *out);
int cpg_stepper_destroy(cpg_stepper S);
/* unp1 = advance(u, dt, t); u and unp1 are device vectors; synchronous */
int cpg_stepper_advance(cpg_stepper S, const double *u, double dt, double t, double *unp1, cpg_step_stats *stats,
                        void *stream);
/* per-kernel timing of BSR matvec and preconditioner apply launches inside GMRES
 * (HIP events on the launch stream); enable before advance, read after */
int cpg_stepper_enable_kernel_timing(cpg_stepper S, int enable);
int cpg_stepper_kernel_timing(cpg_stepper S, int *n_matvec, double *ms_matvec, int *n_apply, double *ms_apply);

/* ---- batched implicit steps --------------------------------------------------
 * Many independent implicit solves advanced together.  No reference counterpart: the
 * reference advances one mesh at a time (DIRK::advance, src/TimeIntegration.cpp:93-151,
 * BackwardEuler::advance :31-66).  Instances are grouped into classes; class c is one
 * equation over a union mesh whose connected components are its ninst[c] instances
 * (instance q of class c owns cells cell_ptr[c][q] .. cell_ptr[c][q+1]-1).  Instances are
 * numbered class by class.  Each stage's Newton systems of all instances are solved by one
 * batched flexible GMRES(params->restart): matvec, DILU factor and sweeps cover whole
 * classes, inner products are segmented per instance, Hessenberg/Givens/least squares run
 * on the device per instance (an instance that converges stops while the others go on),
 * and the solve runs in one of three modes:
 *   CPG_BATCH_EAGER       launches per phase (matvec and sweeps per class, segmented CGS);
 *   CPG_BATCH_GRAPH       the same, a restart cycle captured once into hipGraphs (start, blocks
 *                         of 5 Arnoldi steps, finish) and replayed, the host ending a cycle
 *                         early once every instance has converged;
 *   CPG_BATCH_PERSISTENT  one cooperative launch per batched solve, every restart cycle of
 *                         every instance inside it: instance q is solved by its own team of
 *                         workgroups (sized by its bytes per Arnoldi step times its Arnoldi
 *                         count in the previous solve), with team barriers between the DILU
 *                         levels, the matvec and the Gram-Schmidt reductions, so a converged
 *                         instance leaves while the others go on and no instance waits on a
 *                         launch boundary.  DILU (not the BF16 factors with FP32 inverses),
 *                         block-Jacobi or no preconditioner; block sizes up to 64.
 * The row-stream launch shapes of every class's block size are autotuned at creation (modes
 * 0 and 1); the persistent mode tunes its workgroups per CU over its first solves.
 * DIRK with nstages 1 (backward Euler) or 2 (SDIRK2) and the other tableaux of
 * cpg_stepper_create; each instance has its own dt (dt[q]); the stage time passed to the
 * residual is t + c_s dt of a class's first instance, so the equations must have steady
 * coefficients and time-independent Dirichlet data.  u[c], unp1[c]: device vectors of
 * class c.  Synchronous.  inst_gmres_iters (optional, ninst entries): Arnoldi steps per
 * instance in this step. */
typedef struct cpg_batch_s *cpg_batch;
enum { CPG_BATCH_EAGER = 0, CPG_BATCH_GRAPH = 1, CPG_BATCH_PERSISTENT = 2 };
int cpg_batch_create(int nclass, const cpg_equation *eqs, const int *ninst, const int *const *cell_ptr, int nstages,
                     const cpg_solver_params *params, int mode, cpg_batch *out);
int cpg_batch_destroy(cpg_batch B);
int cpg_batch_info(cpg_batch B, int *ninst, int *nchunk, int *graph_captured);
/* Persistent mode: workgroups per CU of the cooperative launch (0: autotune over the first
 * solves, the default), the most workgroups one instance's team may have (0: as many as its
 * rows can use) and the wait limit of a team barrier in sleep rounds (0: the default, a few
 * seconds; a tiny limit forces the abort path: the solve then fails with CPG_ERR_HIP and every
 * workgroup leaves the kernel). */
int cpg_batch_set_teams(cpg_batch B, int wg_per_cu, int max_team, unsigned spin_limit);
/* Persistent mode, after a solve: the grid, the workgroups per CU it used, each instance's team
 * size and the time its team took in the last solve in ms (team_wg, team_ms: ninst entries each,
 * optional). */
int cpg_batch_teams(cpg_batch B, int *grid, int *wg_per_cu, int *team_wg, double *team_ms);
int cpg_batch_advance(cpg_batch B, const double *const *u, const double *dt, double t, double *const *unp1,
                      cpg_step_stats *stats, int *inst_gmres_iters, void *stream);

#ifdef __cplusplus
}
#endif
#endif
