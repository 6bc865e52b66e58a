/*
 * oracle.h -- FP64 CPU restatement of cpolyg's implicit DG path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the correctness oracle: the test
 * suite, the smoke check and the CPU-baseline leg of bench.py link it as the
 * checker.  The product path (cpolyg_amd/libcpolyg_hip.so) never calls it.
 *
 * Every function follows the reference algorithm it names (file:line relative to
 * the cpolyg source tree).  Array layouts:
 *   vectors  : u[i*bl + c*bs + k]      (MeshFn cube, src/MeshFn.cpp:14-21)
 *   blocks   : bl x bl column-major    (arma::mat), entry (r,c) at [c*bl + r]
 *   BSR      : rowptr[n+1], colind[nnzb], blocks[nnzb*bl*bl]
 */
#ifndef CPOLYG_ORACLE_H
#define CPOLYG_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int np;
    const double *v;        /* nv x 2 */
    const int *pptr;        /* np+1 */
    const int *pv;          /* polygon vertices, CCW */
    const int *p2p;         /* neighbour per edge, <0 boundary id */
    int nbc;
    const int *bc;          /* nbc x 7: periodic,p1,p2,a1,b1,a2,b2 ; row k = id -1-k */
    const int *tptr;        /* np+1 */
    const double *txy;      /* ntri x 6 */
    const double *bb;       /* np x 4: x0,y0,w,h */
} o_mesh;

enum { O_ADVECTION = 0, O_EULER = 1 };
enum { O_BC_ZERO = 0, O_BC_CONSTANT = 1, O_BC_ISENTROPIC_VORTEX = 2, O_BC_REF_VORTEX = 3 };

typedef struct {
    int kind;               /* O_ADVECTION or O_EULER */
    double beta[6];         /* advection: bx = b0+b1 x+b2 y, by = b3+b4 x+b5 y */
    double gamma;           /* Euler */
    int bc_kind;            /* Dirichlet functor */
    double bc_par[8];
    int periodic_trace;     /* 0: reference trace (src/BoundaryConditions.cpp:18), 1: exact */
    int jacobian;           /* Euler Jacobian form.  0: reference (row-cell state in off-diagonal
                               blocks, Equation.cpp:84,109; energy-row entry of A_x, Euler.cpp:157);
                               1: exact derivative of the residual */
} o_eq;

int o_nc(const o_eq *eq);
/* host threads for timing (1, the default, keeps every loop in its serial order) */
void o_set_threads(int n);

/* geometry / basis helpers exposed for tests */
void o_legendre(int p, double x, double *P, double *dP);
void o_basis(int p, double xi, double eta, double *phi, double *dxi, double *deta);
void o_outward_normal(const o_mesh *m, int i, int a, int b, double *nx, double *ny);
int o_rule_sizes(int deg2, int *ntri, int *nlin);
int o_rule_tables(int deg2, double *tri /* ntri*3 */, double *lin /* nlin*2 */);
double o_polygon_integral_monomial(const o_mesh *m, int i, int deg2, int px, int py);

/* assembly (src/Equation.cpp, src/MassMatrix.cpp, src/MeshFn.cpp) */
int o_mass(const o_mesh *m, int deg, int nc, double *blocks);
int o_assemble(const o_mesh *m, const o_eq *eq, int deg, const double *u, double t, double *b);
int o_jacobian_pattern(const o_mesh *m, int *rowptr, int *colind);
int o_jacobian(const o_mesh *m, const o_eq *eq, int deg, const double *u, double t,
               const int *rowptr, double *blocks);
void o_l2norm(const o_mesh *m, int deg, int nc, const double *u, double *out);
/* MeshFn::interp (src/MeshFn.cpp:56-145): the (deg+1)^2 fitting points of every cell (x, y
 * pairs, cell-major), and the per-cell least-squares fit to values at them (vals: np x
 * (deg+1)^2 x nc) */
void o_gauss_nodes(int m, double *x);
void o_interp_points(const o_mesh *m, int deg, double *xy);
int o_interp_values(const o_mesh *m, int deg, int nc, const double *vals, double *u);
/* MeshFn::L2Error (src/MeshFn.cpp:277-296): the degree-deg2 rule points of every cell (x, y,
 * weight x sub-triangle area; o_quad_count of them), and the per-component error of u against
 * values f at those points for deg2 = 2 deg */
int o_quad_count(const o_mesh *m, int deg2);
void o_quad_points(const o_mesh *m, int deg2, double *xyw);
void o_l2_error_values(const o_mesh *m, int deg, int nc, const double *u, const double *f, double *out);
void o_eval_bc(const o_eq *eq, double x, double y, double t, double *out);

/* block sparse algebra (src/BlockMatrix.cpp, src/Preconditioners.cpp) */
void o_bsr_matvec(int n, int bl, const int *rowptr, const int *colind, const double *A,
                  const double *x, double *y);
int o_inv(int bl, const double *A, double *Ainv);
int o_dilu_factor(int n, int bl, const int *rowptr, const int *colind, const double *A, double *F);
void o_dilu_apply(int n, int bl, const int *rowptr, const int *colind, const double *F, double *x);
int o_bjacobi_factor(int n, int bl, const int *rowptr, const int *colind, const double *A,
                     double *LU, int *piv);
void o_bjacobi_apply(int n, int bl, const double *LU, const int *piv, double *x);

enum { O_PC_NONE = 0, O_PC_BJACOBI = 1, O_PC_DILU = 2 };
enum { O_GMRES_LEFT_REF = 0, O_GMRES_FLEXIBLE = 1 };

/* pc_data: DILU factors (A layout) or block-Jacobi LU (n*bl*bl) + piv (n*bl) */
int o_gmres(int n, int bl, const int *rowptr, const int *colind, const double *A,
            int pc_kind, const double *pc_data, const int *piv,
            const double *b, double *x, int m, double *tol, int *maxit, int mode);
int o_jacobi_solve(int n, int bl, const int *rowptr, const int *colind, const double *A,
                   const double *LU, const int *piv, const double *b, double *x,
                   double *tol, int *maxit);

/* time stepping (src/TimeIntegration.cpp) */
typedef struct {
    int gmres_mode;         /* O_GMRES_* */
    int restart;            /* m */
    double rtol;            /* GMRES tolerance */
    int maxit;              /* GMRES max iterations */
    int pc_kind;            /* O_PC_* */
    double newton_tol;      /* absolute, on max_c L2Norm (kNewtonTolerance) */
    int newton_maxit;       /* kNewtonMaxIterations */
    int dirk_tableau;       /* 3-stage DIRK: O_DIRK3_REFERENCE or O_DIRK3_L_STABLE */
} o_solver;

typedef struct {
    int newton_iters;       /* total over stages */
    int gmres_iters;        /* total */
    int linear_solves;
    double last_resid;
} o_stats;

/* 3-stage DIRK tableaux: the reference's (src/TimeIntegration.cpp:71-77) and Alexander's
 * L-stable SDIRK3 (not in the reference) */
enum { O_DIRK3_REFERENCE = 0, O_DIRK3_L_STABLE = 1 };
void o_dirk_tableau(int nstages, double *A, double *b, double *c);
void o_dirk_tableau_v(int nstages, int variant, double *A, double *b, double *c);
void o_irk_tableau(int nstages, double *A, double *b, double *c);
int o_dirk_step(const o_mesh *m, const o_eq *eq, int deg, int nstages, const double *u,
                double dt, double t, const o_solver *s, double *unp1, o_stats *st);
int o_irk_step(const o_mesh *m, const o_eq *eq, int deg, int nstages, const double *u,
               double dt, double t, const o_solver *s, double *unp1, o_stats *st);
int o_forward_euler_step(const o_mesh *m, const o_eq *eq, int deg, const double *u, double dt, double t,
                         double *unp1);
int o_rk2_step(const o_mesh *m, const o_eq *eq, int deg, const double *u, double dt, double t, double *unp1);
int o_rk4_step(const o_mesh *m, const o_eq *eq, int deg, const double *u, double dt, double t,
               double *unp1);

#ifdef __cplusplus
}
#endif
#endif
