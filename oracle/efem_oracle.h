/*
 * efem_oracle.h -- CPU restatement of EFEM's Total-Lagrangian solve loop.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it is loaded by
 * tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and
 * by nothing else.  The product (efem_amd, libefem_gpu.so) never links,
 * loads or calls it.
 *
 * Plain C99, FP64, single thread.  Every function cites the reference
 * function (file:line, paths relative to the SeanPei/EFEM source tree) whose
 * arithmetic it restates.
 *
 * Conventions
 *   - nodes / elements are 0-based; X is [n_nodes][3] (AoS, FP64);
 *   - element node order is the reference's (VTK / Abaqus order, with the
 *     reference's own parametric coordinates, see or_shape);
 *   - a 3x3 matrix is row-major: M[3*i + j] = M_ij;
 *   - vectors u, r live on the free DOFs: entry 3*id + c, where id is the
 *     node's rank among free nodes in node order (Constraint.cpp:61-72).
 */
#ifndef EFEM_ORACLE_H
#define EFEM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* constitutive laws (Material.h:9-16) */
enum { OR_NEOHOOKEAN = 0, OR_MOONEY_RIVLIN = 1, OR_YEOH = 2 };

/* material parameter vector p[4]:
 *   Neo-Hookean    : p = {C10, 0,   0,   D1}
 *   Mooney-Rivlin  : p = {C10, C01, 0,   D1}
 *   Yeoh           : p = {C10, C20, C30, D1}                              */

typedef struct {
    int n_nodes;
    const double* X;            /* [n_nodes][3] reference positions */
    int n_elems;
    int nen;                    /* 4, 8, 10 or 20 */
    const int* conn;            /* [n_elems][nen] */
    int reduced;                /* 0: the reference's (full) rule, 1: reduced rule */
    int law;
    const double* params;       /* [4] shared, or [n_elems][4] if params_per_elem */
    int params_per_elem;
    const unsigned char* fixed; /* [n_nodes] 1 = fixed (NULL: none) */
    int n_faces;
    int fnn;                    /* face node count: 3, 6, 4 or 8 */
    const int* faces;           /* [n_faces][fnn] */
    double pressure;            /* follower pressure value (<= 0: none) */
    const double* pdir;         /* non-NULL: directed traction vector [3] instead */
    int n_fnodes;
    const int* fnodes;          /* nodes carrying a constant nodal force */
    const double* fvec;         /* [3] */
    const double* volumes;      /* optional per-element volumes (NULL: computed) */
    int precision;              /* unused here (keeps the layout of efem_problem) */
    int gravity;                /* 0 none, 1..6: +x -x +y -y +z -z (9.8) */
    int consistent_inertia;     /* 1: consistent mass matrix, 0: lumped */
    const double* density;      /* [1] or [n_elems] as params; NULL: 0 */
} or_problem;

typedef struct or_model or_model;

/* ---- element library ---------------------------------------------------- */
int or_shape(int nen, const double xi[3], double* N, double* dN);
int or_n_gauss(int nen, int reduced);
int or_gauss(int nen, int reduced, int q, double xi[3], double* w);
double or_element_volume(int nen, const double* Xe);

/* ---- materials ---------------------------------------------------------- */
void or_piola(int law, const double p[4], const double F[9], double P[9]);
double or_elasticity_coeff(int law, const double p[4]);
/* strain energy density whose derivative dW/dF is or_piola (an addition:
 * the reference never evaluates an energy) */
double or_energy_density(int law, const double p[4], const double F[9]);

/* ---- model -------------------------------------------------------------- */
or_model* or_model_create(const or_problem* pb, char* err, int errlen);
void or_model_destroy(or_model* m);
int or_ndof(const or_model* m);
int or_free_id(const or_model* m, int node);
double or_volume_of(const or_model* m, int e);
void or_residual(or_model* m, const double* u, double* r);
void or_internal_force(or_model* m, const double* u, double* r);
void or_external_force(or_model* m, const double* u, double* r);
int or_face_parent(const or_model* m, int f);
/* total potential of a conservative problem: strain energy
 * sum_e sum_q V_e w_q W(F_q) minus the work of the constant loads (nodal
 * forces, gravity), so that its gradient is or_residual; NAN when the model
 * has pressure faces (a follower load has no potential) */
double or_potential(or_model* m, const double* u);
/* Pi(x) - Pi(xp) element by element (keeps its relative accuracy for close
 * states); *scale (may be NULL) = sum of |energy terms| at xp, the size of
 * Pi's rounding.  NAN with pressure faces. */
double or_potential_diff(or_model* m, const double* x, const double* xp, double* scale);

/* approximate Jacobian, full symmetric CSC with sorted rows.
 * scalar = 0: the reference's 3N_f matrix with explicit zero cross terms
 * scalar = 1: the N_f x N_f scalar Laplacian (K = L (x) I3 interleaved) */
long or_K_nnz(or_model* m, int scalar);
void or_K_csc(or_model* m, int scalar, int* Ap, int* Ai, double* Ax);

/* sparse LL^T preconditioner H0 = K^{-1} (LBFGS.h:65-81) */
int or_factor(or_model* m, int scalar);     /* 0 ok, 1 not SPD / empty */
long or_factor_nnz(const or_model* m);
void or_apply_h0(or_model* m, const double* v, double* out);

/* ---- L-BFGS (LBFGS.h:96-174, LBFGSMat.h, LineSearchFEA.h) --------------- */
typedef struct {
    double epsilon;        /* absolute ||r||_inf tolerance */
    double epsilon_rel;    /* relative tolerance (x ||x||) */
    int max_iterations;    /* 0: unlimited */
    int max_linesearch;
    int m;                 /* history size (reference: 6) */
    double lns_fea;        /* projection acceptance factor (reference: 1.0) */
    int use_precond;       /* 1: H0 = K^{-1} (lbfgs_ah), 0: H0 = I/theta (lbfgs) */
    int factor_scalar;     /* factor mode, see or_factor */
    /* additions (no counterpart in the reference): */
    int line_search;       /* 0: the reference's projection rule; 1: energy
                              rule (Armijo decrease of the potential with
                              c1 = 1e-4, plus the projection rule) */
    const double* x_target;/* non-NULL: also stop once ||x - x_target||_2 <=
                              target_tol ||x_target||_2 after an iteration */
    double target_tol;
} or_lbfgs_params;

typedef struct {
    int iterations;
    int r_evals;
    int ls_trials;
    int precond_failed;
    double rinf;
    double xinf;
    double factor_seconds;
    double solve_seconds;
    int target_reached;
    double target_err;
} or_lbfgs_stats;

typedef void (*or_resid_fn)(void* ctx, const double* x, double* r);
typedef void (*or_h0_fn)(void* ctx, const double* v, double* out);
/* Pi(x) - Pi(xp), and in *scale the size of Pi's rounding at xp */
typedef double (*or_energy_diff_fn)(void* ctx, const double* x, const double* xp, double* scale);

/* generic solver: f(x) -> r; h0 NULL means H0 = I/theta.
 * trace (optional) receives per iteration {||r||_inf, step, trials}. */
int or_lbfgs(int n, or_resid_fn f, or_h0_fn h0, void* ctx, double* x,
             const or_lbfgs_params* prm, or_lbfgs_stats* st,
             double* trace, int trace_cap);
/* the same with the difference e(x, xp) of a potential whose gradient is f
 * (needed by the energy line search; may be NULL for the projection rule).
 * Energy rule: a trial is accepted when the projection rule holds and
 * Pi(x) - Pi(xp) <= 1e-4 step (r(xp) . d) or |Pi(x) - Pi(xp)| <= 1e-12 scale;
 * without the decrease the step halves, otherwise it grows by 1.5 or halves
 * as the projection rule does. */
int or_lbfgs_energy(int n, or_resid_fn f, or_energy_diff_fn e, or_h0_fn h0, void* ctx, double* x,
                    const or_lbfgs_params* prm, or_lbfgs_stats* st, double* trace, int trace_cap);
int or_lbfgs_solve_model(or_model* m, double* x, const or_lbfgs_params* prm,
                         or_lbfgs_stats* st, double* trace, int trace_cap);

/* two-loop recursion of LBFGSMat::apply_Hv on explicit history, for tests.
 * S, Y column-major [n][mcap]; h0 NULL -> divide by theta. */
void or_two_loop(int n, int mcap, int ncorr, int ptr, const double* S,
                 const double* Y, const double* ys, double theta,
                 or_h0_fn h0, void* ctx, const double* v, double a,
                 double* res);

#ifdef __cplusplus
}
#endif
#endif
