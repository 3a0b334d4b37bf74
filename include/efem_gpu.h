/*
 * efem_gpu.h -- C ABI of the MI355X (gfx950) Total-Lagrangian solve loop.
 *
 * Drop-in boundary for EFEM's solver concept.  The reference is in-process
 * C++ (no FFI); each entry point below replaces the call named in its
 * comment (paths relative to the SeanPei/EFEM source tree):
 *
 *   efem_create     Constraint::Constraint + IntegralElement::initial +
 *                   ElementConstraint ctor + ExternalPressureFace::initial
 *                   (src/fe/Constraint.cpp:22-91,
 *                    src/fe/element/IntegralElement.cpp:3-19,
 *                    src/fe/ElementConstraint.cpp:69-83,
 *                    src/fe/ExternalConstraint.cpp:249-827)
 *   efem_residual   SoftObjective::residual (src/simulation/Integrator.cpp:49-55)
 *                   -> Constraint::getResidual (src/fe/Constraint.cpp:115-126)
 *   efem_factor     SoftObjective::jacobi + LBFGSSolver::updateApproxJacobi
 *                   (Integrator.cpp:57-70, src/simulation/nlsolver/LBFGS.h:65-81)
 *   efem_apply_h0   SimplicialLLT::solve as called at LBFGS.h:130 and
 *                   LBFGSMat.h:120 (external/Eigen/src/SparseCholesky/
 *                   SimplicialCholesky.h:155-181)
 *   efem_solve      LBFGSSolver<SoftObjective>::solve (LBFGS.h:96-174) as
 *                   called by Integrator::integral (Integrator.cpp:22-42)
 *   efem_commit     Constraint::updateDx (src/fe/Constraint.cpp:103-113)
 *   efem_write_vtk  Model::saveVTK (src/model/Model.cpp:50-117)
 *
 * Conventions: plain pointers and sizes; host pointers unless the name ends
 * in _d (device pointers, ordered on the context's HIP stream).  Vectors on
 * the free DOFs are interleaved: entry 3*id + c, id = rank of the node among
 * free (non-fixed) nodes in node order (Constraint.cpp:61-72).  Every call
 * returns EFEM_OK (0) or a negative code; efem_last_error() gives the text.
 * No C++ exception crosses this boundary.  One context per host thread.
 */
#ifndef EFEM_GPU_H
#define EFEM_GPU_H

#ifdef __cplusplus
extern "C" {
#endif

#define EFEM_OK 0
#define EFEM_ERR_ARG (-1)       /* invalid argument / unsupported input */
#define EFEM_ERR_HIP (-2)       /* HIP runtime error (no device, OOM, ...) */
#define EFEM_ERR_NONFINITE (-3) /* NaN/Inf in the result (Integrator.cpp:28-34) */
#define EFEM_ERR_STATE (-4)     /* call out of order */

/* constitutive law (src/material/Material.h:9-16) */
#define EFEM_NEOHOOKEAN 0
#define EFEM_MOONEY_RIVLIN 1
#define EFEM_YEOH 2

/* arithmetic of the element kernels and the preconditioner (bit mask) */
#define EFEM_FP64 0
#define EFEM_F32_ELEMENTS 1 /* FP32 element kernels; FP64 J-1, volumetric
                               coefficient and residual assembly */
#define EFEM_F32_PRECOND 2  /* FP32 factor blocks and triangular solves */
#define EFEM_MIXED 3        /* both; L-BFGS vectors, scalars and norms stay FP64 */

/* line-search rule */
#define EFEM_LS_PROJECTION 0 /* the reference's |r.d| rule (LineSearchFEA.h:37-68) */
/* energy rule (no counterpart in the reference; conservative loads only --
 * nodal forces, gravity -- since a follower pressure has no potential):
 * a trial is accepted when the projection rule holds and the total
 * potential Pi = strain energy - constant loads' work decreased enough,
 * Pi(x) - Pi(xp) <= 1e-4 step (r(xp) . d), or changed by less than
 * 1e-12 of its size (rounding); without the decrease the step halves,
 * otherwise it grows by 1.5 or halves as in the projection rule.  Pi(x) -
 * Pi(xp) is summed element by element in FP64 with compensation. */
#define EFEM_LS_ENERGY 1

/*
 * Problem description, copied at efem_create (the caller keeps ownership).
 * Material parameters p[4]: NH {C10,0,0,D1}; MR {C10,C01,0,D1};
 * Yeoh {C10,C20,C30,D1} (initFromJson of src/material/NeoHookean.cpp,
 * MooneyRivlin.cpp, Yeoh.cpp).
 */
typedef struct efem_problem {
    int n_nodes;
    const double* X;            /* [n_nodes][3] reference positions */
    int n_elems;
    int nen;                    /* 4 (C3D4), 8 (C3D8), 10 (C3D10), 20 (C3D20) */
    const int* conn;            /* [n_elems][nen], 0-based, reference node order */
    int reduced;                /* 0 full integration (the reference), 1 reduced */
    int law;                    /* EFEM_NEOHOOKEAN / _MOONEY_RIVLIN / _YEOH */
    const double* params;       /* [4] or [n_elems][4] */
    int params_per_elem;        /* 1: params is per element */
    const unsigned char* fixed; /* [n_nodes] 1 = fixed, or NULL */
    int n_faces;
    int fnn;                    /* pressure-face node count: 3, 6, 4 or 8 */
    const int* faces;           /* [n_faces][fnn] */
    double pressure;            /* follower pressure (> 0), ExternalConstraint.cpp:184-247 */
    const double* pdir;         /* non-NULL: directed traction [3] (:937-940) */
    int n_fnodes;
    const int* fnodes;          /* nodes with a constant force (:132-175) */
    const double* fvec;         /* [3] */
    const double* volumes;      /* optional [n_elems] element volumes; NULL: computed */
    int precision;              /* EFEM_FP64, EFEM_MIXED or a mask of EFEM_F32_* */
    /* gravity (InertiaConstraint.cpp:34-84, analysis.gravity of
     * Config.cpp:45-65): 0 none, 1 +x, 2 -x, 3 +y, 4 -y, 5 +z, 6 -z, of
     * magnitude 9.8; the residual gains -M g over the free DOFs, M the
     * consistent (consistent_inertia = 1) or lumped mass matrix
     * (ElementConstraint.cpp:85-137) built with the elements' densities */
    int gravity;
    int consistent_inertia;
    const double* density;      /* [1] or [n_elems] (indexed as params); NULL: 0 */
} efem_problem;

/* L-BFGS parameters (src/simulation/nlsolver/Param.h:92-103 defaults) */
typedef struct efem_lbfgs_params {
    double epsilon;      /* absolute ||r||_inf tolerance (residual_criterion) */
    double epsilon_rel;  /* relative tolerance, times ||x||_inf */
    int max_iterations;  /* 0: unlimited */
    int max_linesearch;  /* 20 */
    int m;               /* history, 6 in the reference (1..64) */
    double lns_fea;      /* 1.0 */
    int use_precond;     /* 1: H0 = K^-1 (nl_solver "lbfgs_ah"), 0: H0 = I/theta */
    int line_search;     /* EFEM_LS_PROJECTION or EFEM_LS_ENERGY */
    /* optional measurement hook (no counterpart in the reference): when
     * x_target is non-NULL the solve also stops as soon as
     * ||x - x_target||_2 <= target_tol ||x_target||_2 (checked on the device
     * after every iteration); stats->target_reached reports it */
    const double* x_target;
    double target_tol;
    /* EFEM_F32_ELEMENTS only: the element kernels switch from FP32 to FP64 once
     * ||r||_inf <= fp64_switch * ||r(x0)||_inf (0: stay FP32).  The FP32
     * preconditioner is kept to the end. */
    double fp64_switch;
} efem_lbfgs_params;

typedef struct efem_solve_stats {
    int iterations;      /* LBFGSSolver::solve return value */
    int r_evals;         /* SoftObjective::r_evalutes */
    int ls_trials;
    int precond_failed;  /* factorisation failed -> H0 = I/theta (LBFGS.h:76-80) */
    double rinf;         /* final ||r||_inf */
    double xinf;         /* final ||x - u_c||_inf (u_c: efem_commit, else 0) */
    double factor_seconds;
    double solve_seconds;
    long long factor_nnz; /* nnz of the (scalar) Cholesky factor */
    int target_reached;   /* see efem_lbfgs_params.x_target */
    double target_err;    /* last ||x - x_target|| / ||x_target|| */
    int fp64_switch_iteration; /* iteration at which EFEM_MIXED went FP64 (0: never) */
} efem_solve_stats;

typedef struct efem_ctx efem_ctx;

int efem_create(efem_ctx** out, const efem_problem* pb);
int efem_destroy(efem_ctx* ctx);
const char* efem_last_error(const efem_ctx* ctx);
const char* efem_global_error(void); /* error text of a failed efem_create */

int efem_ndof(const efem_ctx* ctx);
int efem_precision(const efem_ctx* ctx);

/* r(u) = f_int(u) - f_ext(u) on the free DOFs (host pointers, synchronous) */
int efem_residual(efem_ctx* ctx, const double* u, double* r);
/* device-pointer variant: u_d and r_d are device buffers of ndof doubles.
 * The caller's writes of u_d must be complete (its stream synchronised);
 * the call returns once r_d is written and u_d is no longer read. */
int efem_residual_d(efem_ctx* ctx, const double* u_d, double* r_d);

/* assemble the approximate Jacobian, order and factor it on the host,
 * upload the factor.  Returns EFEM_OK; *failed = 1 if not SPD (fallback). */
int efem_factor(efem_ctx* ctx, int* failed);
long long efem_factor_nnz(const efem_ctx* ctx);
/* layout of the factor's apply (diagnostic, no counterpart in the
 * reference): {entries of L (nnz, lower triangles of the diagonal blocks),
 * of which those the level sweeps read (the supernodes outside the fused
 * top), fused-top columns nT (0: none),
 * its Z's leading dimension, its depth (-1: none), kernel launches per
 * apply inside the solve loop, elimination-tree levels, supernodes} */
int efem_factor_layout(const efem_ctx* ctx, long long* out8);
/* drop the factor (the next efem_solve with use_precond factors again) */
int efem_factor_release(efem_ctx* ctx);
/* seconds spent by the last factorisation (diagnostic, no counterpart in
 * the reference): {host model: Laplacian assembly + ordering + symbolic +
 * numeric, of which symbolic, of which numeric, device descriptors and maps,
 * block repack, block upload, other device buffers, total} */
int efem_factor_phases(const efem_ctx* ctx, double* out8);
/* out = K^-1 v (host pointers); requires a successful efem_factor */
int efem_apply_h0(efem_ctx* ctx, const double* v, double* out);

/* L-BFGS from u (in/out, host pointer).  Factors first when
 * prm->use_precond and no factor is present. */
int efem_solve(efem_ctx* ctx, double* u, const efem_lbfgs_params* prm,
               efem_solve_stats* stats);

/* Constraint::updateDx (src/fe/Constraint.cpp:103-113): u (3 x free nodes,
 * host) becomes the committed displacement u_c, the reference's node->dx.
 * The reference solves for the increment dx from 0 on top of it; efem_solve
 * iterates on the total displacement u = u_c + dx (pass u = u_c to start
 * where the reference does) and measures the increment for the relative
 * criterion: ||r||_inf <= epsilon_rel ||u - u_c||_inf (LBFGS.h:118, 153), and
 * stats->xinf = ||u - u_c||_inf (Integrator.cpp:28).  u = NULL clears the
 * commit (u_c = 0, the state after efem_create).  Single-problem contexts;
 * the residual and the factor do not depend on it. */
int efem_commit(efem_ctx* ctx, const double* u);
/* copies u_c out (zeros when nothing is committed) */
int efem_committed(efem_ctx* ctx, double* u);

/* Model::saveVTK (src/model/Model.cpp:50-117), host only: the grid at
 * X + u (u over the free DOFs as above, NULL for the undeformed frame) as an
 * ASCII legacy VTK unstructured grid: every node a FP32 point (vtkPoints'
 * default type), every element a cell with its node list unchanged and type
 * VTK_TETRA / VTK_QUADRATIC_TETRA / VTK_HEXAHEDRON / VTK_QUADRATIC_HEXAHEDRON
 * for nen 4 / 10 / 8 / 20.  Errors through efem_global_error. */
int efem_write_vtk(const efem_problem* pb, const double* u, const char* path);
/* the reference's results file name, directories created:
 * <working_dir>/results/YYYY-MM-DD/YYYY-MM-DD-HH-MM-SS-mmm-<name>.vtk in
 * local time (Model.cpp:93-110); out gets the NUL-terminated path */
int efem_results_path(const char* working_dir, const char* name, char* out, int cap);

/* host-only diagnostic (no GPU): assemble, order and factor the approximate
 * Jacobian on the host, then run the device's block-inverse triangular-solve
 * algorithm on the host: x = K^-1 b (interleaved).  info (may be NULL) gets
 * {nnz(L), supernodes, levels, factor flops, factor us (symbolic+numeric),
 *  numeric us, Laplacian assembly us, ordering us}. */
int efem_host_factor_solve(const efem_problem* pb, const double* b, double* x,
                           long long* info8);

/* device memory of the current HIP device (hipMemGetInfo): free and total
 * bytes; lets a caller check that refactorisations replace, not add to, the
 * factor's device buffers */
int efem_device_memory(long long* free_bytes, long long* total_bytes);
/* the approximate Jacobian as triplets, host only
 * (ElementConstraint::conEvaluteJacobi / conGetJacobi,
 * src/fe/ElementConstraint.cpp:164-210): K = L (x) I3 over the free DOFs, L
 * the elasticity_coeff-weighted scalar Laplacian; every 3 x 3 block of a
 * nonzero L_ij with L_ij on its diagonal and explicit zeros off it, each
 * entry once (the matrix setFromTriplets builds from the reference's
 * per-element triplets).  *nnz = the entry count; rows, cols, vals (cap
 * entries) NULL to query it.  Errors through efem_global_error. */
int efem_host_jacobi(const efem_problem* pb, long long* nnz, int* rows, int* cols, double* vals, long long cap);

/* host-only diagnostic (no GPU): the constant part of the external force on
 * the free DOFs as the residual subtracts it -- nodal forces
 * (ExternalForceConstraint) plus gravity M g -- into out[3 * n_free] */
int efem_host_constant_force(const efem_problem* pb, double* out, int n);

/* host-only diagnostic: the dense micro-kernel the host factorisation uses
 * ("avx512 16x8", "avx2 8x6" or "portable 8x6"; the environment variable
 * EFEM_DENSE_KERNEL=avx2|portable, read once per process, caps the choice) */
const char* efem_host_dense_kernel(void);

/* tuning / debug options:
 *   "h0_graph"          0 (default): the preconditioner apply's level
 *                       kernels are launched one by one; 1: replayed from a
 *                       captured HIP graph (same kernels, same results)
 *   "two_loop_compact"  1 (default, history m <= 16): the two-loop recursion
 *                       is evaluated from Gram products s_i.y_k in two
 *                       multi-dot passes; 0: one reduction per step exactly
 *                       as LBFGSMat::apply_Hv orders them
 *   "lookahead"         1 (default): while the host reads a first line-search
 *                       trial's scalars the device already forms the next
 *                       direction and runs the next first trial, discarded
 *                       when the trial is rejected or the solve stops; the
 *                       iterates are those of 0 (no look-ahead) bit for bit
 *   "device_laplacian"  1 (default): the approximate Jacobian is assembled on
 *                       the device (bit-identical to the host assembly);
 *                       0: on the host; releases the factor
 *   "fwd_tables"        1 (default): forward triangular-solve levels with at
 *                       most 4 contributions per front row read their
 *                       children's update values through per-item tables
 *                       (same sums, same order); 0: through the row lists;
 *                       releases the factor
 *   "device_factor"     1 (default): with "device_laplacian", the numeric
 *                       factorisation (assembly of the fronts, dense
 *                       partial Cholesky, update matrices, block inverses)
 *                       runs on the device from the host's ordering and
 *                       symbolic analysis (the same factor up to rounding);
 *                       0: on the host; releases the factor
 *   "fused_root"        1 (default): when the elimination tree's last level
 *                       is one supernode R whose rows are its columns, its
 *                       forward and backward solves run as one launch
 *                       z_R = M f_R with M = L_RR^-T L_RR^-1 formed at the
 *                       factorisation (the same K^-1 up to rounding, two
 *                       launches fewer per apply); 0: level by level;
 *                       releases the factor
 *   "wide_fwd_items", "wide_bwd_items"  (defaults 500, 1000): levels of wide
 *                       (tall) supernodes with at most this many forward
 *                       (backward) items use 16-wave items; release the factor
 *   "dense_subtree"     1..768 (default 64): elimination subtrees of at most
 *                       this many columns become one dense supernode (fewer
 *                       triangular-solve levels, more stored entries);
 *                       releases the factor (the next solve factors again)
 *   "bwd_narrow_items"  >= 1 (default 4096): levels of the backward solve with
 *                       at least this many 4-column work items use
 *                       16-column items instead; releases the factor (the
 *                       next solve factors again)
 *   "residual_terms"    1..7 (default 7): efem_residual evaluates only the
 *                       terms in the mask -- 1 element internal forces, 2
 *                       follower pressure / directed traction, 4 constant
 *                       nodal forces and gravity -- one constraint's part
 *                       (Constrainable::conEvaluteResidual,
 *                       src/fe/Constraint.h:26); solves use every term
 *   "tet4_patches"      0 / 1 (default 1; C3D4, single-problem contexts):
 *                       the residual sums the element forces of patches of
 *                       256 spatially close elements in LDS and the gather
 *                       reads one FP64 record per (patch, node) instead of
 *                       one slot per (element, node); the same forces, the
 *                       sums in another fixed order
 *   "patch_gather_lanes" 1, 2, 4, 8 (default 2): lanes per node of that
 *                       gather
 *   "batch_dense_h0"    0 / 1 (default 1; batch contexts): the persistent
 *                       solver applies each instance's preconditioner as one
 *                       product with its dense Z_i = (L_i L_i^T)^-1 (formed
 *                       on the device at each factorisation; instances of at
 *                       most 1024 free nodes) instead of the level sweeps;
 *                       releases the factor (the next solve factors again)
 *   "gather_lanes", "tet4_threads", "reduce_blocks", "history_blocks":
 *                       launch shapes (see efem_autotune) */
int efem_set_option(efem_ctx* ctx, const char* name, double value);

/* autotuner (no counterpart in the reference): times the candidate launch
 * shapes on this context's problem and keeps the fastest -- gather lanes per
 * node {1,2,4,8} x C3D4 element block size {64,128,256} (residual time; on
 * the C3D4 patch path the record gather's lanes {1,2,4,8}), the
 * reduction grid {32,64,128,256} blocks (line-search dot time), and the grid
 * of the reductions over the L-BFGS history {32,64,128,256} blocks (one
 * compact two-loop pass pair with m = 8; grows the context's history buffers
 * to m = 8 if smaller).  The choice only changes summation orders, never the
 * formulation.  On a batch context it also times the persistent solver
 * (40 iterations of every instance from u = 0, with the preconditioner)
 * at 256, 512 and 1024 lanes per instance and keeps the fastest.
 * report[9] (may be NULL) = {lanes (the record gather's on the C3D4 patch
 * path), tet4 threads, reduce blocks, residual
 * ms, dot ms, history blocks, history ms, batch lanes (256 when not a
 * batch), batch ms (0 when not a batch)}.  The same shapes can be set by
 * efem_set_option("gather_lanes" | "tet4_threads" | "reduce_blocks" |
 * "history_blocks" | "batch_threads", v). */
int efem_autotune(efem_ctx* ctx, double* report);
/* autotuner of the preconditioner's structure (no counterpart in the
 * reference; single-problem contexts): factors the approximate Jacobian
 * with dense elimination subtrees of at most 64, 128, 192 and 256 columns
 * (option "dense_subtree"), times 20 applies of each, keeps the fastest (a
 * larger size must be 1 % faster) and leaves the context factored with it.
 * Every candidate is the same K^-1 up to rounding.  report[4] (may be NULL)
 * = {dense subtree, apply ms, factorisation s (with its ordering and
 * symbolic analysis), stored factor entries}.  EFEM_ERR_STATE when the
 * matrix is not SPD. */
int efem_autotune_factor(efem_ctx* ctx, double* report);

/* per-iteration trace of the last efem_solve: {||r||_inf, step, trials} */
int efem_trace(const efem_ctx* ctx, double* out, int cap);

/* device timing (HIP events on the context stream) of the element-force
 * kernel and of whole H0 applies (forward + backward triangular solves) made
 * after efem_profile_reset, by efem_residual / efem_apply_h0 / efem_solve */
int efem_profile_reset(efem_ctx* ctx);
int efem_profile_read(efem_ctx* ctx, double* elem_ms_total, int* elem_launches);
int efem_profile_read_h0(efem_ctx* ctx, double* h0_ms_total, int* h0_applies);
/* timing helper: `count` back-to-back residual evaluations at u (uploaded
 * once) on the device; mean device time per evaluation (events around all
 * element, pressure and gather launches) */
int efem_time_residual(efem_ctx* ctx, const double* u, int count, double* ms_per_eval);

/* ---- differentiation (no counterpart in the reference) -----------------
 * K_t(u) = d r / d u is the true tangent: the exact material tangent of the
 * element forces (symmetric) minus the follower-pressure load stiffness
 * (not symmetric).  The reference only builds the constant approximate
 * Jacobian (ElementConstraint.cpp:164-210); its never-called pressure
 * stiffness sketch is at ExternalConstraint.cpp:856-869. */
#define EFEM_KRYLOV_AUTO 0     /* CG without pressure faces, else BiCGStab */
#define EFEM_KRYLOV_CG 1
#define EFEM_KRYLOV_BICGSTAB 2

typedef struct efem_krylov_params {
    double tol;          /* stop at ||b - K x||_2 <= tol ||b||_2 */
    int max_iterations;
    int method;          /* EFEM_KRYLOV_* */
    int use_precond;     /* 1: preconditioned by the approximate-Jacobian
                            Cholesky factor (factored here if absent) */
} efem_krylov_params;

typedef struct efem_krylov_stats {
    int iterations;
    int matvecs;         /* tangent products */
    int precond_applies;
    int converged;
    double rel_residual; /* recurrence ||r|| / ||b|| at exit */
    double seconds;      /* device solve (+ gradient assembly for the adjoint) */
    double factor_seconds;
} efem_krylov_stats;

/* out = K_t(u) v, or K_t(u)^T v when transpose != 0 (host pointers) */
int efem_tangent(efem_ctx* ctx, const double* u, const double* v, double* out,
                 int transpose);
/* K_t(u) x = b (transpose: K_t^T x = b) by preconditioned Krylov */
int efem_tangent_solve(efem_ctx* ctx, const double* u, const double* b,
                       double* x, int transpose, const efem_krylov_params* prm,
                       efem_krylov_stats* stats);
/* Newton-Krylov equilibrium solve from u (in/out): K_t(u) du = -r(u) by
 * efem_tangent_solve to the relative tolerance max(kp->tol, min(1e-2,
 * 0.9 (||r_k|| / ||r_k-1||)^2)), backtracking on ||r||_2, until
 * ||r||_inf <= epsilon.
 * stats: iterations = Newton steps, ls_trials = Krylov iterations summed,
 * r_evals, rinf, xinf, factor/solve seconds. */
int efem_newton(efem_ctx* ctx, double* u, double epsilon, int max_iterations,
                const efem_krylov_params* kp, efem_solve_stats* stats);
/* Adjoint gradient of an objective J(u) at an equilibrium u (r(u) = 0):
 * solves K_t(u)^T lambda = g with g = dJ/du, then
 *   grad_params = -lambda . dr/dparams: [n_elems][4] when the problem has
 *                 per-element parameters, else [4] (same slots as params);
 *   grad_load   = -lambda . dr/dload: [0] = dJ/dpressure (follower), or
 *                 [0..2] = dJ/dpdir (directed traction).
 * lambda, grad_params, grad_load may be NULL. */
int efem_adjoint(efem_ctx* ctx, const double* u, const double* g, double* lambda,
                 double* grad_params, double* grad_load,
                 const efem_krylov_params* prm, efem_krylov_stats* stats);
/* total strain energy sum_e sum_q w_q V_e W(F_q) (FP64, compensated sum) */
int efem_energy(efem_ctx* ctx, const double* u, double* W);
/* efem_energy for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * No counterpart in the reference.  u_d: ndof doubles on the device; W_d:
 * one double on the device; the same kernels and summation order as
 * efem_energy, so the same value bit for bit.  Calls no hipMalloc / hipFree:
 * its work memory is one 16 MiB array per GPU (2^21 doubles, held until the
 * process ends), so it needs roundup32(ndof + 3) + roundup32(n_elems) + 512
 * doubles to fit (EFEM_ERR_ARG otherwise).  Every call on a GPU shares that
 * GPU's array: calls that overlap on one GPU (from several threads or
 * contexts) race.  Returns once W_d is written (context stream
 * synchronised). */
int efem_energy_noalloc(efem_ctx* ctx, const double* u_d, double* W_d);
/* replace the material parameters (same layout as efem_problem.params);
 * drops the factor, which depends on them */
int efem_set_params(efem_ctx* ctx, const double* params);
/* (both setters return EFEM_ERR_STATE on a batch context, whose instances'
 * parameters and loads are fixed at efem_batch_create) */
/* replace the follower pressure (pdir NULL) or the directed traction */
int efem_set_load(efem_ctx* ctx, double pressure, const double* pdir);
/* one material point, host only: P(F) (Material::getFirstPiola,
 * src/material/Material.h:18-28), W(F), and dP = (dP/dF)[dF]
 * (row-major 3x3; any output may be NULL; dP needs dF) */
int efem_material_point(int law, const double* params, const double* F,
                        double* P, double* W, const double* dF, double* dP);

/* ---- batches of independent instances (no counterpart in the reference) --
 * efem_batch_create merges `count` problems (same element kind, integration
 * rule and law; any sizes, parameters and follower pressures) into one
 * context.  efem_residual / efem_factor / efem_apply_h0 then act on the
 * concatenated free-DOF vector (instance i at dof_offsets[i]); each instance
 * is factored on its own, exactly as in a single-instance context.
 * efem_batch_solve runs LBFGSSolver::solve (LBFGS.h:96-174) for every
 * instance in ONE persistent launch, one workgroup per instance (256, 512
 * or 1024 lanes: option "batch_threads", picked by efem_autotune on a batch
 * context), with per-instance L-BFGS state; prm->m <= 16,
 * prm->max_iterations > 0; prm->x_target (concatenated like u) stops each
 * instance at its own ||x_i - x_target_i|| <= target_tol ||x_target_i||.
 * u: concatenated in/out; stats: [count].  Returns
 * EFEM_ERR_NONFINITE (with those instances' u zeroed) if any instance
 * produced NaN/Inf.  efem_solve is refused on a batch context. */
int efem_batch_create(efem_ctx** out, const efem_problem* problems, int count,
                      int precision);
int efem_batch_count(const efem_ctx* ctx);
int efem_batch_offsets(const efem_ctx* ctx, int* dof_offsets /* [count+1] */);
int efem_batch_solve(efem_ctx* ctx, double* u, const efem_lbfgs_params* prm,
                     efem_solve_stats* stats);

/* ---- synthetic workloads (host only, no GPU needed) --------------------
 * which: 1 cantilever 20x4x4 C3D8 NH; 2 beam 200x20x20 Kuhn C3D4 MR;
 *        3 hollow cylinder 48x4x40 C3D20R Yeoh; 4 bar 64x16x12 Kuhn C3D10
 *        NH per-element; 5 one cantilever of the batch (opts = {index,
 *        seed, load scale}: Lx4x4 C3D8, L ~ U{10..40}, C10 ~ U[80k,150k],
 *        p ~ U[0.5,5] kPa times the scale);
 *        0 custom box (dims from opts).
 * opts (may be NULL) = {nx, ny, nz, seed, ...} overrides, see efem_mesh_build. */
typedef struct efem_mesh efem_mesh;
efem_mesh* efem_mesh_build(int which, const double* opts, int n_opts);
void efem_mesh_free(efem_mesh* m);
/* sizes: {n_nodes, n_elems, nen, reduced, law, params_per_elem, n_faces, fnn} */
int efem_mesh_sizes(const efem_mesh* m, long long* sizes8);
/* copies: X [n_nodes*3], conn [n_elems*nen], params [4 or n_elems*4],
 * fixed [n_nodes], faces [n_faces*fnn], scalars {pressure} */
int efem_mesh_arrays(const efem_mesh* m, double* X, int* conn, double* params,
                     unsigned char* fixed, int* faces, double* pressure);
/* fill an efem_problem that points into the mesh (valid while m lives) */
int efem_mesh_problem(const efem_mesh* m, efem_problem* pb, int precision);

/* ---- reference input formats (host only, no GPU needed) -----------------
 * efem_input_load reads an EFEM run directory: the JSON config's keys as
 * Config::Config reads them (src/io/Config.cpp:23-81: analysis.nl_solver,
 * max_solver_iters, max_line_search, dx_criterion, residual_criterion,
 * gravity, consistent_inertia), model[0].mesh_file (MeshParser::parse,
 * src/io/MeshParser.h:22-100; element rows by Body::buildElement,
 * src/model/Body.cpp:49-90), material[] (neo_hookean / mo_riv / yeoh by id,
 * ElementConstraint.cpp:11-67; elements without a defined material are
 * skipped), constraint[] fix / force / pressure with constraint_file and
 * value (ExternalConstraint.cpp:50-247; files read as IO::readFile,
 * src/io/IO.cpp:67-96).  Paths are relative to the JSON's directory.  One
 * body, one element kind, one pressure and one force constraint; Neo-Hookean
 * elements mixed with Mooney-Rivlin (or Yeoh) ones become Mooney-Rivlin
 * (Yeoh) elements with C01 = 0 (C20 = C30 = 0), which is the same stress and
 * approximate-Jacobian coefficient.  NULL on error (efem_input_error). */
typedef struct efem_input efem_input;
efem_input* efem_input_load(const char* json_path);
const char* efem_input_error(void);
void efem_input_free(efem_input* in);
/* {n_nodes, n_elems, nen, law, params_per_elem, n_faces, fnn, n_fixed,
 *  n_force_nodes, skipped_elements} */
int efem_input_sizes(const efem_input* in, long long* sizes10);
/* a problem pointing into the input (valid while it lives) */
int efem_input_problem(const efem_input* in, efem_problem* pb, int precision);
/* solver parameters of the analysis block (Integrator.cpp:6-20: epsilon =
 * epsilon_rel = residual_criterion; m = 6; projection line search) */
int efem_input_lbfgs_params(const efem_input* in, efem_lbfgs_params* prm);

#ifdef __cplusplus
}
#endif
#endif
