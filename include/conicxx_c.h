/*
 * conicxx_c.h -- C-ABI of the MI355X conic interior-point solver.
 *
 * One opaque handle per problem; plain pointers and sizes; CSC matrices with
 * int32 indices and double values (column-major, P upper triangle only);
 * cone rows in the fixed order Zero, Nonnegative, SOC...  No exception
 * crosses this boundary: every call returns a status or an int code.
 * A handle is not thread-safe; distinct handles are independent.
 */
#ifndef CONICXX_C_H
#define CONICXX_C_H

#ifdef __cplusplus
extern "C" {
#endif

/* Status values in the order of enum class Status (include/conicxx/solution.h:9-16). */
enum cx_status {
  CX_UNSOLVED = 0,
  CX_SOLVED = 1,
  CX_PRIMAL_INFEASIBLE = 2,
  CX_DUAL_INFEASIBLE = 3,
  CX_MAX_ITERATIONS = 4,
  CX_NUMERICAL_ERROR = 5
};

/* LinearSolverBackend (include/conicxx/settings.h:8-17).  All three select the
 * device supernodal LDL^T; Eigen and Qdldl use the whole-matrix retry ladder on a
 * failed pivot check, RegularizedLdlt corrects each pivot inline (5% cap). */
enum cx_linear_solver { CX_LS_EIGEN = 0, CX_LS_QDLDL = 1, CX_LS_REGULARIZED = 2 };

/* Numeric factorization precision (no counterpart in the reference, which is FP64 only). */
enum cx_factor_precision { CX_FACTOR_FP64 = 0, CX_FACTOR_FP32_REFINED = 1 };

/* Field-for-field the Settings aggregate (include/conicxx/settings.h:21-62), same
 * defaults; the last field extends it. */
typedef struct cx_settings {
  double tol_feas;
  double tol_gap_abs;
  double tol_gap_rel;
  double tol_infeas;
  int max_iter;
  double static_reg_P;
  double static_reg_A;
  double dynamic_reg_eps;
  double dynamic_reg_delta;
  int refine_max_iter;
  double refine_tol;
  double max_step_fraction;
  int equilibrate;
  int equilibrate_max_iter;
  double equilibrate_min_scale;
  double equilibrate_max_scale;
  int warm_start;
  int verbose;         /* accepted, unused (as in the reference) */
  int record_timings;  /* accepted, unused */
  int validate_inputs; /* accepted, unused */
  int linear_solver;   /* enum cx_linear_solver */
  int factor_precision; /* enum cx_factor_precision */
  double fp32_reg;      /* FP32 factorization: diagonal shift and pivot floor (precision form) */
} cx_settings;

/* Info (include/conicxx/solution.h:20-28). */
typedef struct cx_info {
  int iterations;
  double primal_residual;
  double dual_residual;
  double duality_gap;
  double mu;
  double setup_time_s;
  double solve_time_s;
} cx_info;

typedef struct cx_solver cx_solver;

/* Settings{} defaults (settings.h:23-61). */
void cx_default_settings(cx_settings* s);

/* Solver::Solver / ~Solver (include/conicxx/solver.h:36-37).  NULL if no GPU. */
cx_solver* cx_create(void);
void cx_destroy(cx_solver* h);

/* Solver::setup(P, q, A, b, cone_spec, settings) (solver.h:46-47; solver_impl.cpp:8-55).
 * P: n x n upper-triangular CSC (lower entries ignored by the KKT, as in the
 * reference); A: m x n CSC; m must equal zero_dim + nonneg_dim + sum(soc_dims).
 * Copies all data.  Returns 1 on success, 0 on invalid input or KKT setup failure. */
int cx_setup(cx_solver* h, int n, int m, const int* P_colptr, const int* P_rowind,
             const double* P_val, const double* q, const int* A_colptr, const int* A_rowind,
             const double* A_val, const double* b, int zero_dim, int nonneg_dim, int n_soc,
             const int* soc_dims, const cx_settings* settings);

/* Solver::updateData(P*, q*, A*, b*) (solver.h:77; solver_impl.cpp:57-91;
 * kkt_system.cpp:10-19,119-137).  A NULL value pointer leaves that piece
 * unchanged; P_colptr/P_rowind (A_colptr/A_rowind) must accompany P_val (A_val)
 * for the pattern check.  Returns 0 on pattern mismatch or factorization failure. */
int cx_update_data(cx_solver* h, const int* P_colptr, const int* P_rowind, const double* P_val,
                   const double* q, const int* A_colptr, const int* A_rowind,
                   const double* A_val, const double* b);

/* cx_update_data without any allocation.  No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same arguments,
 * checks, K updates and results as cx_update_data, but every staged value
 * goes through one static device array per GPU (1,048,576 doubles = 8 MiB,
 * held by each GPU the library runs on until the process ends) instead of
 * freshly allocated buffers, and it calls no hipMalloc/hipFree.  Each of
 * nnz(P), nnz(A), n and m (of the pieces passed) must be at most 1,048,576:
 * a larger piece is rejected (returns 0, stderr message) before anything
 * changes.  Every call on a GPU shares that GPU's array: calls that overlap
 * on one GPU (other threads, other handles) race; sequential calls are safe. */
int cx_update_data_noalloc(cx_solver* h, const int* P_colptr, const int* P_rowind, const double* P_val,
                           const double* q, const int* A_colptr, const int* A_rowind,
                           const double* A_val, const double* b);

/* Solver::setWarmStart(x, s, z) in original units (solver.h:81; solver_impl.cpp:93-100). */
void cx_set_warm_start(cx_solver* h, const double* x, const double* s, const double* z);

/* Solver::setSettings / settings() (solver.h:83-84). */
void cx_set_settings(cx_solver* h, const cx_settings* settings);
void cx_get_settings(const cx_solver* h, cx_settings* out);

/* Solver::solve() (solver.h:89; solver_impl.cpp:367-466).  Returns the status. */
int cx_solve(cx_solver* h);

/* Solver::solution() (solver.h:91): copies x (n), s (m), z (m), objective, info
 * (any pointer may be NULL).  Returns the status. */
int cx_get_solution(const cx_solver* h, double* x, double* s, double* z, double* objective,
                    cx_info* info);

/* toString(Status) (src/conicxx/solution.cpp:5-21). */
const char* cx_status_string(int status);

/* Problem sizes of the last setup (n, m) and symbolic statistics of the KKT plan:
 * stats[0..7] = N, nnz(K lower), supernodes, levels, nnz(L) stored, factor flops,
 * largest front, nnz(L) exact.  No counterpart in the reference. */
int cx_get_dims(const cx_solver* h, int* n, int* m);
void cx_get_kkt_stats(const cx_solver* h, double* stats8);

/* Breakdown of the last solve: [0] numeric factorizations (ms, host-timed incl.
 * the pivot check), [1] KKT solves incl. refinement (ms), [2] everything else (ms),
 * [3] number of factorizations, [4] number of KKT solves (incl. refinement steps),
 * [5] average device duration of the dense trailing-update kernel (ms), [6] its
 * launches, [7] its algorithmic flops, [8] FP32 factorizations, [9] FP32 -> FP64
 * fallbacks (factor_precision = CX_FACTOR_FP32_REFINED), [10] largest final
 * relative residual |rhs - K x| / |rhs| of an accepted KKT solve, [11] that of
 * the FP32 solve that triggered the fallback (0 if none), [12] the IPM
 * iteration of the fallback (-1 if none).  [5..7] need CONICXX_KERNEL_TIMING=N
 * (N >= 1) in the environment at setup().  No counterpart in the reference. */
void cx_get_profile(const cx_solver* h, double* out13);

/* Device time per kernel class of the last solve, measured with HIP events on the
 * solver's stream around the launches of every Nth numeric factorization and
 * every Nth KKT solve (CONICXX_KERNEL_TIMING=N, N >= 1, at setup(); the
 * schedule of launches is the same in every factorization): for class k,
 * out[3k..3k+2] = (milliseconds, the timed launches' total scaled to all
 * launches; launches; algorithmic work of all launches); work is flops for the trailing
 * update (class 0) and bytes for the forward/backward sweeps (classes 4, 5).
 * Classes: 0 trailing update, 1 sub-panel (diagonal block LDL^T + panel solve,
 * and the store of the diagonal blocks), 2 assembly, 3 small fronts, 4 forward
 * sweep, 5 backward sweep.  A checked factorization that stopped at a failing
 * pivot is not counted (its later launches return at once).  Returns the
 * number of classes written.  No counterpart in the reference. */
int cx_get_kernel_stats(const cx_solver* h, double* out, int max_kinds);

/* Backward (adjoint) pass of the last solve, which must have returned
 * CX_SOLVED.  Given dL/dx (n), dL/ds (m), dL/dz (m) in original units (any may be
 * NULL = zero), writes dL/dP on P's stored pattern (nnz(P) values in setup's CSC
 * order; entries below the diagonal get 0), dL/dq (n), dL/dA on A's pattern
 * (nnz(A)) and dL/db (m); any output may be NULL.  One adjoint KKT solve that
 * reuses the final factorization, then the rank-one gradient assembly on the
 * device.  Returns 1 on success, 0 if there is no solved problem or the solve
 * failed.  No counterpart in the reference (the reference has no derivatives). */
int cx_backward(cx_solver* h, const double* dl_dx, const double* dl_ds, const double* dl_dz,
                double* dP, double* dq, double* dA, double* db);

/* KktSystem::updateScalingAndFactorize followed by KktSystem::solve
 * (src/conicxx/kkt/kkt_system.cpp:139-150, 296-311), as the reference's own KKT
 * tests drive them (test/kkt/test_kkt_system.cpp:32-60): on the data of the last
 * setup()/updateData(), writes the NT scaling at (s, z) into the KKT matrix
 * (s or z NULL: the identity scaling of construction), factorizes it with the
 * x10 retry ladder and solves K x = rhs (N = n + m entries) with up to
 * refine_max_iter refinement steps (negative: settings.refine_max_iter), with
 * the x1000 escalation on a non-finite solve.  K is the statically regularized
 * [P + static_reg_P I, A'; A, -(Hs + static_reg_A I)] in the solver's internal
 * space (the original units when settings.equilibrate = 0).  factor_precision
 * applies as in solve().  Writes the number of numeric factorizations it ran
 * (retries and escalations included) to *factorizations when non-NULL.  Returns
 * |rhs - K x| / |rhs|, NaN if no factorization was accepted or the handle has
 * no setup. */
double cx_kkt_factor_solve(cx_solver* h, const double* s, const double* z, const double* rhs, double* x,
                           int refine_max_iter, int* factorizations);

/* ---- Batched form: many independent problems in one persistent launch ----
 * No counterpart in the reference (callers there loop over Solver objects,
 * README.md:11-14).  Problem k has the same arguments as cx_setup, passed as
 * arrays of per-problem pointers; every problem is solved start to finish by
 * one workgroup of a single persistent kernel, per-problem iteration counts and
 * statuses are independent, and solve() runs as one captured HIP graph
 * (staged q/b upload, rescaling, solve) that is replayed after update_data().
 * The factor of a problem lives in LDS when it fits, otherwise in global memory. */
typedef struct cx_batch cx_batch;
cx_batch* cx_batch_create(void); /* NULL if no GPU */
void cx_batch_destroy(cx_batch* h);
/* Returns 1 on success, 0 on invalid input (any problem) or device failure. */
int cx_batch_setup(cx_batch* h, int nprob, const int* n, const int* m, const int* const* P_colptr,
                   const int* const* P_rowind, const double* const* P_val, const double* const* q,
                   const int* const* A_colptr, const int* const* A_rowind, const double* const* A_val,
                   const double* const* b, const int* zero_dim, const int* nonneg_dim, const int* n_soc,
                   const int* const* soc_dims, const cx_settings* settings);
/* Solver::updateData(P*, q*, A*, b*) (solver.h:77; solver_impl.cpp:57-91) for
 * every problem at once, with per-problem pointer arrays laid out as in
 * cx_batch_setup.  An array, or an entry of one, that is NULL leaves that piece
 * of that problem unchanged; P_val[k] (A_val[k]) must come with
 * P_colptr[k]/P_rowind[k] (A_colptr[k]/A_rowind[k]) equal to setup's pattern
 * (kkt_system.cpp:10-19).  Every pattern is checked before anything changes:
 * returns 0 on any mismatch, 1 otherwise.  New values are scaled by setup's
 * equilibration (equilibration.cpp:102-121), staged on the host and uploaded
 * by the next cx_batch_solve, whose per-problem factorizations use them (the
 * reference refactorizes inside updateData; here every solve factors anyway). */
int cx_batch_update_data(cx_batch* h, const int* const* P_colptr, const int* const* P_rowind,
                         const double* const* P_val, const double* const* q, const int* const* A_colptr,
                         const int* const* A_rowind, const double* const* A_val, const double* const* b);
/* Solves every problem; writes per-problem statuses (nprob) when non-NULL and
 * returns the number of Solved problems (-1 on device failure).  With
 * settings.warm_start each Solved problem seeds its next solve. */
int cx_batch_solve(cx_batch* h, int* status);
/* Solution of problem k, as cx_get_solution.  Returns its status. */
int cx_batch_get_solution(cx_batch* h, int k, double* x, double* s, double* z, double* objective,
                          cx_info* info);
void cx_batch_set_settings(cx_batch* h, const cx_settings* settings);
/* Times the candidate launch configurations on this batch (reps solves each)
 * and keeps the fastest; returns the number of candidates tried. */
int cx_batch_autotune(cx_batch* h, int reps);
/* Pins a launch configuration (no counterpart in the reference: the launch
 * configuration is this implementation's own): lds_doubles of dynamic LDS per
 * workgroup (0 .. the device maximum; factors and vectors that do not fit live
 * in global memory) and grid persistent workgroups (>= 1).  Returns 1, or 0 on
 * an out-of-range value (nothing changed). */
int cx_batch_set_tuning(cx_batch* h, int lds_doubles, int grid);
/* out4 = (packed M doubles kept in LDS, persistent workgroups, largest packed M,
 * autotuned flag). */
void cx_batch_get_tuning(const cx_batch* h, double* out4);
/* Work counters of problem k in the last solve: out13 = (iterations,
 * factorizations, reduced solves incl. refinement, device microseconds its
 * workgroup spent on it, of which in factorizations, in triangular solves, in
 * refinement residuals, KKT solves, device microseconds of its last backward,
 * of which in its KKT solve, reduced solves of that KKT solve, device
 * microseconds of the final IPM iteration's KKT solve, device microseconds of
 * the adjoint factorization at the end of a differentiable forward pass (0
 * without one; its factorizations are among the forward's)). */
void cx_batch_get_stats(cx_batch* h, int k, double* out13);
/* Differentiation of the batch (no counterpart in the reference).  With it on,
 * every Solved problem's forward pass also factors its adjoint KKT matrix (the
 * final scaling with the SOC tangent correction, as cx_backward uses), so that
 * cx_batch_backward is one KKT solve plus the gradient assembly per problem.
 * Gradient layout per problem as in cx_backward; any array (or entry) may be
 * NULL.  Returns the number of problems with gradients (the Solved ones; the
 * others get zeros), -1 if differentiation is off or on device failure. */
void cx_batch_set_differentiable(cx_batch* h, int on);
int cx_batch_backward(cx_batch* h, const double* const* dl_dx, const double* const* dl_ds,
                      const double* const* dl_dz, double* const* dP, double* const* dq, double* const* dA,
                      double* const* db);
/* Device time of the last cx_batch_backward (ms, HIP events around the launch). */
double cx_batch_last_backward_ms(const cx_batch* h);
/* Device time of the last cx_batch_solve (ms, HIP events around the graph). */
double cx_batch_last_solve_ms(const cx_batch* h);

/* Host-only Ruiz equilibration of one problem as setup() computes it
 * (Equilibration::compute, src/conicxx/kkt/equilibration.cpp:8-99), for
 * testing: writes the column scale d (n), the row scale e (m) and the cost
 * scale c for the given data and settings (NULL = defaults).  Returns 1 on
 * success, 0 on invalid sizes. */
int cx_equilibrate(int n, int m, const int* P_colptr, const int* P_rowind, const double* P_val,
                   const int* A_colptr, const int* A_rowind, const double* A_val, int zero_dim,
                   int nonneg_dim, int n_soc, const int* soc_dims, const cx_settings* settings, double* d,
                   double* e, double* c);

/* Host-only symbolic analysis of a symmetric pattern (lower CSC, diagonal
 * included), for testing: writes the final permutation (new->old, N ints) and
 * the column counts of L (N ints) when non-NULL, and stats8 as above.
 * Returns 1 on success.  Replaces qdldl_ldlt.cpp:22-56 (ordering + etree + Lnz). */
int cx_analyze_pattern(int N, const int* colptr, const int* rowind, int* perm, int* colcount,
                       double* stats8);

/* Input conversions to the compressed-column layout (host only, no GPU
 * needed; conicxx::toSparse / toSparseUpperTriangular, include/conicxx/
 * convert.h:17-41, src/conicxx/convert.cpp:22-79).  The caller allocates
 * colptr (cols + 1) and rowind / values (the returned entry count: rows*cols,
 * n(n+1)/2, or the CSR's nnz).  Each prints a one-line warning on stderr
 * unless warn is 0 (name, if not NULL, names the matrix in it).  Return the
 * number of entries written, or -1 on invalid arguments.
 *   cx_dense_to_csc: every entry explicit, exact zeros included (a pattern
 *     independent of the values, as updateData needs);
 *   cx_dense_to_csc_upper: the entries i <= j of a square matrix (for P), with
 *     an extra warning when it is not symmetric to 1e-8 of its largest entry;
 *   cx_csr_to_csc: a row-major matrix re-laid out column by column.
 * dense is column-major when col_major != 0, else row-major. */
long cx_dense_to_csc(int rows, int cols, const double* dense, int col_major, const char* name, int warn,
                     int* colptr, int* rowind, double* values);
long cx_dense_to_csc_upper(int n, const double* dense, int col_major, const char* name, int warn, int* colptr,
                           int* rowind, double* values);
long cx_csr_to_csc(int rows, int cols, const int* rowptr, const int* colind, const double* values_in,
                   const char* name, int warn, int* colptr, int* rowind, double* values);

#ifdef __cplusplus
}
#endif

#endif /* CONICXX_C_H */
