/*
 * conic_oracle.h -- CPU oracle for the conic primal-dual interior-point path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C, single-threaded, FP64
 * restatement of the reference method (homogeneous self-dual embedding,
 * Nesterov-Todd scaling, Mehrotra predictor-corrector, quasi-definite KKT
 * factorized by an up-looking sparse LDL^T with static + dynamic
 * regularization and iterative refinement).  Only the test-suite, the
 * smoke entry and the CPU-baseline leg of the benchmark may load it; the
 * product library never links or calls it.
 *
 * Parity pinning (what checks each part):
 *   - cones (product / inverse product / NT scaling / maxStep / margins):
 *     the reference's own unit-test vectors (known answers) plus the NT
 *     identities Hs*z == s and W^-1 s == W z;
 *   - equilibration: identity when disabled, KKT residual preservation under
 *     scale/unscale, uniform SOC block scale;
 *   - KKT solve: dense FP64 LDL^T / LU of the same matrix (numpy) to 1e-8;
 *   - end-to-end: closed-form known answers of the reference's solver tests
 *     and the KKT optimality certificate on random feasible instances;
 *   - ordering (own exact minimum degree, replacing the reference's AMD):
 *     parity unpinned -- it changes rounding only, never the result beyond it.
 */
#ifndef CONIC_ORACLE_H
#define CONIC_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Arithmetic type of the whole oracle: double (liboracle.so, the restatement
 * of the FP64 reference), or long double (liboracle_xp.so, built with
 * -DOC_REAL="long double": the same method in x87 extended precision, used by
 * the tests only to compute the problems' optima to far below the FP64
 * method's own rounding-level spread). */
#ifndef OC_REAL
#define OC_REAL double
#endif
typedef OC_REAL oc_real;

/* Field-for-field the reference Settings aggregate (include/conicxx/settings.h:77-118). */
typedef struct oc_settings {
  oc_real tol_feas;
  oc_real tol_gap_abs;
  oc_real tol_gap_rel;
  oc_real tol_infeas;
  int max_iter;
  oc_real static_reg_P;
  oc_real static_reg_A;
  oc_real dynamic_reg_eps;
  oc_real dynamic_reg_delta;
  int refine_max_iter;
  oc_real refine_tol;
  oc_real max_step_fraction;
  int equilibrate;
  int equilibrate_max_iter;
  oc_real equilibrate_min_scale;
  oc_real equilibrate_max_scale;
  int warm_start;
  int verbose;
  int record_timings;
  int validate_inputs;
  int linear_solver; /* 0 Eigen, 1 Qdldl, 2 RegularizedLdlt */
} oc_settings;

void oc_default_settings(oc_settings* s);

typedef struct oc_solver oc_solver;

oc_solver* oc_create(void);
void oc_destroy(oc_solver* h);

/* CSC, int32 indices, P upper triangle.  Returns 1 on success, 0 on failure. */
int oc_setup(oc_solver* h, int n, int m, const int* Pp, const int* Pi, const oc_real* Px,
             const oc_real* q, const int* Ap, const int* Ai, const oc_real* Ax, const oc_real* b,
             int zero_dim, int nonneg_dim, int n_soc, const int* soc_dims, const oc_settings* st);

/* Any pointer group may be NULL (unchanged).  P given iff Px != NULL (Pp/Pi then required). */
int oc_update_data(oc_solver* h, const int* Pp, const int* Pi, const oc_real* Px, const oc_real* q,
                   const int* Ap, const int* Ai, const oc_real* Ax, const oc_real* b);
void oc_set_warm_start(oc_solver* h, const oc_real* x, const oc_real* s, const oc_real* z);
void oc_set_settings(oc_solver* h, const oc_settings* st);

/* Returns the Status code (0 Unsolved .. 5 NumericalError). */
int oc_solve(oc_solver* h);

/* info[0..4] = primal_residual, dual_residual, duality_gap, mu, objective. */
int oc_get_solution(const oc_solver* h, oc_real* x, oc_real* s, oc_real* z, oc_real* info,
                    int* iterations);

/* Per-iteration trace: rows of 8 doubles
 * (mu, norm_rx, norm_rz, pobj, dobj, alpha_aff, sigma, alpha).  Returns rows written. */
int oc_get_trace(const oc_solver* h, oc_real* out, int max_rows);

/* Symbolic statistics of the last setup: [N, nnz(K lower), nnz(L), factor flops]. */
void oc_get_kkt_stats(const oc_solver* h, oc_real* out4);

/* ---- per-kernel entry points (the same code the IPM uses) ---- */

/* SOC NT scaling of one block (dim d): dense column-major W and Hs (d*d each). */
void oc_soc_scaling(int d, const oc_real* s, const oc_real* z, oc_real* W, oc_real* Hs);
oc_real oc_soc_max_step(int d, const oc_real* x, const oc_real* dx, oc_real alpha_max);
void oc_soc_product(int d, const oc_real* u, const oc_real* v, oc_real* out);
void oc_soc_inverse_product(int d, const oc_real* u, const oc_real* w, oc_real* out);
/* Cone-set level ops on stacked m-vectors. */
void oc_cone_margins(int zero_dim, int nonneg_dim, int n_soc, const int* soc_dims,
                     const oc_real* x, oc_real* min_margin, oc_real* pos_sum);
oc_real oc_cone_max_step(int zero_dim, int nonneg_dim, int n_soc, const int* soc_dims,
                        const oc_real* x, const oc_real* dx, oc_real alpha_max);
/* NT scaling of the whole cone set at (s, z); writes lambda = W z, Hz = Hs*v,
 * Winv_v = W^-1 v for a probe vector v. */
void oc_cone_scaling_apply(int zero_dim, int nonneg_dim, int n_soc, const int* soc_dims,
                           const oc_real* s, const oc_real* z, const oc_real* v, oc_real* lambda,
                           oc_real* Hv, oc_real* Wv, oc_real* Winv_v);

/* Ruiz equilibration only: returns d (n), e (m), c. */
void oc_equilibrate(int n, int m, const int* Pp, const int* Pi, const oc_real* Px, const int* Ap,
                    const int* Ai, const oc_real* Ax, int zero_dim, int nonneg_dim, int n_soc,
                    const int* soc_dims, const oc_settings* st, oc_real* d, oc_real* e, oc_real* c);

/* One KKT solve on a freshly set-up system (no equilibration) with cone scaling at (s,z)
 * (NULL: identity scaling as after construction).  Returns the refined relative residual. */
oc_real oc_kkt_solve(int n, int m, const int* Pp, const int* Pi, const oc_real* Px, const int* Ap,
                    const int* Ai, const oc_real* Ax, int zero_dim, int nonneg_dim, int n_soc,
                    const int* soc_dims, const oc_real* s, const oc_real* z, const oc_settings* st,
                    const oc_real* rhs, oc_real* x_out, int* ok);

#ifdef __cplusplus
}
#endif

#endif
