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

/* Field-for-field the reference Settings aggregate (include/conicxx/settings.h:77-118). */
typedef struct oc_settings {
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
int oc_setup(oc_solver* h, int n, int m, const int* Pp, const int* Pi, const double* Px,
             const double* q, const int* Ap, const int* Ai, const double* Ax, const double* b,
             int zero_dim, int nonneg_dim, int n_soc, const int* soc_dims, const oc_settings* st);

/* Any pointer group may be NULL (unchanged).  P given iff Px != NULL (Pp/Pi then required). */
int oc_update_data(oc_solver* h, const int* Pp, const int* Pi, const double* Px, const double* q,
                   const int* Ap, const int* Ai, const double* Ax, const double* b);
void oc_set_warm_start(oc_solver* h, const double* x, const double* s, const double* z);
void oc_set_settings(oc_solver* h, const oc_settings* st);

/* Returns the Status code (0 Unsolved .. 5 NumericalError). */
int oc_solve(oc_solver* h);

/* info[0..4] = primal_residual, dual_residual, duality_gap, mu, objective. */
int oc_get_solution(const oc_solver* h, double* x, double* s, double* z, double* info,
                    int* iterations);

/* Per-iteration trace: rows of 8 doubles
 * (mu, norm_rx, norm_rz, pobj, dobj, alpha_aff, sigma, alpha).  Returns rows written. */
int oc_get_trace(const oc_solver* h, double* out, int max_rows);

/* Symbolic statistics of the last setup: [N, nnz(K lower), nnz(L), factor flops]. */
void oc_get_kkt_stats(const oc_solver* h, double* out4);

/* ---- per-kernel entry points (the same code the IPM uses) ---- */

/* SOC NT scaling of one block (dim d): dense column-major W and Hs (d*d each). */
void oc_soc_scaling(int d, const double* s, const double* z, double* W, double* Hs);
double oc_soc_max_step(int d, const double* x, const double* dx, double alpha_max);
void oc_soc_product(int d, const double* u, const double* v, double* out);
void oc_soc_inverse_product(int d, const double* u, const double* w, double* out);
/* Cone-set level ops on stacked m-vectors. */
void oc_cone_margins(int zero_dim, int nonneg_dim, int n_soc, const int* soc_dims,
                     const double* x, double* min_margin, double* pos_sum);
double oc_cone_max_step(int zero_dim, int nonneg_dim, int n_soc, const int* soc_dims,
                        const double* x, const double* dx, double alpha_max);
/* NT scaling of the whole cone set at (s, z); writes lambda = W z, Hz = Hs*v,
 * Winv_v = W^-1 v for a probe vector v. */
void oc_cone_scaling_apply(int zero_dim, int nonneg_dim, int n_soc, const int* soc_dims,
                           const double* s, const double* z, const double* v, double* lambda,
                           double* Hv, double* Wv, double* Winv_v);

/* Ruiz equilibration only: returns d (n), e (m), c. */
void oc_equilibrate(int n, int m, const int* Pp, const int* Pi, const double* Px, const int* Ap,
                    const int* Ai, const double* Ax, int zero_dim, int nonneg_dim, int n_soc,
                    const int* soc_dims, const oc_settings* st, double* d, double* e, double* c);

/* One KKT solve on a freshly set-up system (no equilibration) with cone scaling at (s,z)
 * (NULL: identity scaling as after construction).  Returns the refined relative residual. */
double oc_kkt_solve(int n, int m, const int* Pp, const int* Pi, const double* Px, const int* Ap,
                    const int* Ai, const double* Ax, int zero_dim, int nonneg_dim, int n_soc,
                    const int* soc_dims, const double* s, const double* z, const oc_settings* st,
                    const double* rhs, double* x_out, int* ok);

#ifdef __cplusplus
}
#endif

#endif
