// Device interior-point solver behind the C-ABI (one problem per handle).
#pragma once

#include <memory>
#include <vector>

#include "../../include/conicxx_c.h"
#include "common.hpp"
#include "ipm_kernels.hpp"
#include "kkt_device.hpp"

namespace cx {

struct HostCSC {
  int rows = 0, cols = 0;
  std::vector<int> p, i;
  std::vector<double> x;
  int nnz() const { return p.empty() ? 0 : p[cols]; }
};

class Solver {
 public:
  Solver();
  ~Solver();
  bool setup(int n, int m, const int* Pp, const int* Pi, const double* Px, const double* q,
             const int* Ap, const int* Ai, const double* Ax, const double* b, int zero_dim,
             int nonneg_dim, int nsoc, const int* soc_dims, const cx_settings& st);
  bool update_data(const int* Pp, const int* Pi, const double* Px, const double* q, const int* Ap,
                   const int* Ai, const double* Ax, const double* b);
  // update_data without any device or host allocation: values are staged in
  // the library's per-GPU static scratch array (at most kScratchDoubles values
  // per piece); calls that overlap on one GPU race on that array
  bool update_data_noalloc(const int* Pp, const int* Pi, const double* Px, const double* q, const int* Ap,
                           const int* Ai, const double* Ax, const double* b);
  static constexpr long kScratchDoubles = 1L << 20;
  void set_warm_start(const double* x, const double* s, const double* z);
  void set_settings(const cx_settings& st) { st_ = st; }
  const cx_settings& settings() const { return st_; }
  int solve();
  int get_solution(double* x, double* s, double* z, double* obj, cx_info* info) const;
  void dims(int* n, int* m) const { *n = n_; *m = m_; }
  void kkt_stats(double* out8) const;
  void profile(double* out13) const;
  int kernel_stats(double* out, int max_kinds) const;
  int backward(const double* wx, const double* ws, const double* wz, double* gP, double* gq,
               double* gA, double* gb);
  // One KKT factorization (NT scaling at (s, z), NULL = identity) and one
  // refined solve on the current data (kkt_system.cpp:139-150, 296-311)
  double kkt_factor_solve(const double* s, const double* z, const double* rhs, double* x, int refine_max_iter,
                          int* factorizations);

 private:
  // KKT (kkt_system.cpp)
  bool kkt_update_pa(bool p, bool a);
  bool kkt_update_scaling();
  bool factor_retry();
  bool try_factor(double ep, double ea);
  bool escalate(const double* rhs, double* x);
  void leave_fp32(double rel);  // FP64 factorizations for the rest of this solve()
  double kkt_solve(const double* rhs, double* x);
  // kkt_solve on two independent right-hand sides with their sweeps paired;
  // returns whether the second result is valid (see solver.hip)
  bool kkt_solve_pair(const double* rhs0, double* x0, const double* rhs1, double* x1);
  // GMRES on K with the FP32 factorization as right preconditioner, from x
  // with residual r_ (relative residual rel); returns the final relative residual
  double gmres_refine(const double* rhs, double* x, double rn, double rel);
  // IPM pieces (solver_impl.cpp)
  bool initial_point();
  void shift_cold(double* v);
  void ensure_interior_warm(double* v);
  bool constant_solve();
  void constant_rhs(double* rhs);
  bool constant_finish(const double* sol);
  void residuals();
  bool tau_elim(const double* xv, const double* zv, double dtau_rhs, double dkappa_rhs, double* dtau);
  bool affine_step();
  void affine_rhs(double* rhs);
  bool affine_finish(const double* sol);
  bool combined_step(double sigma, double mu);
  double step_length(const double* ds, const double* dz, double dtau, double dkappa);
  void finalize();
  double mu() const { return (dsz_ + tau_ * kappa_) / (double)(degree_ + 1); }
  // helpers
  std::vector<double> reduce(const RedBatch& b);
  void pmul(const double* x, double* y);
  void margins(const double* v, double* mn, double* ps);
  void zero_zero_block(double* s);
  int count_nonfinite(const double* a, int la, const double* b = nullptr, int lb = 0,
                      const double* c = nullptr, int lc = 0);

  hipStream_t stream_ = nullptr;
  int device_ = 0;  // the GPU current at construction (owner of stream_)
  cx_settings st_{};
  bool setup_done_ = false;
  int n_ = 0, m_ = 0, N_ = 0;
  int zero_dim_ = 0, nonneg_dim_ = 0, degree_ = 0;
  std::vector<int> soc_dims_, soc_off_;
  int P_nnz_ = 0;
  // pattern fingerprints (kkt_system.cpp:110-113)
  std::vector<int> Pp_ref_, Pi_ref_, Ap_ref_, Ai_ref_;
  // equilibration (host) + device copies
  std::vector<double> ed_, ee_;
  double ec_ = 1.0;
  DevBuf<double> d_ed_, d_ee_;
  // problem data on device (scaled)
  DevBuf<double> dPx_, dAx_, dq_, db_, dPx_new_, dAx_new_;
  DevBuf<int> dPp_, dPi_, dAp_, dAi_;
  DevBuf<int> pf_ptr_, pf_col_, pf_src_;  // full symmetric P (CSR)
  DevBuf<int> ar_ptr_, ar_col_, ar_src_;  // A by rows
  DevBuf<int> at_src_;                    // identity map for A' by columns
  // KKT values and slots
  DevBuf<double> Kx_;
  int nnzK_ = 0;
  DevBuf<int> pd_slot_, pd_src_, po_slot_, po_src_, a_slot_, a_src_;
  int npd_ = 0, npo_ = 0, na_ = 0;
  DevBuf<int> hs_slot_, hs_kind_, hs_blk_, hs_a_, hs_b_, hs_row_;
  int nhs_ = 0;
  DeviceKKT kkt_;
  bool factorized_ = false;
  double last_p_ = 0, last_a_ = 0;
  // cones
  ConeDev cone_{};
  DevBuf<int> c_soc_off_, c_soc_dim_, c_small_, c_large_;
  DevBuf<double> c_beta_, c_vv_, c_v_, c_w_, c_hsd_;
  DevBuf<int> c_ident_;
  DevBuf<double> identity_;
  // iterate and workspace
  DevBuf<double> x_, s_, z_, Px_, rx_, rz_, x1_, z1_, Px1_, dxa_, dza_, dsa_, dx_, dz_, ds_;
  DevBuf<double> wx_, ws_, wz_;
  DevBuf<double> rhs_, sol_, r_, t_, dsol_, wn_, wn2_, wm_, wm2_, dsk_;
  DevBuf<double> rhs2_, sol2_, r2_, t2_, dsol2_;  // second side of kkt_solve_pair
  DevBuf<double> red_partial_, red_out_;
  double* host_red_ = nullptr;
  double tau_ = 1.0, kappa_ = 1.0;
  double rtau_ = 0, dqx_ = 0, dbz_ = 0, dsz_ = 0, dxPx_ = 0, nrx2_ = 0, nrz2_ = 0;
  double dtaua_ = 0, dkappaa_ = 0, dtau_ = 0, dkappa_ = 0;
  bool have_warm_ = false;
  // backward pass
  bool bw_valid_ = false;  // the factorization in memory belongs to the last Solved solve
  bool bw_factored_ = false;  // the adjoint (tangent-corrected) matrix is factored
  DevBuf<double> bw_tcorr_, bw_tdir_;
  DevBuf<double> bw_rhs_, bw_sol_, bw_us_, bw_ntz_, bw_gP_, bw_gA_, bw_gq_, bw_gb_;
  // solution
  int status_ = CX_UNSOLVED, iterations_ = 0;
  std::vector<double> xo_, so_, zo_;
  double objective_ = 0, pres_ = 0, dres_ = 0, gap_ = 0, mu_out_ = 0;
  double setup_time_ = 0, solve_time_ = 0;
  // profile
  double prof_fact_ms_ = 0, prof_solve_ms_ = 0, prof_total_ms_ = 0;
  int prof_nfact_ = 0, prof_nsolve_ = 0;
  // FP32 factorization with FP64 refinement (factor_precision); falls back to
  // FP64 for the rest of a solve when refinement misses refine_tol
  bool fp32_ = false;
  int prof_nfact32_ = 0, prof_fp64_fallback_ = 0, prof_fallback_iter_ = -1;
  double prof_max_rel_ = 0, prof_fallback_rel_ = 0;
  static constexpr int GM_RESTART = 20;     // Arnoldi steps per GMRES cycle
  static constexpr int GM_MAX_SOLVES = 20;  // preconditioner applications per KKT solve
  static constexpr double FP32_ACCEPT_REL = 1e-9;
  DevBuf<double> gm_v_, gm_z_, gm_part_, gm_h_;
  double* gm_host_ = nullptr;  // pinned: two Gram-Schmidt passes' coefficients
};

}  // namespace cx
