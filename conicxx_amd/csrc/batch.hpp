// Batched solver: many independent conic QPs, each solved start to finish by
// one workgroup of a persistent kernel, with every per-problem datum on the
// device.  New in this design (the reference solves one problem per Solver
// object, src/conicxx/solver_impl.cpp:367-466); each problem follows exactly
// the same method as the single-problem path.
//
// Per problem the KKT system [P+eI, A'; A, -(Hs+eI)] is factored as a dense
// LDL^T in the order "variables first, then constraint rows", the order in
// which both pivot blocks stay definite (positive, then negative) and no
// pivot divides by a small (Hs + eI) eigenvalue.  Bound rows (nonnegative
// rows with a single entry, such as box constraints) are diagonal and are
// folded into the P block first (a^2 / (Hs_ii + e) on one diagonal entry, a
// harmless scaling); every other row is kept.  M, of order n + kept rows, is
// held packed (LDS when it fits, else global memory) and factored by the
// workgroup; solves are refined against the full K.
#pragma once

#include <vector>

#include "../../include/conicxx_c.h"
#include "common.hpp"

namespace cx {

struct BatchDesc {
  int n, m, nz, nn, nsoc, nk, pnnz, nr;  // nk kept rows, nr = n + nk
  int xo, co, so, ko;  // bases of n-vectors, m-vectors, SOC tables, kept-row lists
  int pfp, acp, arp;    // bases of the per-problem row/column pointer arrays
  long wofs;            // base of the per-problem workspace
  long mofs;            // base of M in global memory (problems whose M does not fit in LDS)
  double ec;            // cost scale
};

struct BatchTuning {
  int lds_doubles = 0;  // packed M up to this size is kept in LDS, larger in global memory
  int grid = 0;         // persistent workgroups
};

class BatchSolver {
 public:
  BatchSolver();
  ~BatchSolver();
  BatchSolver(const BatchSolver&) = delete;
  BatchSolver& operator=(const BatchSolver&) = delete;

  bool setup(int nprob, const int* n, const int* m, const int* const* Pp, const int* const* Pi,
             const double* const* Px, const double* const* q, const int* const* Ap,
             const int* const* Ai, const double* const* Ax, const double* const* b, const int* zero_dim,
             const int* nonneg_dim, const int* nsoc, const int* const* soc_dims, const cx_settings& st);
  // new q and/or b values for every problem (NULL = unchanged)
  bool update_data(const double* const* q, const double* const* b);
  // new P, q, A and/or b values (NULL = unchanged); P and A values must come
  // with their index arrays, which must equal setup's pattern
  bool update_data(const int* const* Pp, const int* const* Pi, const double* const* Px,
                   const double* const* q, const int* const* Ap, const int* const* Ai,
                   const double* const* Ax, const double* const* b);
  int solve(int* status_out);  // returns the number of Solved problems, -1 on failure
  int get_solution(int k, double* x, double* s, double* z, double* obj, cx_info* info);
  void set_settings(const cx_settings& st);
  int autotune(int reps);
  bool set_tuning(int lds_doubles, int grid);  // pins a launch configuration
  // differentiation: when on, every Solved problem's forward pass also factors
  // its adjoint KKT matrix; backward() is then one KKT solve per problem
  void set_differentiable(bool on);
  int backward(const double* const* wx, const double* const* ws, const double* const* wz,
               double* const* gP, double* const* gq, double* const* gA, double* const* gb);
  double last_backward_ms() const { return last_bw_ms_; }
  void tuning(double* out4) const;
  void stats(int k, double* out13);  // work counters of problem k (cx_batch_get_stats)
  double last_solve_ms() const { return last_ms_; }

 private:
  void launch_all();
  void build_args(void* args);
  void fill_diff_args(void* args);
  void fetch();

  hipStream_t st_ = nullptr;
  cx_settings set_{};
  bool ready_ = false;
  int np_ = 0;
  std::vector<int> n_, m_, nz_, nn_;
  std::vector<std::vector<int>> socd_;
  std::vector<BatchDesc> desc_h_;
  std::vector<double> ed_h_, ee_h_;
  long sum_n_ = 0, sum_m_ = 0;
  // device structure
  DevBuf<BatchDesc> desc_;
  DevBuf<int> order_, counter_, xprob_;
  DevBuf<int> pf_ptr_, pf_col_, pf_src_, ac_ptr_, ac_row_, ar_ptr_, ar_col_, ar_src_;
  DevBuf<int> soc_off_, soc_dim_;
  DevBuf<int> rk_, kr_;
  DevBuf<double> Pv_, Av_, q_, b_, qraw_, braw_, ed_, ee_, W_, Mg_;
  DevBuf<double> beta_, vv_, wxs_, wss_, wzs_, xo_, so_, zo_, info_;
  DevBuf<int> ident_, status_, iters_, have_warm_;
  double* stage_q_ = nullptr;  // pinned staging for update_data
  double* stage_b_ = nullptr;
  // P / A value updates: setup's patterns (host, for the pattern check), the
  // per-entry equilibration factors (c d_r d_c for P, e_r d_c for A), raw
  // values staged in pinned memory and uploaded + scaled before the next solve
  std::vector<std::vector<int>> Pp_h_, Pi_h_, Ap_h_, Ai_h_;
  std::vector<double> Praw_h_, Araw_h_;  // setup's values, unscaled
  DevBuf<double> sP_, sA_, Praw_, Araw_;
  double* stage_P_ = nullptr;
  double* stage_A_ = nullptr;
  bool P_dirty_ = false, A_dirty_ = false;
  void upload_values();
  int max_packed_ = 0;
  long max_lds_need_ = 0;
  BatchTuning tune_{};
  bool tuned_ = false;
  // graph of one timestep: staged q/b upload, rescale, counter reset, solve
  hipGraph_t graph_ = nullptr;
  hipGraphExec_t gexec_ = nullptr;
  // host copies of the results
  bool fetched_ = false;
  std::vector<double> hx_, hs_, hz_, hinfo_;
  std::vector<int> hstatus_, hiters_;
  double last_ms_ = 0, last_bw_ms_ = 0;
  // differentiation state
  bool diff_ = false;
  std::vector<long> pv_base_, av_base_, av_len_;
  long nnzP_total_ = 0, nnzA_total_ = 0;
  size_t nsoc_total_ = 0;
  DevBuf<double> tcorr_, tdir_, xh_, sh_, zh_, diff_a_, gwx_, gws_, gwz_, gP_, gA_, gq_, gb_;
  DevBuf<int> diff_ok_;
  double clock_khz_ = 100000;
  hipEvent_t ev0_ = nullptr, ev1_ = nullptr;
};

}  // namespace cx
