// Device supernodal (multifrontal) LDL^T of the quasi-definite KKT matrix.
//
// Numeric counterpart of the reference's LDL^T backends: QdldlLdlt::factorize/
// solve (src/conicxx/kkt/qdldl_ldlt.cpp:58-75) and RegularizedLdlt::factorize
// (src/conicxx/kkt/regularized_ldlt.cpp:142-254).  The symbolic plan comes
// from symbolic.cpp; every launch of a factorization or solve follows a static
// schedule built once per setup().
#pragma once

#include <vector>

#include "common.hpp"
#include "symbolic.hpp"

namespace cx {

struct FactorStats {
  double min_abs_d = 0;  // NaN if any pivot is NaN
  int zero_pivot = 0;    // exact zero pivot met (checked pivot mode)
  int nreg = 0;          // pivots corrected inline (regularizing mode)
  int stopped = 0;       // checked mode: a pivot |d| <= eps ended the factorization early
};

struct AsmTask { int s, j0, j1, seg_beg, seg_end; };
struct AsmSeg { int child, i0, i1; };
// rows [r0, r1) of sub-panel [a, a+w) of front s; the task with r0 == a + w
// also stores the factored diagonal block, in slot blk of the staging buffer
struct PanelTask { int s, a, w, r0, r1, blk; };
struct TrailTask { int s, ti, tj, k0, k1, rlo, chi; };
struct SolveTask { int s, blk; };
// Interleaved right-hand sides of one sweep launch: ticket t runs task t / nctx
// of context t % nctx, each context with its own work vector, hand-off
// buffer, update vectors, block flags and done counters at these strides.
struct SweepCtx { int nctx; long vstride, hstride, ustride, fstride, dstride; };

enum LaunchKind { L_ASSEMBLE = 0, L_SMALL = 1, L_PANEL = 3, L_TRAIL = 4, L_TRAIL64 = 5 };
struct Launch { int kind, beg, end; };

// Pivot modes.
enum PivotMode { PIVOT_CHECK = 0, PIVOT_REGULARIZE = 1 };

class DeviceKKT {
 public:
  DeviceKKT() = default;
  DeviceKKT(const DeviceKKT&) = delete;
  DeviceKKT& operator=(const DeviceKKT&) = delete;

  // Symbolic analysis + device plan.  K: lower CSC pattern (diagonal present
  // for every column).  nx: columns < nx expect positive pivots, others negative.
  void setup(const LowerPattern& K, int nx, hipStream_t stream);

  // Numeric factorization of the K lower values Kx (device, CSC order) with the
  // diagonal bumped by +extra_p (rows < nx) and -extra_a (rows >= nx).
  // prec 1: the fronts are factored in FP32 (separate FP32 front storage,
  // FP32 MFMA trailing updates); pivots and the solve sweeps stay FP64.
  void factorize_async(const double* Kx, double extra_p, double extra_a, int mode, double eps,
                       double delta, int prec = 0);
  int factor_precision() const { return prec_; }
  FactorStats factor_stats();  // synchronizes the stream

  // x = K^-1 rhs with the current factorization (device vectors of length N).
  void solve_async(const double* rhs, double* x);
  // x0 = K^-1 rhs0 and x1 = K^-1 rhs1 in one launch per sweep (the two
  // contexts' tasks interleaved); each result equals solve_async's
  void solve2_async(const double* rhs0, double* x0, const double* rhs1, double* x1);
  // y = K x, K symmetric from its lower values Kx.
  void spmv_async(const double* Kx, const double* x, double* y);
  // nonzero if a bounded wait in a solve sweep timed out (synchronizes)
  int take_solve_error();

  const SupernodalPlan& plan() const { return P_; }
  int N() const { return P_.N; }
  hipStream_t stream() const { return st_; }

  // Per-kernel-class device timing with pooled HIP events recorded on this
  // stream around each launch of every Nth factorization and every Nth solve
  // (CONICXX_KERNEL_TIMING=N, N >= 1; every launch and its work are counted);
  // events are resolved only where the solver synchronizes anyway.
  enum TimedKind { T_TRAIL = 0, T_PANEL, T_ASM, T_SMALL, T_FWD, T_BWD, T_NKINDS };
  struct KindStats {
    double ms = 0;    // device time: the timed launches' total scaled to all launches
    double work = 0;  // algorithmic flops (factorization kinds) or bytes (sweeps), all launches
    long launches = 0;
  };
  bool timing() const { return timing_; }
  void collect_timing();
  void reset_timing();
  KindStats kind_stats(int k) const;
  double sweep_bytes() const { return sweep_bytes_; }

 private:
  void build_schedule();
  void sweeps(int nctx, const double* const* rhs, double* const* x);
  template <typename T>
  void launch_factor(T* F, const double* Kx, double extra_p, double extra_a, int mode, double eps,
                     double delta);

  hipStream_t st_ = nullptr;
  SupernodalPlan P_;
  int nx_ = 0;
  // host copies of the schedule
  std::vector<Launch> launches_;
  std::vector<AsmTask> asm_h_;
  std::vector<AsmSeg> seg_h_;
  std::vector<int> small_h_;
  std::vector<PanelTask> panel_h_;
  std::vector<PanelTask> dslot_h_;  // one per sub-panel: where its diagonal block goes
  std::vector<TrailTask> trail_h_;
  std::vector<double> trail_launch_flops_;
  // device plan
  DevBuf<long> fofs_;
  DevBuf<int> nr_, nc_, first_, rowptr_, rows_, child_ptr_, child_;
  DevBuf<double> sgn_;
  DevBuf<int> kcolptr_, ksrc_;
  DevBuf<long> kdst_;
  DevBuf<signed char> kdiag_;
  DevBuf<int> rmap_ofs_, rmap_;
  DevBuf<AsmTask> asm_;
  DevBuf<AsmSeg> seg_;
  DevBuf<int> small_;
  DevBuf<PanelTask> panel_, dslot_;
  DevBuf<double> Fd_;  // factored diagonal blocks (NB x NB per sub-panel), stored into F after the sweep of launches
  DevBuf<TrailTask> trail_;
  DevBuf<double> F_;   // lower-packed fronts: FP64 factor, or the FP32 factor in its first half
  int prec_ = 0;       // precision of the current factorization
  DevBuf<double> D_;   // pivots, permuted order
  DevBuf<int> fstat_;  // [0] zero pivot, [1] nreg
  DevBuf<double> dstat_;
  // solve plan
  DevBuf<int> perm_;
  DevBuf<SolveTask> fwd_, bwd_;
  int nfwd_ = 0, nbwd_ = 0;
  DevBuf<int> gptr_, gidx_;
  DevBuf<long> uofs_;
  DevBuf<int> flag_ofs_, nblk_, nblkb_, parent_;
  DevBuf<int> flags_, done_, tickets_, serr_;
  DevBuf<double> xp_, U_;
  DevBuf<double> hand_;  // hand-off buffers of the forward [0, N) and backward [N, 2N) sweeps
  int gen_ = 0;
  // K full symmetric CSR for spmv
  DevBuf<int> kfptr_, kfcol_, kfsrc_;
  // timing
  void tbegin();
  void tend(int kind, double work);
  bool timing_ = false;
  int timing_every_ = 1;
  bool sample_ = false;               // the current factorization / solve is timed
  long fact_seq_ = 0, sweep_seq_ = 0;  // since the last reset_timing()
  std::vector<hipEvent_t> evpool_;
  size_t evused_ = 0;
  std::vector<int> evpending_;  // kind of each timed launch
  struct KindRaw {
    double ms_timed = 0, work = 0;
    long timed = 0, all = 0;
  };
  KindRaw kraw_[T_NKINDS];
  // launches and work of the factorization in flight, added to kraw_ by
  // factor_stats() only if it ran to completion (a checked factorization that
  // stops at a failing pivot leaves its remaining launches returning at once:
  // neither its time nor its work is kept)
  KindRaw fpend_[T_NKINDS];
  bool in_factor_ = false;
  size_t fact_ev0_ = 0;  // first timed launch of that factorization in evpending_
  double sweep_bytes_ = 0;
};

}  // namespace cx
