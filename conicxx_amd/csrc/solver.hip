// Device interior-point solver: host orchestration of the homogeneous
// self-dual embedding IPM (src/conicxx/solver_impl.cpp) over device vectors,
// with the KKT system of src/conicxx/kkt/kkt_system.cpp (slot layout, static
// regularization, x10 retry ladder, x1000 escalation, iterative refinement)
// factorized by the device multifrontal LDL^T.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <numeric>

#include "solver.hpp"
#include "equilibrate.hpp"

namespace cx {

namespace {

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

bool valid_csc(int rows, int cols, const int* p, const int* i) {
  if (!p || p[0] != 0) return false;
  for (int c = 0; c < cols; ++c) {
    if (p[c + 1] < p[c]) return false;
    for (int k = p[c]; k < p[c + 1]; ++k) {
      if (!i || i[k] < 0 || i[k] >= rows) return false;
      if (k > p[c] && i[k] <= i[k - 1]) return false;  // sorted, no duplicates
    }
  }
  return true;
}

RedOp dot_op(const double* a, const double* b, int len) { return RedOp{RED_DOT, len, a, b, nullptr, nullptr, 0.0}; }
RedOp nonfin_op(const double* a, int len) { return RedOp{RED_NONFIN, len, a, nullptr, nullptr, nullptr, 0.0}; }

}  // namespace

// Staging for update_data_noalloc: one array per GPU for the whole process,
// shared by every handle on that GPU.  Host code never uses the name itself
// (a host-side stand-in): it takes the device address from scratch_address().
static __device__ __align__(256) double scratch[Solver::kScratchDoubles];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch[kMaxDevices] = {};

// Device address of `scratch` on GPU dev, looked up once per device with that
// device current; nullptr if the lookup fails.
double* scratch_address(int dev) {
  if (dev < 0 || dev >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[dev]) {
    int cur = -1;
    if (hipGetDevice(&cur) != hipSuccess) return nullptr;
    if (cur != dev && hipSetDevice(dev) != hipSuccess) return nullptr;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != dev) (void)hipSetDevice(cur);
    if (e != hipSuccess || !p) {
      std::fprintf(stderr, "[conicxx_amd] hipGetSymbolAddress(scratch) failed on device %d: %s\n", dev,
                   hipGetErrorString(e));
      return nullptr;
    }
    g_scratch[dev] = static_cast<double*>(p);
  }
  return g_scratch[dev];
}
}  // namespace

Solver::Solver() {
  CX_HIP(hipGetDevice(&device_));
  CX_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
  CX_HIP(hipHostMalloc((void**)&host_red_, 64 * sizeof(double), hipHostMallocDefault));
  red_partial_.alloc(RED_MAX_OPS * RED_BLOCKS * 3);
  red_out_.alloc(64);
  cx_default_settings(&st_);
}

Solver::~Solver() {
  if (host_red_) (void)hipHostFree(host_red_);
  if (gm_host_) (void)hipHostFree(gm_host_);
  if (stream_) (void)hipStreamDestroy(stream_);
}

std::vector<double> Solver::reduce(const RedBatch& b) {
  launch_reduce(b, red_partial_.p, red_out_.p, stream_);
  CX_HIP(hipMemcpyAsync(host_red_, red_out_.p, b.nops * sizeof(double), hipMemcpyDeviceToHost, stream_));
  CX_HIP(hipStreamSynchronize(stream_));
  return std::vector<double>(host_red_, host_red_ + b.nops);
}

void Solver::pmul(const double* x, double* y) {
  launch_csr_mv(n_, pf_ptr_.p, pf_col_.p, pf_src_.p, dPx_.p, x, y, stream_);
}

// ---------------------------------------------------------------------------
// setup: solver_impl.cpp:8-55, equilibration.cpp:8-139, kkt_system.cpp:22-117
// ---------------------------------------------------------------------------
bool Solver::setup(int n, int m, const int* Pp, const int* Pi, const double* Px, const double* q,
                   const int* Ap, const int* Ai, const double* Ax, const double* b, int zero_dim,
                   int nonneg_dim, int nsoc, const int* soc_dims, const cx_settings& st) {
  const double t0 = now_s();
  setup_done_ = false;
  bw_valid_ = false;
  if (zero_dim < 0 || nonneg_dim < 0 || n < 0 || m < 0 || nsoc < 0) return false;
  long total = (long)zero_dim + nonneg_dim;
  for (int k = 0; k < nsoc; ++k) {
    if (soc_dims[k] < 1) return false;
    total += soc_dims[k];
  }
  if (total != m) return false;
  if (!valid_csc(n, n, Pp, Pi) || !valid_csc(m, n, Ap, Ai)) return false;
  if ((n > 0 && !q) || (m > 0 && !b)) return false;
  st_ = st;
  n_ = n, m_ = m, N_ = n + m;
  zero_dim_ = zero_dim, nonneg_dim_ = nonneg_dim;
  soc_dims_.assign(soc_dims, soc_dims + nsoc);
  soc_off_.resize(nsoc);
  degree_ = nonneg_dim;
  {
    int off = zero_dim + nonneg_dim;
    for (int k = 0; k < nsoc; ++k) soc_off_[k] = off, off += soc_dims_[k], degree_ += 1;
  }
  const int nP = Pp[n], nA = Ap[n];
  P_nnz_ = nP;
  Pp_ref_.assign(Pp, Pp + n + 1);
  Pi_ref_.assign(Pi, Pi + nP);
  Ap_ref_.assign(Ap, Ap + n + 1);
  Ai_ref_.assign(Ai, Ai + nA);
  std::vector<double> Pv(Px, Px + nP), Av(Ax, Ax + nA), qv(q, q + n), bv(b, b + m);

  // --- Ruiz equilibration (equilibration.cpp:8-127) ---
  ruiz_equilibrate(n, m, Pp, Pi, Pv, Ap, Ai, Av, soc_off_, soc_dims_, st_, ed_, ee_, ec_);
  if (st_.equilibrate && n > 0) {
    for (int i = 0; i < n; ++i) qv[i] *= ec_ * ed_[i];
    for (int i = 0; i < m; ++i) bv[i] *= ee_[i];
  }
  d_ed_.upload(ed_, stream_);
  d_ee_.upload(ee_, stream_);

  // --- device problem data ---
  dPp_.upload(Pp_ref_, stream_);
  dPi_.upload(Pi_ref_, stream_);
  dAp_.upload(Ap_ref_, stream_);
  dAi_.upload(Ai_ref_, stream_);
  dPx_.upload(Pv, stream_);
  dAx_.upload(Av, stream_);
  dPx_new_.alloc(std::max(1, nP));
  dAx_new_.alloc(std::max(1, nA));
  dq_.upload(qv, stream_);
  db_.upload(bv, stream_);
  {  // full symmetric P by rows (selfadjointView<Upper>)
    std::vector<int> ptr(n + 1, 0);
    for (int c = 0; c < n; ++c)
      for (int k = Pp[c]; k < Pp[c + 1]; ++k) {
        int r = Pi[k];
        if (r > c) continue;
        ptr[r + 1]++;
        if (r < c) ptr[c + 1]++;
      }
    for (int i = 0; i < n; ++i) ptr[i + 1] += ptr[i];
    std::vector<int> col(ptr[n]), src(ptr[n]), fill(ptr.begin(), ptr.end() - 1);
    for (int c = 0; c < n; ++c)
      for (int k = Pp[c]; k < Pp[c + 1]; ++k) {
        int r = Pi[k];
        if (r > c) continue;
        col[fill[r]] = c, src[fill[r]++] = k;
        if (r < c) col[fill[c]] = r, src[fill[c]++] = k;
      }
    pf_ptr_.upload(ptr, stream_);
    pf_col_.upload(col.empty() ? std::vector<int>{0} : col, stream_);
    pf_src_.upload(src.empty() ? std::vector<int>{0} : src, stream_);
  }
  {  // A by rows, A' by columns
    std::vector<int> ptr(m + 1, 0);
    for (int k = 0; k < nA; ++k) ptr[Ai[k] + 1]++;
    for (int i = 0; i < m; ++i) ptr[i + 1] += ptr[i];
    std::vector<int> col(nA), src(nA), fill(ptr.begin(), ptr.end() - 1);
    for (int c = 0; c < n; ++c)
      for (int k = Ap[c]; k < Ap[c + 1]; ++k) col[fill[Ai[k]]] = c, src[fill[Ai[k]]++] = k;
    ar_ptr_.upload(ptr, stream_);
    ar_col_.upload(col.empty() ? std::vector<int>{0} : col, stream_);
    ar_src_.upload(src.empty() ? std::vector<int>{0} : src, stream_);
    std::vector<int> iota(std::max(1, nA));
    std::iota(iota.begin(), iota.end(), 0);
    at_src_.upload(iota, stream_);
  }

  // --- cones ---
  {
    std::vector<int> small, large;
    for (int k = 0; k < nsoc; ++k) (soc_dims_[k] <= 32 ? small : large).push_back(k);
    c_soc_off_.upload(soc_off_.empty() ? std::vector<int>{0} : soc_off_, stream_);
    c_soc_dim_.upload(soc_dims_.empty() ? std::vector<int>{0} : soc_dims_, stream_);
    c_small_.upload(small.empty() ? std::vector<int>{0} : small, stream_);
    c_large_.upload(large.empty() ? std::vector<int>{0} : large, stream_);
    c_beta_.alloc(std::max(1, nsoc));
    c_vv_.alloc(std::max(1, nsoc));
    c_ident_.alloc(std::max(1, nsoc));
    c_v_.alloc(std::max(1, m));
    c_w_.alloc(std::max(1, m));
    c_hsd_.alloc(std::max(1, m));
    cone_.zero_off = 0;
    cone_.zero_dim = zero_dim;
    cone_.nn_off = zero_dim;
    cone_.nn_dim = nonneg_dim;
    cone_.nsoc = nsoc;
    cone_.soc_off = c_soc_off_.p;
    cone_.soc_dim = c_soc_dim_.p;
    cone_.small_list = c_small_.p;
    cone_.nsmall = (int)small.size();
    cone_.large_list = c_large_.p;
    cone_.nlarge = (int)large.size();
    cone_.beta = c_beta_.p;
    cone_.vv = c_vv_.p;
    cone_.ident = c_ident_.p;
    cone_.v = c_v_.p;
    cone_.w = c_w_.p;
    cone_.hsd = c_hsd_.p;
    std::vector<double> e(std::max(1, m), 0.0);
    for (int i = 0; i < nonneg_dim; ++i) e[zero_dim + i] = 1.0;
    for (int k = 0; k < nsoc; ++k) e[soc_off_[k]] = 1.0;
    identity_.upload(e, stream_);
  }

  // --- KKT slots and pattern (kkt_system.cpp:54-107, sparsity_map.cpp:14-56) ---
  struct Key { int r, c, id; };
  std::vector<Key> keys;
  std::vector<int> pd_slot(n), pd_src(n, -1), po_slot, po_src, a_slot, a_src;
  std::vector<int> hs_slot, hs_kind, hs_blk, hs_a, hs_b, hs_row;
  auto add = [&](int r, int c) {
    keys.push_back({r, c, (int)keys.size()});
    return (int)keys.size() - 1;
  };
  for (int i = 0; i < n; ++i) {
    pd_slot[i] = add(i, i);
    for (int k = Pp[i]; k < Pp[i + 1]; ++k)
      if (Pi[k] == i) pd_src[i] = k;
  }
  for (int c = 0; c < n; ++c)
    for (int k = Pp[c]; k < Pp[c + 1]; ++k) {
      int r = Pi[k];
      if (r >= c) continue;
      po_slot.push_back(add(c, r));
      po_src.push_back(k);
    }
  for (int c = 0; c < n; ++c)
    for (int k = Ap[c]; k < Ap[c + 1]; ++k) {
      a_slot.push_back(add(n + Ai[k], c));
      a_src.push_back(k);
    }
  auto add_hs = [&](int kind, int blk, int off, int a, int b2) {
    hs_slot.push_back(add(n + off + a, n + off + b2));
    hs_kind.push_back(kind);
    hs_blk.push_back(blk);
    hs_a.push_back(a);
    hs_b.push_back(b2);
    hs_row.push_back(off + a);
  };
  for (int a = 0; a < zero_dim; ++a) add_hs(0, 0, 0, a, a);
  for (int a = 0; a < nonneg_dim; ++a) add_hs(1, 0, zero_dim, a, a);
  for (int k = 0; k < nsoc; ++k)
    for (int a = 0; a < soc_dims_[k]; ++a)
      for (int b2 = 0; b2 <= a; ++b2) add_hs(2, k, soc_off_[k], a, b2);
  std::vector<Key> sorted(keys);
  std::sort(sorted.begin(), sorted.end(), [](const Key& x, const Key& y) {
    return x.c != y.c ? x.c < y.c : (x.r != y.r ? x.r < y.r : x.id < y.id);
  });
  LowerPattern K;
  K.N = N_;
  K.colptr.assign(N_ + 1, 0);
  std::vector<int> slot_off(keys.size());
  for (size_t t = 0; t < sorted.size(); ++t) {
    bool fresh = t == 0 || sorted[t].r != sorted[t - 1].r || sorted[t].c != sorted[t - 1].c;
    if (fresh) {
      K.rowind.push_back(sorted[t].r);
      K.colptr[sorted[t].c + 1]++;
    }
    slot_off[sorted[t].id] = (int)K.rowind.size() - 1;
  }
  for (int c = 0; c < N_; ++c) K.colptr[c + 1] += K.colptr[c];
  nnzK_ = (int)K.rowind.size();
  auto remap = [&](std::vector<int>& v) {
    for (int& s : v) s = slot_off[s];
  };
  remap(pd_slot), remap(po_slot), remap(a_slot), remap(hs_slot);
  npd_ = n, npo_ = (int)po_slot.size(), na_ = (int)a_slot.size(), nhs_ = (int)hs_slot.size();
  auto up = [&](DevBuf<int>& d, const std::vector<int>& v) { d.upload(v.empty() ? std::vector<int>{0} : v, stream_); };
  up(pd_slot_, pd_slot), up(pd_src_, pd_src), up(po_slot_, po_slot), up(po_src_, po_src);
  up(a_slot_, a_slot), up(a_src_, a_src);
  up(hs_slot_, hs_slot), up(hs_kind_, hs_kind), up(hs_blk_, hs_blk), up(hs_a_, hs_a), up(hs_b_, hs_b);
  up(hs_row_, hs_row);
  Kx_.alloc(std::max(1, nnzK_));
  Kx_.zero(stream_);
  kkt_.setup(K, n, stream_);
  factorized_ = false;
  last_p_ = last_a_ = 0;

  // --- workspace ---
  for (DevBuf<double>* v : {&x_, &Px_, &rx_, &x1_, &Px1_, &dxa_, &dx_, &wx_, &wn_, &wn2_})
    v->alloc(std::max(1, n)), v->zero(stream_);
  for (DevBuf<double>* v : {&s_, &z_, &rz_, &z1_, &dza_, &dsa_, &dz_, &ds_, &ws_, &wz_, &wm_, &wm2_, &dsk_})
    v->alloc(std::max(1, m)), v->zero(stream_);
  for (DevBuf<double>* v : {&rhs_, &sol_, &r_, &t_, &dsol_, &rhs2_, &sol2_, &r2_, &t2_, &dsol2_})
    v->alloc(std::max(1, N_)), v->zero(stream_);

  // updateData(P, A) then updateScalingAndFactorize with the constructor scaling
  if (!kkt_update_pa(true, true)) return false;
  launch_cone_scaling(cone_, identity_.p, identity_.p, stream_);
  if (!kkt_update_scaling()) return false;
  have_warm_ = false;
  status_ = CX_UNSOLVED;
  iterations_ = 0;
  xo_.assign(n, 0.0), so_.assign(m, 0.0), zo_.assign(m, 0.0);
  objective_ = pres_ = dres_ = gap_ = mu_out_ = 0;
  setup_done_ = true;
  setup_time_ = now_s() - t0;
  return true;
}

// KktSystem::updateData value writes (kkt_system.cpp:119-137), from the staged or
// current device values, followed by factorizeWithRetry.
bool Solver::kkt_update_pa(bool p, bool a) {
  const double* Pv = dPx_.p;
  const double* Av = dAx_.p;
  if (p) {
    launch_scatter(npd_, pd_slot_.p, pd_src_.p, Pv, st_.static_reg_P, Kx_.p, stream_);
    launch_scatter(npo_, po_slot_.p, po_src_.p, Pv, 0.0, Kx_.p, stream_);
  }
  if (a) launch_scatter(na_, a_slot_.p, a_src_.p, Av, 0.0, Kx_.p, stream_);
  return factor_retry();
}

bool Solver::kkt_update_scaling() {
  launch_hs_scatter(cone_, nhs_, hs_slot_.p, hs_kind_.p, hs_blk_.p, hs_a_.p, hs_b_.p, hs_row_.p,
                    st_.static_reg_A, Kx_.p, stream_);
  return factor_retry();
}

bool Solver::try_factor(double ep, double ea) {
  const double t0 = now_s();
  if (fp32_) {
    // FP32 factorization of K + r*S (S = +1 on variable pivots, -1 on cone
    // pivots), every pivot floored at magnitude r.  Single precision cannot
    // carry the O(1e-8) static shifts: a pivot that small grows the update
    // by 1e8 and wipes out the O(1) entries it is added to.  The shift keeps
    // the growth near 1/r, and FP64 refinement against K removes it.
    const double r = st_.fp32_reg;
    kkt_.factorize_async(Kx_.p, ep + r, ea + r, PIVOT_REGULARIZE, 0.5 * r, r, 1);
    FactorStats fs = kkt_.factor_stats();
    prof_nfact32_++;
    prof_fact_ms_ += (now_s() - t0) * 1e3;
    prof_nfact_++;
    return std::isfinite(fs.min_abs_d) && fs.min_abs_d > 0.0;
  }
  const int mode = st_.linear_solver == CX_LS_REGULARIZED ? PIVOT_REGULARIZE : PIVOT_CHECK;
  kkt_.factorize_async(Kx_.p, ep, ea, mode, st_.dynamic_reg_eps, st_.dynamic_reg_delta, 0);
  FactorStats fs = kkt_.factor_stats();
  prof_fact_ms_ += (now_s() - t0) * 1e3;
  prof_nfact_++;
  if (mode == PIVOT_CHECK && (fs.zero_pivot || fs.stopped)) return false;
  if (mode == PIVOT_REGULARIZE && (double)fs.nreg > 0.05 * (double)N_) return false;
  return std::isfinite(fs.min_abs_d) && fs.min_abs_d > st_.dynamic_reg_eps;
}

void Solver::leave_fp32(double rel) {
  fp32_ = false;
  prof_fp64_fallback_++;
  prof_fallback_rel_ = rel;
  prof_fallback_iter_ = iterations_;
}

bool Solver::factor_retry() {
  double ep = 0.0, ea = 0.0;
  if (fp32_) {
    if (try_factor(ep, ea)) {
      factorized_ = true;
      last_p_ = ep, last_a_ = ea;
      return true;
    }
    leave_fp32(NAN);  // the FP32 factorization broke down: FP64 from here on
  }
  for (int att = 0; att < 18; ++att) {
    if (try_factor(ep, ea)) {
      factorized_ = true;
      last_p_ = ep, last_a_ = ea;
      return true;
    }
    ep = std::max(ep * 10.0, st_.dynamic_reg_delta);
    ea = std::max(ea * 10.0, st_.dynamic_reg_delta);
  }
  factorized_ = false;
  return false;
}

bool Solver::escalate(const double* rhs, double* x) {
  if (fp32_) {  // a non-finite FP32 solve: FP64 from here on
    leave_fp32(NAN);
    if (!factor_retry()) return false;
    kkt_.solve_async(rhs, x);
    if (count_nonfinite(x, N_) == 0) return true;
  }
  double ep = last_p_, ea = last_a_;
  for (int att = 0; att < 10; ++att) {
    ep = std::max(ep * 1000.0, st_.dynamic_reg_delta);
    ea = std::max(ea * 1000.0, st_.dynamic_reg_delta);
    if (!try_factor(ep, ea)) continue;
    kkt_.solve_async(rhs, x);
    if (count_nonfinite(x, N_) == 0) {
      last_p_ = ep, last_a_ = ea;
      return true;
    }
  }
  return false;
}

int Solver::count_nonfinite(const double* a, int la, const double* b, int lb, const double* c, int lc) {
  RedBatch B{};
  B.op[B.nops++] = nonfin_op(a, la);
  if (b) B.op[B.nops++] = nonfin_op(b, lb);
  if (c) B.op[B.nops++] = nonfin_op(c, lc);
  std::vector<double> r = reduce(B);
  double t = 0;
  for (double v : r) t += v;
  return (int)t;
}

// KktSystem::solve (kkt_system.cpp:296-311)
double Solver::kkt_solve(const double* rhs, double* x) {
  const double t0 = now_s();
  kkt_.solve_async(rhs, x);
  kkt_.spmv_async(Kx_.p, x, t_.p);
  launch_lin2(N_, 1.0, rhs, -1.0, t_.p, r_.p, stream_);
  RedBatch B{};
  B.op[B.nops++] = nonfin_op(x, N_);
  B.op[B.nops++] = dot_op(rhs, rhs, N_);
  B.op[B.nops++] = dot_op(r_.p, r_.p, N_);
  std::vector<double> v = reduce(B);
  prof_nsolve_++;
  if (kkt_.take_solve_error()) {
    // a bounded wait timed out: treat the solve as failed
    launch_fill(x, N_, NAN, stream_);
    prof_solve_ms_ += (now_s() - t0) * 1e3;
    return NAN;
  }
  const double rn = std::max(std::sqrt(v[1]), 1e-30);
  double rr = v[2];
  if (v[0] != 0.0) {
    escalate(rhs, x);
    kkt_.spmv_async(Kx_.p, x, t_.p);
    launch_lin2(N_, 1.0, rhs, -1.0, t_.p, r_.p, stream_);
    RedBatch B2{};
    B2.op[B2.nops++] = dot_op(r_.p, r_.p, N_);
    rr = reduce(B2)[0];
  }
  double rel = std::sqrt(rr) / rn;
  // An FP32 factorization is the preconditioner of FP64 GMRES on K; when the
  // residual misses the acceptance bound below, the solve falls back to an
  // FP64 factorization for the rest of this solve().
  if (fp32_) rel = gmres_refine(rhs, x, rn, rel);
  for (int it = 0; !fp32_ && it < st_.refine_max_iter && rel > st_.refine_tol; ++it) {
    kkt_.solve_async(r_.p, dsol_.p);
    prof_nsolve_++;
    launch_axpby(N_, 1.0, dsol_.p, 1.0, x, stream_);
    kkt_.spmv_async(Kx_.p, x, t_.p);
    launch_lin2(N_, 1.0, rhs, -1.0, t_.p, r_.p, stream_);
    RedBatch B3{};
    B3.op[B3.nops++] = dot_op(r_.p, r_.p, N_);
    rel = std::sqrt(reduce(B3)[0]) / rn;
  }
  prof_solve_ms_ += (now_s() - t0) * 1e3;
  // An FP32 solve is accepted when GMRES reached refine_tol, or stopped within
  // FP32_ACCEPT_REL: a residual no FP64 solve with refine_max_iter steps
  // beats on the same ill-conditioned late-iteration systems.
  const double accept = std::max(st_.refine_tol, FP32_ACCEPT_REL);
  if (!fp32_ || rel <= accept) prof_max_rel_ = std::max(prof_max_rel_, rel);
  if (fp32_ && !(rel <= accept)) {
    leave_fp32(rel);
    if (!factor_retry()) return NAN;
    return kkt_solve(rhs, x);
  }
  return rel;
}

// The constant and affine solves of one iteration (kkt_solve on [-q; b] and
// on [rx; s - rz]) have independent right-hand sides: both go through one
// launch per sweep, and their refinement steps are paired while both still
// refine.  Each side takes exactly the steps kkt_solve takes on it alone, so
// the results are the same.  The first side gets kkt_solve's escalation; the
// second is returned (true) only when it came out finite on the factorization
// the first ended with: otherwise the caller solves it again with kkt_solve,
// after the convergence tests, where the reference solves it
// (solver_impl.cpp:165-231, kkt_system.cpp:266-311).  FP64 factorizations only.
bool Solver::kkt_solve_pair(const double* rhs0, double* x0, const double* rhs1, double* x1) {
  const double t0 = now_s();
  kkt_.solve2_async(rhs0, x0, rhs1, x1);
  kkt_.spmv_async(Kx_.p, x0, t_.p);
  launch_lin2(N_, 1.0, rhs0, -1.0, t_.p, r_.p, stream_);
  kkt_.spmv_async(Kx_.p, x1, t2_.p);
  launch_lin2(N_, 1.0, rhs1, -1.0, t2_.p, r2_.p, stream_);
  RedBatch B{};
  B.op[B.nops++] = nonfin_op(x0, N_);
  B.op[B.nops++] = dot_op(rhs0, rhs0, N_);
  B.op[B.nops++] = dot_op(r_.p, r_.p, N_);
  B.op[B.nops++] = nonfin_op(x1, N_);
  B.op[B.nops++] = dot_op(rhs1, rhs1, N_);
  B.op[B.nops++] = dot_op(r2_.p, r2_.p, N_);
  std::vector<double> v = reduce(B);
  prof_nsolve_ += 2;
  if (kkt_.take_solve_error()) {
    launch_fill(x0, N_, NAN, stream_);
    prof_solve_ms_ += (now_s() - t0) * 1e3;
    return false;
  }
  const double rn0 = std::max(std::sqrt(v[1]), 1e-30), rn1 = std::max(std::sqrt(v[4]), 1e-30);
  double rr0 = v[2];
  bool ok1 = v[3] == 0.0;
  if (v[0] != 0.0) {
    escalate(rhs0, x0);  // a new factorization: the second side is solved again later
    ok1 = false;
    kkt_.spmv_async(Kx_.p, x0, t_.p);
    launch_lin2(N_, 1.0, rhs0, -1.0, t_.p, r_.p, stream_);
    RedBatch B2{};
    B2.op[B2.nops++] = dot_op(r_.p, r_.p, N_);
    rr0 = reduce(B2)[0];
  }
  double rel0 = std::sqrt(rr0) / rn0, rel1 = std::sqrt(v[5]) / rn1;
  for (int it0 = 0, it1 = 0;;) {
    const bool a0 = it0 < st_.refine_max_iter && rel0 > st_.refine_tol;
    const bool a1 = ok1 && it1 < st_.refine_max_iter && rel1 > st_.refine_tol;
    if (!a0 && !a1) break;
    if (a0 && a1) {
      kkt_.solve2_async(r_.p, dsol_.p, r2_.p, dsol2_.p);
      prof_nsolve_ += 2;
    } else {
      kkt_.solve_async(a0 ? r_.p : r2_.p, a0 ? dsol_.p : dsol2_.p);
      prof_nsolve_++;
    }
    RedBatch B3{};
    if (a0) {
      launch_axpby(N_, 1.0, dsol_.p, 1.0, x0, stream_);
      kkt_.spmv_async(Kx_.p, x0, t_.p);
      launch_lin2(N_, 1.0, rhs0, -1.0, t_.p, r_.p, stream_);
      B3.op[B3.nops++] = dot_op(r_.p, r_.p, N_);
    }
    if (a1) {
      launch_axpby(N_, 1.0, dsol2_.p, 1.0, x1, stream_);
      kkt_.spmv_async(Kx_.p, x1, t2_.p);
      launch_lin2(N_, 1.0, rhs1, -1.0, t2_.p, r2_.p, stream_);
      B3.op[B3.nops++] = dot_op(r2_.p, r2_.p, N_);
    }
    std::vector<double> w = reduce(B3);
    if (a0) rel0 = std::sqrt(w[0]) / rn0, ++it0;
    if (a1) rel1 = std::sqrt(w[a0 ? 1 : 0]) / rn1, ++it1;
  }
  prof_max_rel_ = std::max(prof_max_rel_, rel0);
  if (ok1) prof_max_rel_ = std::max(prof_max_rel_, rel1);
  prof_solve_ms_ += (now_s() - t0) * 1e3;
  return ok1;
}

// Restarted GMRES (classical Gram-Schmidt applied twice, Givens rotations) on
// K dx = r, right-preconditioned by the FP32 factorization M: the Krylov basis
// V and the preconditioned vectors Z = M^-1 V stay on the device, the small
// Hessenberg least-squares problem on the host.  Every residual and update is
// FP64; each cycle ends with the true residual rhs - K x.
double Solver::gmres_refine(const double* rhs, double* x, double rn, double rel) {
  const int m = GM_RESTART;
  const long ld = N_;
  if (gm_v_.n < (size_t)(m + 1) * N_) gm_v_.alloc((size_t)(m + 1) * N_);
  if (gm_z_.n < (size_t)m * N_) gm_z_.alloc((size_t)m * N_);
  gm_part_.alloc((size_t)KRY_MAX * KRY_BLOCKS);
  gm_h_.alloc(2 * KRY_MAX);
  if (!gm_host_) CX_HIP(hipHostMalloc((void**)&gm_host_, 2 * KRY_MAX * sizeof(double), hipHostMallocDefault));
  int solves = 0;
  std::vector<double> H((size_t)(m + 1) * m), cs(m), sn(m), g(m + 1);
  auto h_at = [&](int i, int j) -> double& { return H[(size_t)j * (m + 1) + i]; };
  while (std::isfinite(rel) && rel > st_.refine_tol && solves < GM_MAX_SOLVES) {
    const double beta = rel * rn;  // r_ holds rhs - K x
    double* V = gm_v_.p;
    double* Z = gm_z_.p;
    launch_lin2(N_, 1.0 / beta, r_.p, 0.0, r_.p, V, stream_);
    std::fill(g.begin(), g.end(), 0.0);
    g[0] = beta;
    int k = 0;  // completed Arnoldi steps
    bool breakdown = false;
    while (k < m && solves < GM_MAX_SOLVES) {
      const int j = k;
      kkt_.solve_async(V + j * ld, Z + j * ld);
      ++solves;
      prof_nsolve_++;
      double* w = V + (j + 1) * ld;
      kkt_.spmv_async(Kx_.p, Z + j * ld, w);
      for (int pass = 0; pass < 2; ++pass) {
        launch_mdot(N_, j + 1, V, ld, w, gm_part_.p, gm_h_.p + pass * KRY_MAX, stream_);
        launch_mgs_update(N_, j + 1, V, ld, gm_h_.p + pass * KRY_MAX, w, stream_);
      }
      CX_HIP(hipMemcpyAsync(gm_host_, gm_h_.p, 2 * KRY_MAX * sizeof(double), hipMemcpyDeviceToHost, stream_));
      RedBatch B{};
      B.op[B.nops++] = dot_op(w, w, N_);
      const double hn = std::sqrt(reduce(B)[0]);  // synchronizes the stream
      for (int i = 0; i <= j; ++i) h_at(i, j) = gm_host_[i] + gm_host_[KRY_MAX + i];
      h_at(j + 1, j) = hn;
      if (!std::isfinite(hn)) return NAN;
      for (int i = 0; i < j; ++i) {  // previous rotations
        const double a = h_at(i, j), b = h_at(i + 1, j);
        h_at(i, j) = cs[i] * a + sn[i] * b;
        h_at(i + 1, j) = -sn[i] * a + cs[i] * b;
      }
      const double a = h_at(j, j), b = h_at(j + 1, j);
      const double den = std::hypot(a, b);
      cs[j] = den > 0 ? a / den : 1.0;
      sn[j] = den > 0 ? b / den : 0.0;
      h_at(j, j) = den;
      h_at(j + 1, j) = 0.0;
      g[j + 1] = -sn[j] * g[j];
      g[j] = cs[j] * g[j];
      k = j + 1;
      if (hn <= 1e-300 * beta) {
        breakdown = true;
        break;
      }
      launch_scale(N_, 1.0 / hn, w, stream_);
      if (std::fabs(g[k]) <= st_.refine_tol * rn) break;
    }
    (void)breakdown;
    // y = R^-1 g, x += Z y
    KryCoef y{};
    for (int i = k - 1; i >= 0; --i) {
      double v = g[i];
      for (int c = i + 1; c < k; ++c) v -= h_at(i, c) * y.c[c];
      y.c[i] = h_at(i, i) != 0.0 ? v / h_at(i, i) : 0.0;
    }
    launch_multi_axpy(N_, k, Z, ld, y, x, stream_);
    kkt_.spmv_async(Kx_.p, x, t_.p);
    launch_lin2(N_, 1.0, rhs, -1.0, t_.p, r_.p, stream_);
    RedBatch B{};
    B.op[B.nops++] = dot_op(r_.p, r_.p, N_);
    rel = std::sqrt(reduce(B)[0]) / rn;
  }
  return rel;
}

// ---------------------------------------------------------------------------
// updateData / setWarmStart: solver_impl.cpp:57-100
// ---------------------------------------------------------------------------
bool Solver::update_data(const int* Pp, const int* Pi, const double* Px, const double* q,
                         const int* Ap, const int* Ai, const double* Ax, const double* b) {
  if (!setup_done_) return false;
  bw_valid_ = false;
  const bool hasP = Px != nullptr, hasA = Ax != nullptr;
  const int nP = (int)Pi_ref_.size(), nA = (int)Ai_ref_.size();
  auto same = [](const int* p, const int* i, int cols, const std::vector<int>& pr,
                 const std::vector<int>& ir) {
    if (!p || (!i && pr[cols] > 0)) return false;
    for (int c = 0; c <= cols; ++c)
      if (p[c] != pr[c]) return false;
    for (int k = 0; k < pr[cols]; ++k)
      if (i[k] != ir[k]) return false;
    return true;
  };
  // stage scaled values on the device
  if (hasP) {
    if (!same(Pp, Pi, n_, Pp_ref_, Pi_ref_)) return false;
    DevBuf<double> raw;
    raw.upload(Px, nP, stream_);
    if (st_.equilibrate)
      launch_scale_csc(n_, dPp_.p, dPi_.p, raw.p, ec_, d_ed_.p, d_ed_.p, dPx_new_.p, stream_);
    else
      launch_copy(nP, raw.p, dPx_new_.p, stream_);
    launch_scatter(npd_, pd_slot_.p, pd_src_.p, dPx_new_.p, st_.static_reg_P, Kx_.p, stream_);
    launch_scatter(npo_, po_slot_.p, po_src_.p, dPx_new_.p, 0.0, Kx_.p, stream_);
    CX_HIP(hipStreamSynchronize(stream_));
  }
  if (hasA) {
    if (!same(Ap, Ai, n_, Ap_ref_, Ai_ref_)) return false;
    DevBuf<double> raw;
    raw.upload(Ax, nA, stream_);
    if (st_.equilibrate)
      launch_scale_csc(n_, dAp_.p, dAi_.p, raw.p, 1.0, d_ee_.p, d_ed_.p, dAx_new_.p, stream_);
    else
      launch_copy(nA, raw.p, dAx_new_.p, stream_);
    launch_scatter(na_, a_slot_.p, a_src_.p, dAx_new_.p, 0.0, Kx_.p, stream_);
    CX_HIP(hipStreamSynchronize(stream_));
  }
  // The reference refactorizes here even when only q and b change
  // (kkt_system.cpp:136); with K untouched since a successful factorization
  // that factorization would succeed again, so it is skipped.
  if ((hasP || hasA || !factorized_) && !factor_retry()) return false;
  if (hasP) launch_copy(nP, dPx_new_.p, dPx_.p, stream_);
  if (hasA) launch_copy(nA, dAx_new_.p, dAx_.p, stream_);
  if (q) {
    std::vector<double> qv(q, q + n_);
    if (st_.equilibrate)
      for (int i = 0; i < n_; ++i) qv[i] *= ec_ * ed_[i];
    dq_.upload(qv, stream_);
  }
  if (b) {
    std::vector<double> bv(b, b + m_);
    if (st_.equilibrate)
      for (int i = 0; i < m_; ++i) bv[i] *= ee_[i];
    db_.upload(bv, stream_);
  }
  CX_HIP(hipStreamSynchronize(stream_));
  return true;
}

// update_data with every staged value in the per-GPU scratch array instead of
// freshly allocated device buffers, and the q / b rescaling on the device (the
// same products the host does in update_data, so results are identical).  The
// same checks and the same order of K updates as update_data; rejects any
// piece longer than kScratchDoubles before touching the device.
bool Solver::update_data_noalloc(const int* Pp, const int* Pi, const double* Px, const double* q,
                                 const int* Ap, const int* Ai, const double* Ax, const double* b) {
  if (!setup_done_) return false;
  const bool hasP = Px != nullptr, hasA = Ax != nullptr;
  const long nP = (long)Pi_ref_.size(), nA = (long)Ai_ref_.size();
  if ((hasP && nP > kScratchDoubles) || (hasA && nA > kScratchDoubles) || (q && n_ > kScratchDoubles) ||
      (b && m_ > kScratchDoubles)) {
    std::fprintf(stderr,
                 "[conicxx_amd] update_data_noalloc: nnz(P) %ld, nnz(A) %ld, n %d, m %d; each staged piece is "
                 "limited to %ld values\n",
                 nP, nA, n_, m_, kScratchDoubles);
    return false;
  }
  double* stage = scratch_address(device_);
  if (!stage) return false;
  bw_valid_ = false;
  auto same = [](const int* p, const int* i, int cols, const std::vector<int>& pr, const std::vector<int>& ir) {
    if (!p || (!i && pr[cols] > 0)) return false;
    for (int c = 0; c <= cols; ++c)
      if (p[c] != pr[c]) return false;
    for (int k = 0; k < pr[cols]; ++k)
      if (i[k] != ir[k]) return false;
    return true;
  };
  if (hasP) {
    if (!same(Pp, Pi, n_, Pp_ref_, Pi_ref_)) return false;
    if (nP) CX_HIP(hipMemcpyAsync(stage, Px, nP * sizeof(double), hipMemcpyHostToDevice, stream_));
    if (st_.equilibrate)
      launch_scale_csc(n_, dPp_.p, dPi_.p, stage, ec_, d_ed_.p, d_ed_.p, dPx_new_.p, stream_);
    else
      launch_copy((int)nP, stage, dPx_new_.p, stream_);
    launch_scatter(npd_, pd_slot_.p, pd_src_.p, dPx_new_.p, st_.static_reg_P, Kx_.p, stream_);
    launch_scatter(npo_, po_slot_.p, po_src_.p, dPx_new_.p, 0.0, Kx_.p, stream_);
  }
  if (hasA) {
    if (!same(Ap, Ai, n_, Ap_ref_, Ai_ref_)) {
      CX_HIP(hipStreamSynchronize(stream_));
      return false;
    }
    // the stream orders this copy after the kernels that read P's values from the array
    if (nA) CX_HIP(hipMemcpyAsync(stage, Ax, nA * sizeof(double), hipMemcpyHostToDevice, stream_));
    if (st_.equilibrate)
      launch_scale_csc(n_, dAp_.p, dAi_.p, stage, 1.0, d_ee_.p, d_ed_.p, dAx_new_.p, stream_);
    else
      launch_copy((int)nA, stage, dAx_new_.p, stream_);
    launch_scatter(na_, a_slot_.p, a_src_.p, dAx_new_.p, 0.0, Kx_.p, stream_);
  }
  if ((hasP || hasA || !factorized_) && !factor_retry()) return false;
  if (hasP) launch_copy((int)nP, dPx_new_.p, dPx_.p, stream_);
  if (hasA) launch_copy((int)nA, dAx_new_.p, dAx_.p, stream_);
  if (q && n_) {
    CX_HIP(hipMemcpyAsync(stage, q, n_ * sizeof(double), hipMemcpyHostToDevice, stream_));
    if (st_.equilibrate)
      launch_vscale(n_, d_ed_.p, ec_, stage, dq_.p, stream_);
    else
      launch_copy(n_, stage, dq_.p, stream_);
  }
  if (b && m_) {
    CX_HIP(hipMemcpyAsync(stage, b, m_ * sizeof(double), hipMemcpyHostToDevice, stream_));
    if (st_.equilibrate)
      launch_vscale(m_, d_ee_.p, 1.0, stage, db_.p, stream_);
    else
      launch_copy(m_, stage, db_.p, stream_);
  }
  // the array is free for the next call once this returns
  CX_HIP(hipStreamSynchronize(stream_));
  return true;
}

void Solver::set_warm_start(const double* x, const double* s, const double* z) {
  if (!setup_done_) return;
  std::vector<double> xv(x, x + n_), sv(s, s + m_), zv(z, z + m_);
  if (st_.equilibrate) {  // equilibration.cpp:135-139
    for (int i = 0; i < n_; ++i) xv[i] /= ed_[i];
    for (int i = 0; i < m_; ++i) zv[i] = zv[i] * ec_ / ee_[i];
    for (int i = 0; i < m_; ++i) sv[i] *= ee_[i];
  }
  wx_.upload(xv, stream_);
  ws_.upload(sv, stream_);
  wz_.upload(zv, stream_);
  CX_HIP(hipStreamSynchronize(stream_));
  have_warm_ = true;
}

// Backward (adjoint) pass at the optimum of the last solve.  The solution
// map's differential follows from the KKT conditions with complementarity
// linearized as ds = -H dz:
//   [P, A'; A, -H] [dx; dz] = [-(dP x + dq + dA' z); db - dA x].
// H is the final NT scaling Hs with, on every SOC block, the eigenvalue of the
// directions orthogonal to z1 set to s0/z0 (the limit of the exact
// linearization arw(z) ds + arw(s) dz = 0; NT gives it only on the central
// path).  Nonnegative and zero rows are exact already.  The corrected KKT
// matrix (static regularization included) is factored once per solved problem,
// on the first backward call; for L with gradients (wx, ws, wz) in original
// units the adjoint solve K [u; v] = [D wx; E wz / c - H (ws / E)] (plus the
// usual refinement) then gives
//   dL/dq = -c D u,  dL/db = E v,  dL/dA_ij = -(z_i u_j + v_i x_j) E_i D_j,
//   dL/dP_ij (upper) = -(u_i x_j + u_j x_i) c D_i D_j, diagonal -u_i x_i c D_i^2,
// with x, z the normalized optimum in the equilibrated space.
int Solver::backward(const double* wx, const double* ws, const double* wz, double* gP, double* gq,
                     double* gA, double* gb) {
  if (!setup_done_ || !bw_valid_ || status_ != CX_SOLVED || !factorized_) return 0;
  const int nA = Ap_ref_[n_];
  std::vector<double> tx(n_, 0.0), us(m_, 0.0), ntz(m_, 0.0);
  for (int i = 0; i < n_; ++i) tx[i] = wx ? wx[i] * ed_[i] : 0.0;
  for (int i = 0; i < m_; ++i) {
    us[i] = ws ? ws[i] / ee_[i] : 0.0;
    ntz[i] = wz ? -wz[i] * ee_[i] / ec_ : 0.0;
  }
  fp32_ = false;  // the adjoint factorization is FP64
  if (!bw_factored_) {
    if (!soc_dims_.empty()) {
      bw_tcorr_.alloc(soc_dims_.size());
      bw_tdir_.alloc(m_);
      launch_soc_tangent(cone_, s_.p, z_.p, bw_tcorr_.p, bw_tdir_.p, stream_);
      launch_hs_scatter(cone_, nhs_, hs_slot_.p, hs_kind_.p, hs_blk_.p, hs_a_.p, hs_b_.p, hs_row_.p,
                        st_.static_reg_A, Kx_.p, stream_, bw_tcorr_.p, bw_tdir_.p);
      if (!factor_retry()) return 0;
    }
    bw_factored_ = true;
  }
  bw_rhs_.alloc(N_);
  bw_sol_.alloc(N_);
  bw_us_.upload(us, stream_);
  bw_ntz_.upload(ntz, stream_);
  if (n_) CX_HIP(hipMemcpyAsync(bw_rhs_.p, tx.data(), n_ * sizeof(double), hipMemcpyHostToDevice, stream_));
  if (m_) launch_neg_hs_plus(cone_, m_, bw_us_.p, bw_ntz_.p, bw_rhs_.p + n_, stream_);
  if (m_ && ws && !soc_dims_.empty())
    launch_soc_tangent_apply(cone_, bw_tcorr_.p, bw_tdir_.p, bw_us_.p, bw_rhs_.p + n_, stream_);
  const double rel = kkt_solve(bw_rhs_.p, bw_sol_.p);
  if (!std::isfinite(rel)) return 0;
  const double* u = bw_sol_.p;
  const double* v = bw_sol_.p + n_;
  const double inv_tau = 1.0 / std::max(tau_, 1e-30);
  bw_gP_.alloc(std::max(1, P_nnz_));
  bw_gA_.alloc(std::max(1, nA));
  bw_gq_.alloc(std::max(1, n_));
  bw_gb_.alloc(std::max(1, m_));
  launch_grad_csc(n_, P_nnz_, dPp_.p, dPi_.p, 1, u, v, x_.p, z_.p, inv_tau, d_ed_.p, d_ed_.p, ec_,
                  bw_gP_.p, stream_);
  launch_grad_csc(n_, nA, dAp_.p, dAi_.p, 0, u, v, x_.p, z_.p, inv_tau, d_ee_.p, d_ed_.p, 1.0,
                  bw_gA_.p, stream_);
  launch_vmuldiv(n_, u, d_ed_.p, nullptr, -ec_, bw_gq_.p, stream_);
  launch_vmuldiv(m_, v, d_ee_.p, nullptr, 1.0, bw_gb_.p, stream_);
  if (gP) bw_gP_.download(gP, P_nnz_, stream_);
  if (gA) bw_gA_.download(gA, nA, stream_);
  if (gq) bw_gq_.download(gq, n_, stream_);
  if (gb) bw_gb_.download(gb, m_, stream_);
  CX_HIP(hipStreamSynchronize(stream_));
  return 1;
}

// KktSystem::updateScalingAndFactorize + KktSystem::solve on their own
// (kkt_system.cpp:139-150, 296-311): the NT scaling at (s, z) (the identity of
// construction when NULL) is written into K, K is factorized with the retry
// ladder, and rhs is solved with up to refine_max_iter refinement steps
// (negative: the settings' value).  Vectors are in the solver's internal
// (equilibrated) space.  Returns the final relative residual |rhs - K x| / |rhs|
// against the statically regularized K, NaN on a failed factorization.
double Solver::kkt_factor_solve(const double* s, const double* z, const double* rhs, double* x,
                                int refine_max_iter, int* factorizations) {
  if (!setup_done_) return NAN;
  bw_valid_ = false;
  fp32_ = st_.factor_precision == CX_FACTOR_FP32_REFINED;
  prof_nfact_ = prof_nsolve_ = prof_nfact32_ = prof_fp64_fallback_ = 0;
  if (s && z && m_ > 0) {
    s_.upload(s, m_, stream_);
    z_.upload(z, m_, stream_);
    launch_cone_scaling(cone_, s_.p, z_.p, stream_);
  } else {
    launch_cone_scaling(cone_, identity_.p, identity_.p, stream_);
  }
  double rel = NAN;
  if (kkt_update_scaling()) {
    rhs_.upload(rhs, N_, stream_);
    const int keep = st_.refine_max_iter;
    if (refine_max_iter >= 0) st_.refine_max_iter = refine_max_iter;
    rel = kkt_solve(rhs_.p, sol_.p);
    st_.refine_max_iter = keep;
    sol_.download(x, N_, stream_);
    CX_HIP(hipStreamSynchronize(stream_));
  }
  if (factorizations) *factorizations = prof_nfact_;
  fp32_ = false;
  return rel;
}

// ---------------------------------------------------------------------------
// IPM pieces: solver_impl.cpp
// ---------------------------------------------------------------------------
void Solver::margins(const double* v, double* mn, double* ps) {
  launch_margins(cone_, v, red_partial_.p, red_out_.p, stream_);
  CX_HIP(hipMemcpyAsync(host_red_, red_out_.p, 2 * sizeof(double), hipMemcpyDeviceToHost, stream_));
  CX_HIP(hipStreamSynchronize(stream_));
  *mn = host_red_[0];
  *ps = host_red_[1];
}

void Solver::zero_zero_block(double* s) {
  if (zero_dim_) launch_fill(s, zero_dim_, 0.0, stream_);
}

void Solver::shift_cold(double* v) {  // solver_impl.cpp:104-116
  double mn, ps;
  margins(v, &mn, &ps);
  const double deg = (double)std::max(degree_, 1);
  const double target = std::max(1.0, 0.1 * ps / deg);
  if (mn <= 0.0) {
    launch_unit_shift(cone_, v, -mn, stream_);
    launch_unit_shift(cone_, v, target, stream_);
  } else if (mn < target) {
    launch_unit_shift(cone_, v, target - mn, stream_);
  }
}

void Solver::ensure_interior_warm(double* v) {  // solver_impl.cpp:118-124
  double mn, ps;
  margins(v, &mn, &ps);
  if (mn <= 1e-8) launch_unit_shift(cone_, v, -mn + 1e-6, stream_);
}

bool Solver::initial_point() {  // solver_impl.cpp:126-158
  launch_cone_scaling(cone_, identity_.p, identity_.p, stream_);
  if (!kkt_update_scaling()) return false;
  if (P_nnz_ == 0) {
    launch_fill(rhs_.p, n_, 0.0, stream_);
    launch_copy(m_, db_.p, rhs_.p + n_, stream_);
    kkt_solve(rhs_.p, sol_.p);
    launch_copy(n_, sol_.p, x_.p, stream_);
    launch_lin2(m_, -1.0, sol_.p + n_, 0.0, sol_.p + n_, s_.p, stream_);
    launch_lin2(n_, -1.0, dq_.p, 0.0, dq_.p, rhs_.p, stream_);
    launch_fill(rhs_.p + n_, m_, 0.0, stream_);
    kkt_solve(rhs_.p, sol_.p);
    launch_copy(m_, sol_.p + n_, z_.p, stream_);
  } else {
    launch_lin2(n_, -1.0, dq_.p, 0.0, dq_.p, rhs_.p, stream_);
    launch_copy(m_, db_.p, rhs_.p + n_, stream_);
    kkt_solve(rhs_.p, sol_.p);
    launch_copy(n_, sol_.p, x_.p, stream_);
    launch_copy(m_, sol_.p + n_, z_.p, stream_);
    launch_lin2(m_, -1.0, sol_.p + n_, 0.0, sol_.p + n_, s_.p, stream_);
  }
  zero_zero_block(s_.p);
  shift_cold(s_.p);
  shift_cold(z_.p);
  tau_ = 1.0;
  kappa_ = 1.0;
  return count_nonfinite(x_.p, n_, s_.p, m_, z_.p, m_) == 0;
}

bool Solver::constant_solve() {  // solver_impl.cpp:165-175
  constant_rhs(rhs_.p);
  kkt_solve(rhs_.p, sol_.p);
  return constant_finish(sol_.p);
}

void Solver::constant_rhs(double* rhs) {  // [-q; b]
  launch_lin2(n_, -1.0, dq_.p, 0.0, dq_.p, rhs, stream_);
  launch_copy(m_, db_.p, rhs + n_, stream_);
}

bool Solver::constant_finish(const double* sol) {
  launch_copy(n_, sol, x1_.p, stream_);
  launch_copy(m_, sol + n_, z1_.p, stream_);
  pmul(x1_.p, Px1_.p);
  return count_nonfinite(x1_.p, n_, z1_.p, m_) == 0;
}

void Solver::residuals() {  // solver_impl.cpp:177-186
  pmul(x_.p, Px_.p);
  launch_csr_mv(n_, dAp_.p, dAi_.p, at_src_.p, dAx_.p, z_.p, wn_.p, stream_);  // A'z
  launch_lin3(n_, -1.0, wn_.p, -1.0, Px_.p, -tau_, dq_.p, rx_.p, stream_);
  launch_csr_mv(m_, ar_ptr_.p, ar_col_.p, ar_src_.p, dAx_.p, x_.p, wm_.p, stream_);  // A x
  launch_lin3(m_, 1.0, wm_.p, 1.0, s_.p, -tau_, db_.p, rz_.p, stream_);
  RedBatch B{};
  B.op[B.nops++] = dot_op(dq_.p, x_.p, n_);
  B.op[B.nops++] = dot_op(db_.p, z_.p, m_);
  B.op[B.nops++] = dot_op(s_.p, z_.p, m_);
  B.op[B.nops++] = dot_op(x_.p, Px_.p, n_);
  B.op[B.nops++] = dot_op(rx_.p, rx_.p, n_);
  B.op[B.nops++] = dot_op(rz_.p, rz_.p, m_);
  std::vector<double> v = reduce(B);
  dqx_ = v[0], dbz_ = v[1], dsz_ = v[2], dxPx_ = v[3], nrx2_ = v[4], nrz2_ = v[5];
  rtau_ = dqx_ + dbz_ + kappa_ + dxPx_ / tau_;
}

bool Solver::tau_elim(const double* xv, const double* zv, double dtau_rhs, double dkappa_rhs,
                      double* dtau) {
  // xi = x/tau; num uses P xv; den uses (xi - x1)' P (xi - x1)
  pmul(xv, wn_.p);                                              // P xv
  launch_lin2(n_, 1.0 / tau_, x_.p, -1.0, x1_.p, wn2_.p, stream_);  // xi - x1
  pmul(wn2_.p, t_.p);                                           // P (xi - x1)  (t_ has length N >= n)
  RedBatch B{};
  B.op[B.nops++] = dot_op(dq_.p, xv, n_);
  B.op[B.nops++] = dot_op(db_.p, zv, m_);
  B.op[B.nops++] = dot_op(x_.p, wn_.p, n_);
  B.op[B.nops++] = dot_op(dq_.p, x1_.p, n_);
  B.op[B.nops++] = dot_op(db_.p, z1_.p, m_);
  B.op[B.nops++] = dot_op(wn2_.p, t_.p, n_);
  B.op[B.nops++] = dot_op(x1_.p, Px1_.p, n_);
  std::vector<double> v = reduce(B);
  const double num = dtau_rhs - dkappa_rhs / tau_ + v[0] + v[1] + 2.0 * (v[2] / tau_);
  const double den = kappa_ / tau_ - v[3] - v[4] + v[5] - v[6];
  if (den == 0.0 || !std::isfinite(num) || !std::isfinite(den)) return false;
  *dtau = num / den;
  return true;
}

bool Solver::affine_step() {  // solver_impl.cpp:192-231
  affine_rhs(rhs_.p);
  kkt_solve(rhs_.p, sol_.p);
  return affine_finish(sol_.p);
}

void Solver::affine_rhs(double* rhs) {  // [rx; s - rz]
  launch_copy(n_, rx_.p, rhs, stream_);
  launch_lin2(m_, 1.0, s_.p, -1.0, rz_.p, rhs + n_, stream_);
}

bool Solver::affine_finish(const double* sol) {
  if (count_nonfinite(sol, N_) != 0) return false;
  const double dtau_rhs = rtau_, dkappa_rhs = tau_ * kappa_;
  if (!tau_elim(sol, sol + n_, dtau_rhs, dkappa_rhs, &dtaua_)) return false;
  launch_lin2(n_, 1.0, sol, dtaua_, x1_.p, dxa_.p, stream_);
  launch_lin2(m_, 1.0, sol + n_, dtaua_, z1_.p, dza_.p, stream_);
  launch_neg_hs_plus(cone_, m_, dza_.p, s_.p, dsa_.p, stream_);
  dkappaa_ = -(dkappa_rhs + kappa_ * dtaua_) / tau_;
  return std::isfinite(dtaua_) && std::isfinite(dkappaa_) &&
         count_nonfinite(dxa_.p, n_, dza_.p, m_, dsa_.p, m_) == 0;
}

bool Solver::combined_step(double sigma, double mu) {  // solver_impl.cpp:233-286
  launch_combined_rhs(cone_, m_, z_.p, dsa_.p, dza_.p, rz_.p, sigma * mu, 1.0 - sigma, dsk_.p,
                      rhs_.p + n_, stream_);
  launch_lin2(n_, 1.0 - sigma, rx_.p, 0.0, rx_.p, rhs_.p, stream_);
  kkt_solve(rhs_.p, sol_.p);
  if (count_nonfinite(sol_.p, N_) != 0) return false;
  const double dtau_rhs = (1.0 - sigma) * rtau_;
  const double dkappa_rhs = tau_ * kappa_ + dtaua_ * dkappaa_ - sigma * mu;
  if (!tau_elim(sol_.p, sol_.p + n_, dtau_rhs, dkappa_rhs, &dtau_)) return false;
  launch_lin2(n_, 1.0, sol_.p, dtau_, x1_.p, dx_.p, stream_);
  launch_lin2(m_, 1.0, sol_.p + n_, dtau_, z1_.p, dz_.p, stream_);
  launch_neg_hs_plus(cone_, m_, dz_.p, dsk_.p, ds_.p, stream_);
  dkappa_ = -(dkappa_rhs + kappa_ * dtau_) / tau_;
  return std::isfinite(dtau_) && std::isfinite(dkappa_) &&
         count_nonfinite(dx_.p, n_, dz_.p, m_, ds_.p, m_) == 0;
}

double Solver::step_length(const double* ds, const double* dz, double dtau, double dkappa) {
  double alpha = 1.0;  // solver_impl.cpp:288-296
  if (dtau < 0) alpha = std::min(alpha, -tau_ / dtau);
  if (dkappa < 0) alpha = std::min(alpha, -kappa_ / dkappa);
  launch_max_step(cone_, z_.p, dz, red_partial_.p, red_out_.p, stream_);
  launch_max_step(cone_, s_.p, ds, red_partial_.p + RED_BLOCKS, red_out_.p + 1, stream_);
  CX_HIP(hipMemcpyAsync(host_red_, red_out_.p, 2 * sizeof(double), hipMemcpyDeviceToHost, stream_));
  CX_HIP(hipStreamSynchronize(stream_));
  alpha = std::min(alpha, host_red_[0]);
  alpha = std::min(alpha, host_red_[1]);
  return alpha;
}

void Solver::finalize() {  // solver_impl.cpp:342-365
  const bool inf = status_ == CX_PRIMAL_INFEASIBLE || status_ == CX_DUAL_INFEASIBLE;
  const double scale = inf ? 1.0 / std::max(kappa_, 1e-30) : 1.0 / std::max(tau_, 1e-30);
  xo_.assign(n_, 0.0), so_.assign(m_, 0.0), zo_.assign(m_, 0.0);
  if (n_) x_.download(xo_.data(), n_, stream_);
  if (m_) s_.download(so_.data(), m_, stream_), z_.download(zo_.data(), m_, stream_);
  CX_HIP(hipStreamSynchronize(stream_));
  for (auto& v : xo_) v *= scale;
  for (auto& v : so_) v *= scale;
  for (auto& v : zo_) v *= scale;
  if (st_.equilibrate) {
    for (int i = 0; i < n_; ++i) xo_[i] *= ed_[i];
    for (int i = 0; i < m_; ++i) zo_[i] = zo_[i] * ee_[i] / ec_;
    for (int i = 0; i < m_; ++i) so_[i] /= ee_[i];
  }
  double pobj = (0.5 * dxPx_ / tau_ + dqx_) / tau_;
  if (st_.equilibrate) pobj /= ec_;
  objective_ = pobj;
  double gap = dsz_ / (tau_ * tau_);
  if (st_.equilibrate) gap /= ec_;
  gap_ = gap;
  pres_ = std::sqrt(nrz2_) / std::max(tau_, 1e-30);
  dres_ = std::sqrt(nrx2_) / std::max(tau_, 1e-30);
  mu_out_ = mu();
}

int Solver::solve() {  // solver_impl.cpp:367-466
  const double t0 = now_s();
  status_ = CX_UNSOLVED;
  bw_valid_ = false;
  fp32_ = st_.factor_precision == CX_FACTOR_FP32_REFINED;
  prof_nfact32_ = prof_fp64_fallback_ = 0;
  prof_fallback_iter_ = -1;
  prof_max_rel_ = prof_fallback_rel_ = 0;
  iterations_ = 0;
  objective_ = pres_ = dres_ = gap_ = mu_out_ = 0;
  prof_fact_ms_ = prof_solve_ms_ = 0;
  prof_nfact_ = prof_nsolve_ = 0;
  kkt_.reset_timing();
  if (!setup_done_) {
    status_ = CX_NUMERICAL_ERROR;
    return status_;
  }
  try {
    const bool use_warm = st_.warm_start && have_warm_;
    if (use_warm) {
      launch_copy(n_, wx_.p, x_.p, stream_);
      launch_copy(m_, ws_.p, s_.p, stream_);
      launch_copy(m_, wz_.p, z_.p, stream_);
      tau_ = 1.0;
      kappa_ = 1.0;
      zero_zero_block(s_.p);
      ensure_interior_warm(s_.p);
      ensure_interior_warm(z_.p);
    } else if (!initial_point()) {
      status_ = CX_NUMERICAL_ERROR;
      finalize();
      solve_time_ = now_s() - t0;
      return status_;
    }
    int final_status = CX_MAX_ITERATIONS, iterations = 0;
    for (int iter = 0; iter < st_.max_iter; ++iter) {
      iterations = iter + 1;
      iterations_ = iter;  // current iteration, for the profile
      launch_cone_scaling(cone_, s_.p, z_.p, stream_);
      if (!kkt_update_scaling()) {
        final_status = CX_NUMERICAL_ERROR;
        break;
      }
      // The residuals do not use the constant solve, so they come first; with
      // an FP64 factorization the affine solve (whose right-hand side they
      // give) is then paired with the constant solve
      residuals();
      bool affine_ready = false;
      bool const_ok;
      if (fp32_) {
        const_ok = constant_solve();
      } else {
        constant_rhs(rhs_.p);
        affine_rhs(rhs2_.p);
        affine_ready = kkt_solve_pair(rhs_.p, sol_.p, rhs2_.p, sol2_.p);
        const_ok = constant_finish(sol_.p);
      }
      if (!const_ok) {
        final_status = CX_NUMERICAL_ERROR;
        break;
      }
      const double mu_now = mu();
      if (st_.verbose >= 1) {
        std::fprintf(stderr, "[conicxx_amd] iter %3d mu %.3e |rx| %.3e |rz| %.3e t %.3fs (fact %.1f ms, solve %.1f ms)\n",
                     iter, mu_now, std::sqrt(nrx2_) / tau_, std::sqrt(nrz2_) / tau_, now_s() - t0,
                     prof_fact_ms_, prof_solve_ms_);
        std::fflush(stderr);
      }
      const double nrx = std::sqrt(nrx2_) / tau_;
      const double nrz = std::sqrt(nrz2_) / tau_;
      const double pobj = (0.5 * dxPx_ / tau_ + dqx_) / tau_;
      const double dobj = (-0.5 * dxPx_ / tau_ - dbz_) / tau_;
      const double gap_abs = std::fabs(pobj - dobj);
      const double gap_tol = st_.tol_gap_abs + st_.tol_gap_rel * std::max(std::fabs(pobj), std::fabs(dobj));
      if (nrx < st_.tol_feas && nrz < st_.tol_feas && gap_abs < gap_tol && tau_ > 1e-6) {
        final_status = CX_SOLVED;
        break;
      }
      {  // checkInfeasibility: solver_impl.cpp:327-340
        const double ratio = tau_ / std::max(kappa_, 1e-30);
        if (ratio <= st_.tol_infeas) {
          RedBatch B{};
          B.op[B.nops++] = dot_op(z_.p, z_.p, m_);
          B.op[B.nops++] = dot_op(x_.p, x_.p, n_);
          std::vector<double> nv = reduce(B);
          if (dbz_ < -st_.tol_infeas * std::max(1.0, std::sqrt(nv[0]))) {
            final_status = CX_PRIMAL_INFEASIBLE;
            break;
          }
          if (dqx_ < -st_.tol_infeas * std::max(1.0, std::sqrt(nv[1]))) {
            final_status = CX_DUAL_INFEASIBLE;
            break;
          }
        }
      }
      if (!(affine_ready ? affine_finish(sol2_.p) : affine_step())) {
        final_status = CX_NUMERICAL_ERROR;
        break;
      }
      const double aa = step_length(dsa_.p, dza_.p, dtaua_, dkappaa_);
      RedBatch B{};
      B.op[B.nops++] = RedOp{RED_DOT_AXPY, m_, s_.p, z_.p, dsa_.p, dza_.p, aa};
      const double saz = reduce(B)[0];
      const double tau_a = tau_ + aa * dtaua_;
      const double kap_a = kappa_ + aa * dkappaa_;
      const double mu_aff = (saz + tau_a * kap_a) / (double)(degree_ + 1);
      double sigma;
      if (mu_now > 0) {
        const double r = std::max(mu_aff / mu_now, 0.0);
        sigma = std::min(std::max(r * r * r, 0.0), 1.0);
      } else {
        sigma = 1.0;
      }
      if (!combined_step(sigma, mu_now)) {
        final_status = CX_NUMERICAL_ERROR;
        break;
      }
      const double alpha = step_length(ds_.p, dz_.p, dtau_, dkappa_) * st_.max_step_fraction;
      if (!std::isfinite(alpha) || alpha <= 0.0) {
        final_status = CX_NUMERICAL_ERROR;
        break;
      }
      launch_axpby(n_, alpha, dx_.p, 1.0, x_.p, stream_);
      launch_axpby(m_, alpha, ds_.p, 1.0, s_.p, stream_);
      launch_axpby(m_, alpha, dz_.p, 1.0, z_.p, stream_);
      tau_ += alpha * dtau_;
      kappa_ += alpha * dkappa_;
      const double sc = std::max(tau_, kappa_);
      if (sc > 1e8 || sc < 1e-8) {
        const double inv = 1.0 / sc;
        launch_scale(n_, inv, x_.p, stream_);
        launch_scale(m_, inv, s_.p, stream_);
        launch_scale(m_, inv, z_.p, stream_);
        tau_ *= inv;
        kappa_ *= inv;
      }
    }
    status_ = final_status;
    iterations_ = iterations;
    bw_valid_ = final_status == CX_SOLVED;
    bw_factored_ = false;
    finalize();
    if (final_status == CX_SOLVED && st_.warm_start) {
      launch_copy(n_, x_.p, wx_.p, stream_);
      launch_copy(m_, s_.p, ws_.p, stream_);
      launch_copy(m_, z_.p, wz_.p, stream_);
      CX_HIP(hipStreamSynchronize(stream_));
      have_warm_ = true;
    }
  } catch (const std::exception&) {
    status_ = CX_NUMERICAL_ERROR;
  }
  solve_time_ = now_s() - t0;
  prof_total_ms_ = solve_time_ * 1e3;
  return status_;
}

int Solver::get_solution(double* x, double* s, double* z, double* obj, cx_info* info) const {
  if (x && !xo_.empty()) std::memcpy(x, xo_.data(), sizeof(double) * xo_.size());
  if (s && !so_.empty()) std::memcpy(s, so_.data(), sizeof(double) * so_.size());
  if (z && !zo_.empty()) std::memcpy(z, zo_.data(), sizeof(double) * zo_.size());
  if (obj) *obj = objective_;
  if (info) {
    info->iterations = iterations_;
    info->primal_residual = pres_;
    info->dual_residual = dres_;
    info->duality_gap = gap_;
    info->mu = mu_out_;
    info->setup_time_s = setup_time_;
    info->solve_time_s = solve_time_;
  }
  return status_;
}

void Solver::kkt_stats(double* out8) const {
  const SupernodalPlan& P = kkt_.plan();
  out8[0] = P.N;
  out8[1] = nnzK_;
  out8[2] = P.ns;
  out8[3] = P.nlevels;
  out8[4] = P.nnzL;
  out8[5] = P.flops;
  out8[6] = P.max_front;
  out8[7] = P.nnzL_exact;
}

void Solver::profile(double* out13) const {
  double* out8 = out13;
  const DeviceKKT::KindStats& t = kkt_.kind_stats(DeviceKKT::T_TRAIL);
  out8[0] = prof_fact_ms_;
  out8[1] = prof_solve_ms_;
  out8[2] = prof_total_ms_ - prof_fact_ms_ - prof_solve_ms_;
  out8[3] = prof_nfact_;
  out8[4] = prof_nsolve_;
  out8[5] = t.launches ? t.ms / t.launches : 0.0;
  out8[6] = (double)t.launches;
  out8[7] = t.work;
  out13[8] = prof_nfact32_;
  out13[9] = prof_fp64_fallback_;
  out13[10] = prof_max_rel_;
  out13[11] = prof_fallback_rel_;
  out13[12] = prof_fallback_iter_;
}

int Solver::kernel_stats(double* out, int max_kinds) const {
  int n = std::min(max_kinds, (int)DeviceKKT::T_NKINDS);
  for (int k = 0; k < n; ++k) {
    const DeviceKKT::KindStats& t = kkt_.kind_stats(k);
    out[3 * k + 0] = t.ms;
    out[3 * k + 1] = (double)t.launches;
    out[3 * k + 2] = t.work;
  }
  return n;
}

}  // namespace cx
