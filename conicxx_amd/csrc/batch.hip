// Batched persistent interior-point solver (see batch.hpp).
//
// One workgroup (256 threads) owns one problem from its initial point to its
// final status, then takes the next problem from a shared counter; problems
// are handed out largest first.  The method is the single-problem path's
// (solver_impl.cpp:104-466 restated step by step: initial point, NT scaling,
// constant solve, residuals, tau elimination, Mehrotra predictor-corrector,
// step length, rescaling, convergence and infeasibility tests, finalize,
// warm-start carry), with the KKT factorization specialised as described in
// batch.hpp.  Every loop is bounded (max_iter, retry ladders); there are no
// waits on other workgroups.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <stdexcept>

#include "batch.hpp"
#include "cone_device.hpp"
#include "equilibrate.hpp"

namespace cx {

namespace {

using namespace conedev;

constexpr int BG = 256;  // threads per problem
// per-problem record of the last solve / backward in A.info:
//   0-4 residuals, gap, mu, objective; 5-6 factorizations, reduced solves;
//   7-10 device ticks (total, factorization, solves, residuals); 11 KKT solves;
//   12-14 backward ticks, backward solve ticks, backward reduced solves;
//   15 ticks of the final KKT solve; 16 ticks of the adjoint factorization
constexpr int INFO_STRIDE = 20;
constexpr int BATCH_REFINE_MAX = 40;

enum { PV_CHECK = 0, PV_REGULARIZE = 1 };

struct BatchArgs {
  int nprob;
  const BatchDesc* desc;
  const int* order;
  int* counter;
  // settings
  double tol_feas, tol_gap_abs, tol_gap_rel, tol_infeas, sreg_p, sreg_a, dyn_eps, dyn_delta;
  double refine_tol, step_frac;
  double adj_tol;  // refinement target of the adjoint (backward) solve
  int max_iter, refine_max_iter, mode, warm_start;
  int lds_doubles;
  // structure (entry indices global, row/column values problem-relative)
  const int *pf_ptr, *pf_col, *pf_src;
  const int *ac_ptr, *ac_row;
  const int *ar_ptr, *ar_col, *ar_src;
  const int *soc_off, *soc_dim;
  const int *rk, *kr;  // per row: index among the kept rows or -1 (bound row); kept rows
  // values
  const double *Pv, *Av, *q, *b, *ed, *ee;
  double *W, *Mg;
  double *beta, *vv;
  int* ident;
  // warm start and results
  double *wx, *ws, *wz;
  int* have_warm;
  double *xo, *so, *zo, *info;
  int *status, *iters;
  // differentiation: prepare the adjoint factorization after a Solved problem
  int diff;
  double *tcorr, *tdir, *xh, *sh, *zh, *diff_a;
  int* diff_ok;
  // backward pass inputs and outputs (concatenated like the data)
  const double *gwx, *gws, *gwz;
  double *gP, *gq, *gA, *gb;
  int* counter2;
};
// The kernels read their arguments in place from the kernarg segment through
// a constant-address-space reference: the loads are scalar, and the compiler
// may hoist them past any store (a by-reference copy of the argument struct
// would live in scratch memory and be reloaded after every store).
using CArgs = const __attribute__((address_space(4))) BatchArgs;
__device__ inline CArgs& kernel_args() { return *(CArgs*)__builtin_amdgcn_kernarg_segment_ptr(); }

// ---------------------------------------------------------------------------
// workgroup reductions (all threads receive the result)
// ---------------------------------------------------------------------------
template <int K>
__device__ inline void gsum(double (&v)[K], double* red) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int c = 0; c < K; ++c) {
    double t = v[c];
    for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o, 64);
    if (lane == 0) red[c * 4 + wv] = t;
  }
  __syncthreads();
#pragma unroll
  for (int c = 0; c < K; ++c) v[c] = (red[c * 4] + red[c * 4 + 1]) + (red[c * 4 + 2] + red[c * 4 + 3]);
  __syncthreads();
}
__device__ inline double gsum1(double v, double* red) {
  double a[1] = {v};
  gsum<1>(a, red);
  return a[0];
}
__device__ inline double gmin1(double v, double* red) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
  if (lane == 0) red[wv] = v;
  __syncthreads();
  const double r = fmin(fmin(red[0], red[1]), fmin(red[2], red[3]));
  __syncthreads();
  return r;
}
__device__ inline double dmax(double a, double b) { return a > b ? a : b; }
__device__ inline double dmin(double a, double b) { return a < b ? a : b; }

// ---------------------------------------------------------------------------
// per-problem view
// ---------------------------------------------------------------------------
struct PV {
  int n, m, nz, nn, nsoc, nk, N, nr, pnnz;
  int co, so, ko, pfp, acp, arp;
  double cur_a;  // (2,2) regularization of the current factorization
  int nfact, nsolve, nkkt;  // factorizations, reduced solves and KKT solves (with refinement)
  long long t_fact, t_solve, t_res;  // device clock ticks in factorization, reduced solves, residuals
  long long t_kkt_last;              // device clock ticks of the latest KKT solve (with refinement)
  double* stage;      // LDS staging for panels when M is in global memory (else null)
  long stage_cap;
  const double* tcorr;  // SOC tangent correction of the adjoint matrix (null in the IPM)
  const double* tdir;
  double ec;
  const double *q, *b;
  double *x, *Px, *rx, *x1, *Px1, *dxa, *dx, *wn, *wn2;
  double *s, *z, *rz, *z1, *dza, *dsa, *dz, *ds, *dsk, *wm, *wm2;
  double *rhs, *sol, *res, *tmp, *dsol;
  double* M;
  bool m_lds, v_lds;  // M / the work vectors are in LDS (else in global memory)
  ConeDev C;
  // scalar state
  double tau, kappa, rtau, dqx, dbz, dsz, dxPx, nrx2, nrz2;
  double dtaua, dkappaa, dtau, dkappa;
  double last_p, last_a;
};

__device__ inline long packed_size(int nr) { return (long)nr * (nr + 1) / 2; }
__device__ inline long colbase(int j, int nr) { return (long)j * nr - (long)j * (j - 1) / 2; }

// y = P x (full symmetric CSR)
__device__ void pmul(CArgs& A, const PV& p, const double* x, double* y) {
  for (int i = threadIdx.x; i < p.n; i += BG) {
    double acc = 0.0;
    for (int e = A.pf_ptr[p.pfp + i]; e < A.pf_ptr[p.pfp + i + 1]; ++e) acc += A.Pv[A.pf_src[e]] * x[A.pf_col[e]];
    y[i] = acc;
  }
  __syncthreads();
}
// y = A x (all rows)
__device__ void amul(CArgs& A, const PV& p, const double* x, double* y) {
  for (int i = threadIdx.x; i < p.m; i += BG) {
    double acc = 0.0;
    for (int e = A.ar_ptr[p.arp + i]; e < A.ar_ptr[p.arp + i + 1]; ++e) acc += A.Av[A.ar_src[e]] * x[A.ar_col[e]];
    y[i] = acc;
  }
  __syncthreads();
}
// y = A' z
__device__ void atmul(CArgs& A, const PV& p, const double* z, double* y) {
  for (int j = threadIdx.x; j < p.n; j += BG) {
    double acc = 0.0;
    for (int e = A.ac_ptr[p.acp + j]; e < A.ac_ptr[p.acp + j + 1]; ++e) acc += A.Av[e] * z[A.ac_row[e]];
    y[j] = acc;
  }
  __syncthreads();
}
__device__ inline void vcopy(int len, const double* x, double* y) {
  for (int i = threadIdx.x; i < len; i += BG) y[i] = x[i];
  __syncthreads();
}
__device__ inline void vfill(int len, double v, double* y) {
  for (int i = threadIdx.x; i < len; i += BG) y[i] = v;
  __syncthreads();
}
__device__ inline void vlin2(int len, double a, const double* x, double b, const double* y, double* out) {
  for (int i = threadIdx.x; i < len; i += BG) out[i] = a * x[i] + b * y[i];
  __syncthreads();
}
__device__ inline void vlin3(int len, double a, const double* x, double b, const double* y, double c,
                             const double* z, double* out) {
  for (int i = threadIdx.x; i < len; i += BG) out[i] = a * x[i] + b * y[i] + c * z[i];
  __syncthreads();
}
__device__ inline void vaxpy(int len, double a, const double* x, double* y) {
  for (int i = threadIdx.x; i < len; i += BG) y[i] += a * x[i];
  __syncthreads();
}
__device__ inline void vscale(int len, double a, double* x) {
  for (int i = threadIdx.x; i < len; i += BG) x[i] *= a;
  __syncthreads();
}
__device__ double count_nonfinite(int len, const double* x, double* red) {
  double c = 0.0;
  for (int i = threadIdx.x; i < len; i += BG) c += isfinite(x[i]) ? 0.0 : 1.0;
  return gsum1(c, red);
}

// ---------------------------------------------------------------------------
// cones (ConeSet over one problem)
// ---------------------------------------------------------------------------
__device__ void cone_scaling(const PV& p, const double* s, const double* z) {
  for (int i = threadIdx.x; i < p.nn; i += BG) {
    const int r = p.nz + i;
    const double w = sqrt(s[r] / z[r]);
    p.C.w[r] = w;
    p.C.hsd[r] = w * w;
  }
  for (int k = threadIdx.x; k < p.nsoc; k += BG)
    soc_scaling<1>(0, p.C.soc_off[k], p.C.soc_dim[k], s, z, p.C.v, &p.C.beta[k], &p.C.vv[k], &p.C.ident[k]);
  __syncthreads();
}
// out = -(Hs x + y)
__device__ void neg_hs_plus(const PV& p, const double* x, const double* y, double* out) {
  for (int i = threadIdx.x; i < p.nz; i += BG) out[i] = -y[i];
  for (int i = threadIdx.x; i < p.nn; i += BG) {
    const int r = p.nz + i;
    out[r] = -(p.C.hsd[r] * x[r] + y[r]);
  }
  for (int k = threadIdx.x; k < p.nsoc; k += BG) soc_neg_hs_plus<1>(0, p.C, k, x, y, out);
  __syncthreads();
}
__device__ void combined_rhs(const PV& p, const double* z, const double* dsa, const double* dza,
                             const double* rz, double smu, double oms, double* dsc, double* tail) {
  for (int i = threadIdx.x; i < p.nz; i += BG) {
    dsc[i] = 0.0;
    tail[i] = 0.0 - oms * rz[i];
  }
  for (int i = threadIdx.x; i < p.nn; i += BG) {
    const int r = p.nz + i;
    const double w = p.C.w[r];
    const double lam = w * z[r];
    const double a1 = dsa[r] / w;
    const double a2 = w * dza[r];
    const double dc = lam * lam + a1 * a2 - smu;
    const double lid = dc / lam;
    const double dsv = w * lid;
    dsc[r] = dsv;
    tail[r] = dsv - oms * rz[r];
  }
  for (int k = threadIdx.x; k < p.nsoc; k += BG) soc_combined<1>(0, p.C, k, z, dsa, dza, rz, smu, oms, dsc, tail);
  __syncthreads();
}
__device__ double cone_max_step(const PV& p, const double* x, const double* dx, double* red) {
  double a = INFINITY;
  for (int i = threadIdx.x; i < p.nn; i += BG) {
    const int r = p.nz + i;
    if (dx[r] < 0) a = fmin(a, -x[r] / dx[r]);
  }
  for (int k = threadIdx.x; k < p.nsoc; k += BG)
    a = fmin(a, soc_max_step<1>(0, p.C.soc_off[k], p.C.soc_dim[k], x, dx));
  return gmin1(a, red);
}
// ConeSet::margins: (min margin, sum of positive margins), the orthant counted as one block
__device__ void margins(const PV& p, const double* x, double* mn, double* ps, double* red) {
  double nmin = INFINITY, smin = INFINITY, spos = 0.0;
  for (int i = threadIdx.x; i < p.nn; i += BG) nmin = fmin(nmin, x[p.nz + i]);
  for (int k = threadIdx.x; k < p.nsoc; k += BG) {
    const int off = p.C.soc_off[k], d = p.C.soc_dim[k];
    double ss = 0.0;
    for (int i = 1; i < d; ++i) ss += x[off + i] * x[off + i];
    const double mg = x[off] - sqrt(ss);
    if (isfinite(mg)) smin = fmin(smin, mg), spos += fmax(mg, 0.0);
  }
  nmin = gmin1(nmin, red);
  smin = gmin1(smin, red);
  spos = gsum1(spos, red);
  double m0 = smin, p0 = spos;
  if (p.nn > 0 && isfinite(nmin)) {
    m0 = fmin(m0, nmin);
    p0 += fmax(nmin, 0.0);
  }
  if (!isfinite(m0)) m0 = 0.0;
  *mn = m0;
  *ps = p0;
}
__device__ void unit_shift(const PV& p, double* x, double alpha) {
  for (int i = threadIdx.x; i < p.nn; i += BG) x[p.nz + i] += alpha;
  for (int k = threadIdx.x; k < p.nsoc; k += BG) x[p.C.soc_off[k]] += alpha;
  __syncthreads();
}

// Entry (a, c), a >= c, of SOC block k's Hs as the factored matrix holds it,
// with the adjoint's tangent correction when p.tcorr is set
__device__ inline double hs_stored(const PV& p, int k, int off, bool id, double be, double vvk, const double* v,
                                   int a, int c) {
  if (id) return (a == c) ? 1.0 : 0.0;
  double h = soc_hs_entry(be, vvk, v, a, c);
  if (p.tcorr && a > 0 && c > 0) {
    const double* u = p.tdir + off;
    h += p.tcorr[k] * ((a == c ? 1.0 : 0.0) - u[a] * u[c]);
  }
  return h;
}
// out = -(H x + y) with H the (2,2) block's cone part exactly as assemble_M
// stores it (diagonal orthant, stored SOC entries incl. the tangent correction):
// the product the refinement residual needs, so that refinement converges to
// the solution of the matrix that was factored (kkt_system.cpp:296-311 refines
// against the assembled K)
// out = -(H x + 0) on the orthant and SOC rows, no barrier (the zero rows are
// left to the caller)
__device__ void neg_hs_plus_stored_nb(const PV& p, const double* x, double* out) {
  for (int i = threadIdx.x; i < p.nn; i += BG) {
    const int r = p.nz + i;
    out[r] = -(p.C.hsd[r] * x[r] + 0.0);
  }
  for (int k = threadIdx.x; k < p.nsoc; k += BG) {
    const int off = p.C.soc_off[k], d = p.C.soc_dim[k];
    const double* v = p.C.v + off;
    const double be = p.C.beta[k], vvk = p.C.vv[k];
    const bool id = p.C.ident[k] != 0;
    for (int a = 0; a < d; ++a) {
      double acc = 0.0;
      for (int c = 0; c < d; ++c)
        acc += (a >= c ? hs_stored(p, k, off, id, be, vvk, v, a, c) : hs_stored(p, k, off, id, be, vvk, v, c, a)) *
               x[off + c];
      out[off + a] = -(acc + 0.0);
    }
  }
}
__device__ void neg_hs_plus_stored(const PV& p, const double* x, const double* y, double* out) {
  for (int i = threadIdx.x; i < p.nz; i += BG) out[i] = -y[i];
  for (int i = threadIdx.x; i < p.nn; i += BG) {
    const int r = p.nz + i;
    out[r] = -(p.C.hsd[r] * x[r] + y[r]);
  }
  for (int k = threadIdx.x; k < p.nsoc; k += BG) {
    const int off = p.C.soc_off[k], d = p.C.soc_dim[k];
    const double* v = p.C.v + off;
    const double be = p.C.beta[k], vvk = p.C.vv[k];
    const bool id = p.C.ident[k] != 0;
    for (int a = 0; a < d; ++a) {
      double acc = 0.0;
      for (int c = 0; c < d; ++c)
        acc += (a >= c ? hs_stored(p, k, off, id, be, vvk, v, a, c) : hs_stored(p, k, off, id, be, vvk, v, c, a)) *
               x[off + c];
      out[off + a] = -(acc + y[off + a]);
    }
  }
  __syncthreads();
}

// ---------------------------------------------------------------------------
// KKT: M assembly (bound rows folded into the P block, every other row kept)
// ---------------------------------------------------------------------------
__device__ void assemble_M(CArgs& A, const PV& p, double dp, double da) {
  const int nr = p.nr, n = p.n;
  double* M = p.M;
  const long tot = packed_size(nr);
  for (long e = threadIdx.x; e < tot; e += BG) M[e] = 0.0;
  __syncthreads();
  // variable columns: P + dp I + sum over bound rows a^2 / (hs + da), kept rows of A
  for (int j = threadIdx.x; j < n; j += BG) {
    const long cb = colbase(j, nr);
    for (int e = A.pf_ptr[p.pfp + j]; e < A.pf_ptr[p.pfp + j + 1]; ++e) {
      const int l = A.pf_col[e];
      if (l >= j) M[cb + (l - j)] += A.Pv[A.pf_src[e]];
    }
    double dg = dp;
    for (int e = A.ac_ptr[p.acp + j]; e < A.ac_ptr[p.acp + j + 1]; ++e) {
      const int i = A.ac_row[e];
      const double aij = A.Av[e];
      const int kk = A.rk[p.co + i];
      if (kk < 0)
        dg += aij * aij / (p.C.hsd[i] + da);
      else
        M[cb + (n + kk - j)] = aij;
    }
    M[cb] += dg;
  }
  // constraint block: -(H + da I) on the kept rows (zero: 0, nonnegative: Hs_ii, SOC: Hs block)
  for (int kk = threadIdx.x; kk < p.nk; kk += BG) {
    const int i = A.kr[p.ko + kk];
    if (i < p.nz)
      M[colbase(n + kk, nr)] = -da;
    else if (i < p.nz + p.nn)
      M[colbase(n + kk, nr)] = -(p.C.hsd[i] + da);
  }
  for (int k = threadIdx.x; k < p.nsoc; k += BG) {
    const int off = p.C.soc_off[k], d = p.C.soc_dim[k];
    const int k0 = n + A.rk[p.co + off];
    const double* v = p.C.v + off;
    const double be = p.C.beta[k], vvk = p.C.vv[k];
    const bool id = p.C.ident[k] != 0;
    for (int c = 0; c < d; ++c) {
      const long cb = colbase(k0 + c, nr);
      for (int a = c; a < d; ++a) {
        M[cb + (a - c)] = -(hs_stored(p, k, off, id, be, vvk, v, a, c) + (a == c ? da : 0.0));
      }
    }
  }
  __syncthreads();
}

__device__ inline double bcast(double v, int l) {
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), l),
                          __builtin_amdgcn_readlane(__double2loint(v), l));
}
__device__ inline double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
// Cross-lane moves of a double on the VALU: DPP within rows of 16 lanes, and
// the gfx950 permlane swaps between 16-lane rows (odd rows of x with even rows
// of y) and between 32-lane halves (upper half of x with lower half of y).
template <int CTRL>
__device__ inline double dppd(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xf, 0xf, false);
  return __hiloint2double(hi, lo);
}
__device__ inline void swap16(double& x, double& y) {
  const auto lo = __builtin_amdgcn_permlane16_swap((unsigned)__double2loint(x), (unsigned)__double2loint(y), false, false);
  const auto hi = __builtin_amdgcn_permlane16_swap((unsigned)__double2hiint(x), (unsigned)__double2hiint(y), false, false);
  x = __hiloint2double((int)hi[0], (int)lo[0]);
  y = __hiloint2double((int)hi[1], (int)lo[1]);
}
__device__ inline void swap32(double& x, double& y) {
  const auto lo = __builtin_amdgcn_permlane32_swap((unsigned)__double2loint(x), (unsigned)__double2loint(y), false, false);
  const auto hi = __builtin_amdgcn_permlane32_swap((unsigned)__double2hiint(x), (unsigned)__double2hiint(y), false, false);
  x = __hiloint2double((int)hi[0], (int)lo[0]);
  y = __hiloint2double((int)hi[1], (int)lo[1]);
}
// Butterfly sum of 16 per-lane values over the 64 lanes of a wave: every lane
// l returns the wave total of value (l >> 2) & 15.  Each step pairs every lane
// with one in the other half of a lane group and halves the values a lane
// holds: swapping halves of two registers and adding them does a whole step
// on two values (32- and 16-lane steps); within a row, DPP rotates by 8 and
// mirrors 8-lane halves (pairing lanes that differ in bit 2), then quad
// permutations add the last four lanes.  All on the VALU, no LDS traffic.
__device__ inline double wave_sum16(const double (&v)[16]) {
  const int lane = threadIdx.x & 63;
  const bool h3 = lane & 8, h2 = lane & 4;
  double a[8], b[4], c[2];
#pragma unroll
  for (int k = 0; k < 8; ++k) {  // lanes < 32 keep value k, lanes >= 32 value k + 8
    double x = v[k], y = v[k + 8];
    swap32(x, y);
    a[k] = x + y;
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) {  // even rows keep a[k], odd rows a[k + 4]
    double x = a[k], y = a[k + 4];
    swap16(x, y);
    b[k] = x + y;
  }
#pragma unroll
  for (int k = 0; k < 2; ++k) c[k] = (h3 ? b[k + 2] : b[k]) + dppd<0x128>(h3 ? b[k] : b[k + 2]);  // row_ror:8
  double d = (h2 ? c[1] : c[0]) + dppd<0x141>(h2 ? c[0] : c[1]);                                  // row_half_mirror
  d += dppd<0x4E>(d);                                                                              // quad_perm 2,3,0,1
  d += dppd<0xB1>(d);                                                                              // quad_perm 1,0,3,2
  return d;
}
// orders this wave's earlier memory writes before its later reads (M may be in
// LDS or in global memory)
__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
  __builtin_amdgcn_wave_barrier();
}

typedef double v4d __attribute__((ext_vector_type(4)));
// Wave-uniform copies of values the compiler cannot prove uniform: function
// arguments arrive in vector registers and the per-problem view is read from
// private memory.  Uniform copies live in scalar registers, and addresses
// built from them need one vector register (the lane's offset) instead of a
// 64-bit vector address per load.
__device__ inline int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ inline double uni(double v) {
  return __hiloint2double(__builtin_amdgcn_readfirstlane(__double2hiint(v)),
                          __builtin_amdgcn_readfirstlane(__double2loint(v)));
}
template <typename T>
__device__ inline T* uni(T* ptr) {
  if constexpr (sizeof(T*) == 4) {
    return (T*)(unsigned long)(unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(unsigned long)ptr);
  } else {
    const unsigned long v = (unsigned long)ptr;
    const unsigned long lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)v);
    const unsigned long hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(v >> 32));
    return (T*)((hi << 32) | lo);
  }
}
// address-space qualified views of M and the work vectors: LDS accesses then
// compile to ds_read/ds_write and global ones to global_load/global_store with
// a scalar base, instead of flat accesses that wait on both memory counters
typedef __attribute__((address_space(3))) double ldsd;
typedef __attribute__((address_space(1))) double glbd;

// M[idx] as the (uniform) base plus a 32-bit byte offset: a global load then
// takes the scalar-base + vector-offset form, one vector register per address
__device__ inline double ldo(const glbd* M, unsigned idx) {
  return *(const glbd*)((const __attribute__((address_space(1))) char*)M + 8u * idx);
}
__device__ inline double ldo(const ldsd* M, unsigned idx) { return M[idx]; }
__device__ inline double ldo(const double* M, unsigned idx) { return *(const double*)((const char*)M + 8u * idx); }

__device__ inline void sto(glbd* M, unsigned idx, double v) {
  *(glbd*)((__attribute__((address_space(1))) char*)M + 8u * idx) = v;
}
__device__ inline void sto(ldsd* M, unsigned idx, double v) { M[idx] = v; }
__device__ inline void sto(double* M, unsigned idx, double v) { *(double*)((char*)M + 8u * idx) = v; }
// Column offsets along consecutive columns of the packed M: L(i, j) =
// M[o + i] with o = colbase(j) - j, and the next column's o is o + nr - j - 1.
// The running offset is kept opaque in one scalar register; otherwise every
// column's offset of an unrolled loop is precomputed up front and spilled.
__device__ inline unsigned col_off(int j, int nr) { return (unsigned)((int)colbase(j, nr) - j); }
__device__ inline void col_next(unsigned& o, int j, int nr) {
  o += (unsigned)(nr - j - 1);
  __asm__ volatile("" : "+s"(o));
}

constexpr int PB = 32;  // panel width of the blocked factorization
constexpr int JC = 64;  // trailing columns per chunk (16 per wave)
constexpr int COLK = PB * JC + 2 * PB + 8;  // factorization scratch in LDS (doubles)

// Blocked right-looking LDL^T of M (packed lower).  Each 32-column panel is
// factored in LDS (when M itself is in global memory the panel is first staged
// into the LDS region `stage`, row-major, stride 33, and its final values are
// written to M as they are produced): its diagonal block by one wave in
// registers, the rows below by forward substitution against that block, a row
// per thread.  The trailing matrix is then updated on FP64 MFMA 16 x 16 tiles,
// in 64-column chunks of which each wave owns 16 columns.  Pivot rule of the
// reference's backends.  Returns acceptance as in try_factor.
constexpr int PS = PB + 1;  // row stride of a staged panel

template <typename MT, bool STAGED>
__device__ bool factor_M_t(CArgs& A, const PV& p, MT* M, ldsd* wbuf, ldsd* pk, ldsd* stage, double* red) {
  // scalars in scalar registers: A and p live in memory that every store in
  // the loops below might alias, which would force a reload per use
  const int nr = uni(p.nr), pn = uni(p.n), mode = A.mode;
  const double dyn_eps = A.dyn_eps, dyn_delta = A.dyn_delta;
  const int tid = threadIdx.x, lane = tid & 63, wv = uni(tid >> 6);
  M = uni(M), wbuf = uni(wbuf), pk = uni(pk), stage = uni(stage);
  // LDS scratch: Wt of the diagonal block (PB rows of stride 2 PB) in wbuf;
  // the pivot reciprocals follow the pivots pk
  static_assert(JC == 2 * PB, "Wt rows of the diagonal block fill wbuf");
  ldsd* ipk = pk + PB;
  int zero_piv = 0, nreg = 0;
  for (int k0 = 0; k0 < nr; k0 += PB) {
    const int kb = min(PB, nr - k0);
    constexpr bool staged = STAGED;
    // offset of L(i, k0 + c) in M (32-bit: the host bounds the order),
    // colbase(k0 + c) - (k0 + c) + i incrementally from colbase(k0)
    const int cb0 = (int)colbase(k0, nr), w = nr - k0;
    auto off = [&](int i, int c) { return cb0 + c * w - (c * (c + 1)) / 2 + (i - k0); };
    // L(i, k0 + c) for i >= k0 + c (address arithmetic: an indexed local
    // array would be placed in scratch memory)
    auto L = [&](int i, int c) -> decltype(auto) {
      if constexpr (STAGED)
        return (stage[(i - k0) * PS + c]);
      else
        return (M[off(i, c)]);
    };
    if constexpr (staged) {
      if (kb == PB) {
        // a row per thread with all its loads in flight; entries above the
        // diagonal of a column read the previous column's tail (in bounds)
        // and are never used
        for (int i = k0 + tid; i < nr; i += BG) {
          double v[PB];
          unsigned o = col_off(k0, nr);
#pragma unroll
          for (int c = 0; c < PB; ++c) {
            v[c] = ldo(M, o + i);
            col_next(o, k0 + c, nr);
          }
#pragma unroll
          for (int c = 0; c < PB; ++c) stage[(i - k0) * PS + c] = v[c];
        }
      } else {
        for (int c = 0; c < kb; ++c)
          for (int i = k0 + c + tid; i < nr; i += BG) stage[(i - k0) * PS + c] = M[off(i, c)];
      }
      __syncthreads();
    }
    // panel: wave 0 factors the kb x kb diagonal block in registers (lane i
    // holds row i; at each pivot the entries of the scaled column d l reach
    // the other lanes by readlane broadcasts) and leaves D, 1/D and
    // Wt(q, c) = d_q L(c, q) in LDS (rows of stride 2 PB); then every thread
    // solves its own rows below the block against it
    // (L(i, c) = (a_ic - sum_q L(i, q) W(c, q)) / d_c, the row in registers).
    // Two barriers per panel.
    if (wv == 0) {
      double r[PB];
#pragma unroll
      for (int c = 0; c < PB; ++c) {
        const bool ok = c <= lane && lane < kb;
        const double v = L(k0 + (ok ? lane : 0), ok ? c : 0);
        r[c] = ok ? v : 0.0;
      }
#pragma unroll
      for (int j = 0; j < PB; ++j) {
        if (j < kb) {
          double d = bcast(r[j], j);
          const double sg = k0 + j < pn ? 1.0 : -1.0;
          bool reg = false;
          if (mode == PV_REGULARIZE && !(d * sg >= dyn_eps)) {  // regularized_ldlt.cpp:132-139
            d += sg * dyn_delta;
            if (d * sg < dyn_eps) d = sg * dyn_delta;
            reg = true;
          }
          if (tid == 0) {
            nreg += reg ? 1 : 0;
            if (mode != PV_REGULARIZE && d == 0.0) zero_piv = 1;
          }
          const double inv = 1.0 / d;
          const double l = lane > j ? r[j] * inv : 0.0;
          const double w = d * l;  // lane k holds W(k, j) = d L(k, j)
          r[j] = lane > j ? l : (lane == j ? d : r[j]);
          wbuf[j * 2 * PB + lane] = w;
          if (lane == j) pk[j] = d, ipk[j] = inv;
#pragma unroll
          for (int k = j + 1; k < PB; ++k) r[k] -= l * bcast(w, k);
        }
      }
      unsigned o = col_off(k0, nr);
#pragma unroll
      for (int c = 0; c < PB; ++c) {
        if (c <= lane && lane < kb) {
          L(k0 + lane, c) = r[c];
          if constexpr (STAGED) sto(M, o + k0 + lane, r[c]);
        }
        if (c + 1 < kb) col_next(o, k0 + c, nr);
      }
    }
    __syncthreads();
    for (int i = k0 + kb + tid; i < nr; i += BG) {
      // keeps the block's W in LDS instead of hoisting all of it into registers
      __asm__ volatile("" ::: "memory");
      double a[PB];
#pragma unroll
      for (int c = 0; c < PB; ++c) a[c] = L(i, c < kb ? c : 0) * (c < kb ? 1.0 : 0.0);
#pragma unroll
      for (int c = 0; c < PB; ++c) {
        if (c < kb) {
          double acc = a[c];
#pragma unroll
          for (int q = 0; q < c; ++q) acc -= a[q] * wbuf[q * 2 * PB + c];
          a[c] = acc * ipk[c];
        }
      }
      unsigned o = col_off(k0, nr);
#pragma unroll
      for (int c = 0; c < PB; ++c)
        if (c < kb) {
          L(i, c) = a[c];
          if constexpr (STAGED) sto(M, o + i, a[c]);
          if (c + 1 < kb) col_next(o, k0 + c, nr);
        }
    }
    __syncthreads();
    // trailing update C -= L W', W(j, c) = L(j, c) d_c, on v_mfma_f64_16x16x4f64
    // computing transposed 16 x 16 tiles: in each 64-column chunk wave w owns
    // the tile column [jc0 + 16w, jc0 + 16w + 16) from its diagonal tile down
    // to the last row.  Lane l feeds A(row l & 15, k l >> 4) = W(tj + (l & 15), c)
    // and B(k l >> 4, column l & 15) = L(ti + (l & 15), c), c = 4k + (l >> 4),
    // and holds the updates of M(ti + (l & 15), tj + (l >> 4) + 4r): each of its
    // four accesses of M covers 4 columns x 16 consecutive rows.
    for (int jc0 = k0 + kb; jc0 < nr; jc0 += JC) {
      const int tj = jc0 + 16 * wv;
      if (tj < nr) {
        const int l15 = lane & 15, kq = lane >> 4;
        double bf[PB / 4];
#pragma unroll
        for (int k = 0; k < PB / 4; ++k) {
          const int c = 4 * k + kq, jl = tj + l15;
          const bool ok = c < kb && jl < nr;
          const double v = L(ok ? jl : k0, ok ? c : 0), d = pk[ok ? c : 0];
          bf[k] = ok ? v * d : 0.0;
        }
        // M(i, jr[r]) = M[cb[r] + i]; outside the tile the access is clamped to
        // the diagonal entry of a valid column and its value discarded
        unsigned cb[4];
        int jr[4], jd[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          jr[r] = tj + kq + 4 * r;
          jd[r] = min(jr[r], nr - 1);
          cb[r] = (unsigned)((int)colbase(jd[r], nr) - jd[r]);
        }
        // a tile below the diagonal tile, inside the matrix, in a tile column
        // of 16 valid columns needs no masks
        const bool cols_ok = tj + 16 <= nr;
        auto interior = [&](int ti) { return cols_ok && ti > tj && ti + 16 <= nr; };
        auto load_m = [&](int ti, double (&mv)[4]) {
          const int i = ti + l15;
          if (interior(ti)) {
#pragma unroll
            for (int r = 0; r < 4; ++r) mv[r] = ldo(M, cb[r] + i);
          } else {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              // masked-out values only reach masked-out stores: no select
              // (which would wait for the load here)
              const bool ok = i >= jr[r] && i < nr;
              mv[r] = ldo(M, cb[r] + (unsigned)(ok ? i : jd[r]));
            }
          }
        };
        auto store_m = [&](int ti, const double (&mv)[4], const v4d& acc) {
          const int i = ti + l15;
          if (interior(ti)) {
#pragma unroll
            for (int r = 0; r < 4; ++r) sto(M, cb[r] + i, mv[r] - acc[r]);
          } else {
#pragma unroll
            for (int r = 0; r < 4; ++r)
              if (i >= jr[r] && i < nr) sto(M, cb[r] + i, mv[r] - acc[r]);
          }
        };
        auto load_a = [&](int ti, double (&av)[PB / 4]) {
          const int il = ti + l15;
#pragma unroll
          for (int k = 0; k < PB / 4; ++k) {
            const int c = 4 * k + kq;
            const bool ok = c < kb && il < nr;
            const double v = L(ok ? il : k0, ok ? c : 0);
            av[k] = ok ? v : 0.0;
          }
        };
        auto tile = [&](const double (&av)[PB / 4]) {
          v4d acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
          for (int k = 0; k < PB / 4; ++k) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(bf[k], av[k], acc, 0, 0, 0);
          return acc;
        };
        // the tiles of a column are disjoint.  Two register sets, each
        // refilled with the tile two ahead as soon as its own tile is stored:
        // no rotation of registers, since moving a register whose load is in
        // flight waits for the load
        double mA[4], mB[4], aA[PB / 4], aB[PB / 4];
        load_m(tj, mA);
        load_a(tj, aA);
        load_m(tj + 16, mB);
        load_a(tj + 16, aB);
        for (int ti = tj; ti < nr; ti += 32) {
          store_m(ti, mA, tile(aA));
          load_m(ti + 32, mA);
          load_a(ti + 32, aA);
          if (ti + 16 < nr) store_m(ti + 16, mB, tile(aB));
          load_m(ti + 48, mB);
          load_a(ti + 48, aB);
        }
      }
    }
    __syncthreads();
  }
  double mind = INFINITY, nanf = 0.0;
  for (int k = tid; k < nr; k += BG) {
    const double d = M[colbase(k, nr)];
    if (isnan(d)) nanf = 1.0;
    mind = fmin(mind, fabs(d));
  }
  mind = gmin1(mind, red);
  nanf = gsum1(nanf, red);
  double zr[2] = {(double)zero_piv, (double)nreg};
  gsum<2>(zr, red);
  if (nanf > 0) mind = NAN;
  if (A.mode == PV_CHECK && zr[0] > 0) return false;
  if (A.mode == PV_REGULARIZE && zr[1] > 0.05 * (double)p.N) return false;
  return isfinite(mind) && mind > A.dyn_eps;
}

// Every panel staged in LDS when the whole first panel fits the stage (the
// usual case when M lives in global memory), else direct access to M.
__device__ bool factor_M(CArgs& A, const PV& p, double* wbuf, double* pk, double* stage, long stage_cap,
                         double* red) {
  ldsd *w = (ldsd*)wbuf, *d = (ldsd*)pk;
  if (p.m_lds) return factor_M_t<ldsd, false>(A, p, (ldsd*)p.M, w, d, nullptr, red);
  if (stage != nullptr && (long)p.nr * PS <= stage_cap)
    return factor_M_t<glbd, true>(A, p, (glbd*)p.M, w, d, (ldsd*)stage, red);
  return factor_M_t<glbd, false>(A, p, (glbd*)p.M, w, d, nullptr, red);
}

__device__ bool try_factor(CArgs& A, PV& p, double ep, double ea, double* colk, double* red) {
  p.cur_a = A.sreg_a + ea;
  ++p.nfact;
  const long long t0 = wall_clock64();
  assemble_M(A, p, A.sreg_p + ep, p.cur_a);
  const bool ok = factor_M(A, p, colk, colk + PB * JC, p.stage, p.stage_cap, red);
  p.t_fact += wall_clock64() - t0;
  return ok;
}

// factorizeWithRetry (kkt_system.cpp:242-264)
__device__ bool factor_retry(CArgs& A, PV& p, double* colk, double* red) {
  double ep = 0.0, ea = 0.0;
  for (int att = 0; att < 18; ++att) {
    if (try_factor(A, p, ep, ea, colk, red)) {
      p.last_p = ep, p.last_a = ea;
      return true;
    }
    ep = dmax(ep * 10.0, A.dyn_delta);
    ea = dmax(ea * 10.0, A.dyn_delta);
  }
  return false;
}

// x = K^-1 rhs through the factored M (no refinement).  Bound rows:
// y_i = (a_i x_j - r2_i) / (hs_i + da) after the solve of M.
template <typename MT, typename VT>
__device__ void reduced_solve_t(CArgs& A, const PV& p, const double* rhs, double* x, const MT* M, VT* t,
                                ldsd* part);
// colk: LDS scratch of the factorization (256 doubles of partial sums here)
__device__ void reduced_solve(CArgs& A, PV& p, const double* rhs, double* x, double* colk) {
  ++p.nsolve;
  const long long t0 = wall_clock64();
  // the intermediate t lives in LDS: among the vectors, or else in the panel
  // stage, which only the factorization uses
  if (p.m_lds && p.v_lds)
    reduced_solve_t(A, p, rhs, x, (const ldsd*)p.M, (ldsd*)p.tmp, (ldsd*)colk);
  else if (p.v_lds)
    reduced_solve_t(A, p, rhs, x, (const glbd*)p.M, (ldsd*)p.tmp, (ldsd*)colk);
  else if (!p.m_lds && p.stage != nullptr && p.stage_cap >= p.N)
    reduced_solve_t(A, p, rhs, x, (const glbd*)p.M, (ldsd*)p.stage, (ldsd*)colk);
  else
    reduced_solve_t(A, p, rhs, x, (const double*)p.M, p.tmp, (ldsd*)colk);
  p.t_solve += wall_clock64() - t0;
}
template <typename MT, typename VT>
__device__ void reduced_solve_t(CArgs& A, const PV& p, const double* rhs, double* x, const MT* M, VT* t,
                                ldsd* part) {
  // scalars and pointers in scalar registers (A and p live in memory that the
  // stores below might alias, which would force a reload per use)
  const int n = uni(p.n), m = uni(p.m), nr = uni(p.nr), nk = uni(p.nk), tid = threadIdx.x;
  const double da = uni(p.cur_a);
  M = uni(M), t = uni(t), part = uni(part), rhs = uni(rhs), x = uni(x);
  const int* __restrict__ acp = A.ac_ptr + uni(p.acp);
  const int* __restrict__ acr = A.ac_row;
  const int* __restrict__ rk = A.rk + uni(p.co);
  const int* __restrict__ kr = A.kr + uni(p.ko);
  const double* __restrict__ Av = A.Av;
  const double* __restrict__ hsd = uni(p.C.hsd);
  for (int j = tid; j < n; j += BG) {
    double acc = rhs[j];
    for (int e = acp[j]; e < acp[j + 1]; ++e) {
      const int i = acr[e];
      if (rk[i] < 0) acc += Av[e] * rhs[n + i] / (hsd[i] + da);
    }
    t[j] = acc;
  }
  for (int kk = tid; kk < nk; kk += BG) t[n + kk] = rhs[n + kr[kk]];
  __syncthreads();
  const int lane = tid & 63, wv = uni(tid >> 6);
  const int nb = (nr + 63) >> 6;
  // L(i, c) = M[lx(i, c)] for i >= c (32-bit offsets: the host bounds the order)
  auto lx = [&](int i, int c) { return (unsigned)((int)colbase(c, nr) - c + i); };
  auto col0 = [&](int j) { return col_off(j, nr); };
  auto next_col = [&](unsigned& o, int j) { col_next(o, j, nr); };
  // L y = t in 64-row blocks (block b holds rows 64b .. 64b + kb - 1; only the
  // last block can be short).  Two phases per block, each closed by a barrier:
  //   A(b), b >= 1: the rows of block b take the contribution of block b - 1,
  //     each wave 16 of its columns (operands loaded during B(b - 1)), the four
  //     partial sums to LDS;
  //   B(b): wave 0 adds them and runs the dependent recurrence of the diagonal
  //     block on registers (operands loaded during B(b - 1)); meanwhile waves
  //     1..3 apply block b - 1 to every row below block b, a row per thread
  //     with all 64 of its loads in flight.  Every wave then loads its
  //     operands of A(b + 1), and wave 0 the next diagonal block.
  // The critical path per block is the recurrence and one 16-term update.
  // Wave 0 and waves 1..3 run separate loops with the same sequence of
  // barriers (the branch is uniform per wave), so the registers of the
  // diagonal block and of the row updates are never live together.
  double La[16];
  // Loads of a full 64-row block (every block but possibly the last) need no
  // masks: an index above the diagonal of a column lands in the tail of the
  // previous column, in bounds, and the value is discarded.  Each load is then
  // a scalar base plus the lane's offset.
  // La[u] = L(kr + lane, kc + 16 wv + u): rows of one block against this
  // wave's 16 columns of the (full) block left of it
  auto load_a = [&](int kr, int kc) {
    const int i = kr + lane, j0 = kc + 16 * wv;
    unsigned o = col0(j0);
    if (kr + 64 <= nr) {
#pragma unroll
      for (int u = 0; u < 16; ++u) {
        La[u] = ldo(M, o + i);
        next_col(o, j0 + u);
      }
    } else {
      const bool ok = i < nr;
#pragma unroll
      for (int u = 0; u < 16; ++u) {
        const double v = ldo(M, ok ? o + i : 0u);
        La[u] = ok ? v : 0.0;
        next_col(o, j0 + u);
      }
    }
  };
  auto phase_a_fwd = [&](int k0) {
    const int kp = k0 - 64;
    double acc = 0.0;
#pragma unroll
    for (int u = 0; u < 16; ++u) acc += La[u] * t[kp + 16 * wv + u];
    part[wv * 64 + lane] = acc;
    __syncthreads();
  };
  if (wv == 0) {
    double Lr[64];  // Lr[c] = L(k0 + lane, k0 + c), c < lane < kb
    auto load_diag = [&](int k0, int kb) {
      unsigned o = col0(k0);
      if (kb == 64) {
#pragma unroll
        for (int c = 0; c < 64; ++c) {
          const double v = ldo(M, o + k0 + lane);
          Lr[c] = c < lane ? v : 0.0;
          next_col(o, k0 + c);
        }
      } else {
#pragma unroll
        for (int c = 0; c < 64; ++c) {
          const bool ok = c < lane && lane < kb;
          const double v = ldo(M, ok ? o + k0 + lane : 0u);
          Lr[c] = ok ? v : 0.0;
          if (c + 1 < kb) next_col(o, k0 + c);
        }
      }
    };
    load_diag(0, min(64, nr));
    for (int b = 0; b < nb; ++b) {
      const int k0 = 64 * b, kb = min(64, nr - k0);
      if (b > 0) phase_a_fwd(k0);
      double tl = lane < kb ? t[k0 + lane] : 0.0;
      if (b > 0) tl -= (part[lane] + part[64 + lane]) + (part[128 + lane] + part[192 + lane]);
#pragma unroll
      for (int c = 0; c < 64; ++c) tl -= Lr[c] * bcast(tl, c);
      if (lane < kb) t[k0 + lane] = tl;
      if (b + 1 < nb) {
        load_a(k0 + 64, k0);
        load_diag(k0 + 64, min(64, nr - k0 - 64));
      }
      __syncthreads();
    }
  } else {
    for (int b = 0; b < nb; ++b) {
      const int k0 = 64 * b;
      if (b > 0) {
        phase_a_fwd(k0);
        const int kp = k0 - 64;
        for (int i = k0 + tid; i < nr; i += BG - 64) {  // rows k0 + 64 .. nr - 1
          double lv[64];
          unsigned o = col0(kp);
#pragma unroll
          for (int c = 0; c < 64; ++c) {
            lv[c] = ldo(M, o + i);
            next_col(o, kp + c);
          }
          double acc = 0.0;
#pragma unroll
          for (int c = 0; c < 64; ++c) acc += lv[c] * t[kp + c];
          t[i] -= acc;
        }
      }
      if (b + 1 < nb) load_a(k0 + 64, k0);
      __syncthreads();
    }
  }
  for (int k = tid; k < nr; k += BG) t[k] /= ldo(M, (unsigned)colbase(k, nr));
  __syncthreads();
  // L' x = y, blocks from the last, right-looking: once x_b is final, every
  // column j < 64b takes the sum over block b's rows i of L(i, j) x_i, lanes
  // along the rows (coalesced) and a butterfly sum per 16 columns.  Per block:
  //   A(b), b < nb - 1: the columns of block b take the contribution of block
  //     b + 1, each wave 16 of them (operands loaded during B(b + 1));
  //   B(b): wave 0 runs the recurrence of the diagonal block; meanwhile waves
  //     1..3 apply block b + 1 to the columns left of block b, 16-column chunks
  //     dealt round-robin.  Then the operands of A(b - 1) and wave 0's next
  //     diagonal block are loaded.
  auto phase_a_bwd = [&](int k0) {
    const int i = k0 + 64 + lane;
    const double xv = i < nr ? t[i] : 0.0;
    double pr[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) pr[u] = La[u] * xv;
    const double s = wave_sum16(pr);
    if ((lane & 3) == 0) part[16 * wv + (lane >> 2)] = s;
    __syncthreads();
  };
  if (wv == 0) {
    double Lc[64];  // Lc[c] = L(k0 + c, k0 + lane), lane < c < kb
    auto load_diag = [&](int k0, int kb) {
      if (kb == 64) {
        const unsigned cl = lx(k0, k0 + lane);  // L(k0 + c, k0 + lane) = M[cl + c]
#pragma unroll
        for (int c = 0; c < 64; ++c) {
          const double v = ldo(M, cl + c);
          Lc[c] = c > lane ? v : 0.0;
        }
      } else {
#pragma unroll
        for (int c = 0; c < 64; ++c) {
          const bool ok = c > lane && c < kb;
          const double v = ldo(M, ok ? lx(k0 + c, k0 + lane) : 0u);
          Lc[c] = ok ? v : 0.0;
        }
      }
    };
    load_diag(64 * (nb - 1), nr - 64 * (nb - 1));
    for (int b = nb - 1; b >= 0; --b) {
      const int k0 = 64 * b, kb = min(64, nr - k0);
      if (b < nb - 1) phase_a_bwd(k0);
      double tl = lane < kb ? t[k0 + lane] : 0.0;
      if (b < nb - 1) tl -= part[lane];
#pragma unroll
      for (int c = 63; c > 0; --c) tl -= Lc[c] * bcast(tl, c);
      if (lane < kb) t[k0 + lane] = tl;
      if (b > 0) {
        load_a(k0, k0 - 64);
        load_diag(k0 - 64, 64);
      }
      __syncthreads();
    }
  } else {
    for (int b = nb - 1; b >= 0; --b) {
      const int k0 = 64 * b;
      if (b < nb - 1) {
        phase_a_bwd(k0);
        const int i = k0 + 64 + lane;
        const bool okr = i < nr;
        const double xv = okr ? t[i] : 0.0;
        const unsigned ir = okr ? (unsigned)i : 0u;  // masked rows read column heads
        // 16-column chunks dealt round-robin (stride 48), two register sets:
        // each chunk's loads are issued a chunk ahead of its butterfly sum
        auto load16 = [&](int j0, double (&v)[16]) {
          unsigned o = col0(j0);
#pragma unroll
          for (int u = 0; u < 16; ++u) {
            v[u] = ldo(M, o + ir);
            next_col(o, j0 + u);
          }
        };
        auto apply16 = [&](int j0, const double (&v)[16]) {
          double pr[16];
#pragma unroll
          for (int u = 0; u < 16; ++u) pr[u] = okr ? v[u] * xv : 0.0;
          const double s = wave_sum16(pr);
          if ((lane & 3) == 0) t[j0 + (lane >> 2)] -= s;
        };
        double va[16], vb[16];
        int j0 = 16 * (wv - 1);
        if (j0 < k0) load16(j0, va);
        for (; j0 < k0; j0 += 96) {
          if (j0 + 48 < k0) load16(j0 + 48, vb);
          apply16(j0, va);
          if (j0 + 96 < k0) load16(j0 + 96, va);
          if (j0 + 48 < k0) apply16(j0 + 48, vb);
        }
      }
      if (b > 0) load_a(k0, k0 - 64);
      __syncthreads();
    }
  }
  const int* __restrict__ arp = A.ar_ptr + p.arp;
  const int* __restrict__ ars = A.ar_src;
  const int* __restrict__ arc = A.ar_col;
  for (int j = tid; j < n; j += BG) x[j] = t[j];
  for (int kk = tid; kk < nk; kk += BG) x[n + kr[kk]] = t[n + kk];
  for (int i = tid; i < m; i += BG) {
    if (rk[i] >= 0) continue;
    const int e = arp[i];
    x[n + i] = (Av[ars[e]] * t[arc[e]] - rhs[n + i]) / (hsd[i] + da);
  }
  __syncthreads();
}

// SOC tangent correction of the adjoint system (the limit of the exact
// linearization arw(z) ds + arw(s) dz = 0 at a complementary pair): on the
// directions orthogonal to z1 the scaling eigenvalue beta^2 becomes s0/z0.
// tcorr = s0/z0 - beta^2 and tdir = (0, z1/|z1|) per block (0 where |z1| = 0
// or the scaling fell back to the identity).
__device__ void soc_tangent(const PV& p, const double* sv, const double* zv, double* tc, double* td) {
  for (int k = threadIdx.x; k < p.nsoc; k += BG) {
    const int o = p.C.soc_off[k], d = p.C.soc_dim[k];
    double zn = 0.0;
    for (int j = 1; j < d; ++j) zn += zv[o + j] * zv[o + j];
    zn = sqrt(zn);
    const bool ok = zn > 0.0 && zv[o] > 0.0 && !p.C.ident[k];
    td[o] = 0.0;
    for (int j = 1; j < d; ++j) td[o + j] = ok ? zv[o + j] / zn : 0.0;
    tc[k] = ok ? sv[o] / zv[o] - p.C.beta[k] * p.C.beta[k] : 0.0;
  }
  __syncthreads();
}
// out -= tcorr_k P_T x on every SOC block
__device__ void soc_tangent_apply(const PV& p, const double* x, double* out) {
  for (int k = threadIdx.x; k < p.nsoc; k += BG) {
    const int o = p.C.soc_off[k], d = p.C.soc_dim[k];
    const double c = p.tcorr[k];
    if (c == 0.0) continue;
    double ux = 0.0;
    for (int j = 1; j < d; ++j) ux += p.tdir[o + j] * x[o + j];
    for (int j = 1; j < d; ++j) out[o + j] -= c * (x[o + j] - p.tdir[o + j] * ux);
  }
  __syncthreads();
}

// res = rhs - K x with K the statically regularized KKT matrix (the refinement target)
// Two passes: the first computes the variable rows outright (P x and A'y per
// row) and, independently, A x and -H y per constraint row (the loads of all
// three overlap); the second combines the constraint rows.
__device__ void kkt_residual(CArgs& A, PV& p, const double* rhs, const double* x, double* res) {
  const long long t0 = wall_clock64();
  const int n = uni(p.n), m = uni(p.m), tid = threadIdx.x;
  const double sp = A.sreg_p, sa = A.sreg_a;
  const int* __restrict__ pfp = A.pf_ptr + uni(p.pfp);
  const int* __restrict__ acp = A.ac_ptr + uni(p.acp);
  const int* __restrict__ arp = A.ar_ptr + uni(p.arp);
  double* wm = p.wm;
  double* wm2 = p.wm2;
  for (int j = tid; j < n; j += BG) {
    double px = 0.0, aty = 0.0;
    for (int e = pfp[j]; e < pfp[j + 1]; ++e) px += A.Pv[A.pf_src[e]] * x[A.pf_col[e]];
    for (int e = acp[j]; e < acp[j + 1]; ++e) aty += A.Av[e] * x[n + A.ac_row[e]];
    res[j] = rhs[j] - px - sp * x[j] - aty;
  }
  for (int i = tid; i < m; i += BG) {
    double ax = 0.0;
    for (int e = arp[i]; e < arp[i + 1]; ++e) ax += A.Av[A.ar_src[e]] * x[A.ar_col[e]];
    wm[i] = ax;
  }
  for (int i = tid; i < p.nz; i += BG) wm2[i] = -0.0;
  neg_hs_plus_stored_nb(p, x + n, wm2);  // -H y (each block read and written by one thread)
  __syncthreads();
  for (int i = tid; i < m; i += BG) res[n + i] = rhs[n + i] - wm[i] - wm2[i] + sa * x[n + i];
  __syncthreads();
  p.t_res += wall_clock64() - t0;
}

// escalateAndResolve (kkt_system.cpp:266-294)
__device__ bool escalate(CArgs& A, PV& p, const double* rhs, double* x, double* colk, double* red) {
  double ep = p.last_p, ea = p.last_a;
  for (int att = 0; att < 10; ++att) {
    ep = dmax(ep * 1000.0, A.dyn_delta);
    ea = dmax(ea * 1000.0, A.dyn_delta);
    if (!try_factor(A, p, ep, ea, colk, red)) continue;
    reduced_solve(A, p, rhs, x, colk);
    if (count_nonfinite(p.N, x, red) == 0) {
      p.last_p = ep, p.last_a = ea;
      return true;
    }
  }
  return false;
}

// KktSystem::solve (kkt_system.cpp:296-311)
// tol: relative residual target; steps past min_steps stop once a step no
// longer halves the residual
__device__ double kkt_solve(CArgs& A, PV& p, const double* rhs, double* x, double* colk, double* red,
                           double tol, int min_steps) {
  const long long t_kkt0 = wall_clock64();
  ++p.nkkt;
  reduced_solve(A, p, rhs, x, colk);
  const double nonfin = count_nonfinite(p.N, x, red);
  kkt_residual(A, p, rhs, x, p.res);
  double v[2] = {0.0, 0.0};
  for (int i = threadIdx.x; i < p.N; i += BG) v[0] += rhs[i] * rhs[i], v[1] += p.res[i] * p.res[i];
  gsum<2>(v, red);
  const double rn = dmax(sqrt(v[0]), 1e-30);
  double rr = v[1];
  if (nonfin != 0.0) {
    escalate(A, p, rhs, x, colk, red);
    kkt_residual(A, p, rhs, x, p.res);
    double r = 0.0;
    for (int i = threadIdx.x; i < p.N; i += BG) r += p.res[i] * p.res[i];
    rr = gsum1(r, red);
  }
  double rel = sqrt(rr) / rn;
  // The reduced factorization is a weaker preconditioner than a full LDL^T
  // once (Hs + eI)^-1 has entries near 1/e: refinement then continues past
  // refine_max_iter (up to BATCH_REFINE_MAX steps) while it still reduces the
  // residual, so each solve meets its target (refine_tol for the IPM's own
  // solves, as the reference's solves do).
  const int rmax = A.refine_max_iter > 0 ? max(A.refine_max_iter, BATCH_REFINE_MAX) : 0;
  for (int it = 0; it < rmax && rel > tol; ++it) {
    reduced_solve(A, p, p.res, p.dsol, colk);
    vaxpy(p.N, 1.0, p.dsol, x);
    kkt_residual(A, p, rhs, x, p.res);
    double r = 0.0;
    for (int i = threadIdx.x; i < p.N; i += BG) r += p.res[i] * p.res[i];
    const double rel_new = sqrt(gsum1(r, red)) / rn;
    const bool stalled = it >= min_steps && rel_new > 0.5 * rel;
    rel = rel_new;
    if (stalled) break;
  }
  p.t_kkt_last = wall_clock64() - t_kkt0;
  return rel;
}

// ---------------------------------------------------------------------------
// IPM pieces (solver_impl.cpp)
// ---------------------------------------------------------------------------
__device__ void shift_cold(const PV& p, double* v, double* red) {  // solver_impl.cpp:104-116
  double mn, ps;
  margins(p, v, &mn, &ps, red);
  const int deg = p.nn + p.nsoc;
  const double target = dmax(1.0, 0.1 * ps / (double)(deg > 1 ? deg : 1));
  if (mn <= 0.0) {
    unit_shift(p, v, -mn);
    unit_shift(p, v, target);
  } else if (mn < target) {
    unit_shift(p, v, target - mn);
  }
}
__device__ void ensure_interior_warm(const PV& p, double* v, double* red) {  // solver_impl.cpp:118-124
  double mn, ps;
  margins(p, v, &mn, &ps, red);
  if (mn <= 1e-8) unit_shift(p, v, -mn + 1e-6);
}

__device__ bool initial_point(CArgs& A, PV& p, double* colk, double* red) {  // solver_impl.cpp:126-158
  const int n = p.n, m = p.m;
  // scaling at the identity element e
  double* e = p.dsk;
  for (int i = threadIdx.x; i < m; i += BG) e[i] = (i < p.nz) ? 0.0 : (i < p.nz + p.nn ? 1.0 : 0.0);
  __syncthreads();
  for (int k = threadIdx.x; k < p.nsoc; k += BG) e[p.C.soc_off[k]] = 1.0;
  __syncthreads();
  cone_scaling(p, e, e);
  if (!factor_retry(A, p, colk, red)) return false;
  if (p.pnnz == 0) {
    vfill(n, 0.0, p.rhs);
    vcopy(m, p.b, p.rhs + n);
    kkt_solve(A, p, p.rhs, p.sol, colk, red, A.refine_tol, A.refine_max_iter);
    vcopy(n, p.sol, p.x);
    vlin2(m, -1.0, p.sol + n, 0.0, p.sol + n, p.s);
    vlin2(n, -1.0, p.q, 0.0, p.q, p.rhs);
    vfill(m, 0.0, p.rhs + n);
    kkt_solve(A, p, p.rhs, p.sol, colk, red, A.refine_tol, A.refine_max_iter);
    vcopy(m, p.sol + n, p.z);
  } else {
    vlin2(n, -1.0, p.q, 0.0, p.q, p.rhs);
    vcopy(m, p.b, p.rhs + n);
    kkt_solve(A, p, p.rhs, p.sol, colk, red, A.refine_tol, A.refine_max_iter);
    vcopy(n, p.sol, p.x);
    vcopy(m, p.sol + n, p.z);
    vlin2(m, -1.0, p.sol + n, 0.0, p.sol + n, p.s);
  }
  vfill(p.nz, 0.0, p.s);
  shift_cold(p, p.s, red);
  shift_cold(p, p.z, red);
  p.tau = 1.0;
  p.kappa = 1.0;
  return count_nonfinite(n, p.x, red) + count_nonfinite(m, p.s, red) + count_nonfinite(m, p.z, red) == 0;
}

__device__ bool constant_solve(CArgs& A, PV& p, double* colk, double* red) {  // solver_impl.cpp:165-175
  vlin2(p.n, -1.0, p.q, 0.0, p.q, p.rhs);
  vcopy(p.m, p.b, p.rhs + p.n);
  kkt_solve(A, p, p.rhs, p.sol, colk, red, A.refine_tol, A.refine_max_iter);
  vcopy(p.n, p.sol, p.x1);
  vcopy(p.m, p.sol + p.n, p.z1);
  pmul(A, p, p.x1, p.Px1);
  return count_nonfinite(p.n, p.x1, red) + count_nonfinite(p.m, p.z1, red) == 0;
}

__device__ void residuals(CArgs& A, PV& p, double* red) {  // solver_impl.cpp:177-186
  const int n = p.n, m = p.m;
  pmul(A, p, p.x, p.Px);
  atmul(A, p, p.z, p.wn);
  vlin3(n, -1.0, p.wn, -1.0, p.Px, -p.tau, p.q, p.rx);
  amul(A, p, p.x, p.wm);
  vlin3(m, 1.0, p.wm, 1.0, p.s, -p.tau, p.b, p.rz);
  double v[6] = {0, 0, 0, 0, 0, 0};
  for (int i = threadIdx.x; i < n; i += BG) {
    v[0] += p.q[i] * p.x[i];
    v[3] += p.x[i] * p.Px[i];
    v[4] += p.rx[i] * p.rx[i];
  }
  for (int i = threadIdx.x; i < m; i += BG) {
    v[1] += p.b[i] * p.z[i];
    v[2] += p.s[i] * p.z[i];
    v[5] += p.rz[i] * p.rz[i];
  }
  gsum<6>(v, red);
  p.dqx = v[0], p.dbz = v[1], p.dsz = v[2], p.dxPx = v[3], p.nrx2 = v[4], p.nrz2 = v[5];
  p.rtau = p.dqx + p.dbz + p.kappa + p.dxPx / p.tau;
}

__device__ bool tau_elim(CArgs& A, PV& p, const double* xv, const double* zv, double dtau_rhs,
                         double dkappa_rhs, double* dtau, double* red) {
  const int n = p.n, m = p.m;
  pmul(A, p, xv, p.wn);
  vlin2(n, 1.0 / p.tau, p.x, -1.0, p.x1, p.wn2);
  pmul(A, p, p.wn2, p.tmp);
  double v[7] = {0, 0, 0, 0, 0, 0, 0};
  for (int i = threadIdx.x; i < n; i += BG) {
    v[0] += p.q[i] * xv[i];
    v[2] += p.x[i] * p.wn[i];
    v[3] += p.q[i] * p.x1[i];
    v[5] += p.wn2[i] * p.tmp[i];
    v[6] += p.x1[i] * p.Px1[i];
  }
  for (int i = threadIdx.x; i < m; i += BG) {
    v[1] += p.b[i] * zv[i];
    v[4] += p.b[i] * p.z1[i];
  }
  gsum<7>(v, red);
  const double num = dtau_rhs - dkappa_rhs / p.tau + v[0] + v[1] + 2.0 * (v[2] / p.tau);
  const double den = p.kappa / p.tau - v[3] - v[4] + v[5] - v[6];
  if (den == 0.0 || !isfinite(num) || !isfinite(den)) return false;
  *dtau = num / den;
  return true;
}

__device__ bool affine_step(CArgs& A, PV& p, double* colk, double* red) {  // solver_impl.cpp:192-231
  const int n = p.n, m = p.m;
  vcopy(n, p.rx, p.rhs);
  vlin2(m, 1.0, p.s, -1.0, p.rz, p.rhs + n);
  kkt_solve(A, p, p.rhs, p.sol, colk, red, A.refine_tol, A.refine_max_iter);
  if (count_nonfinite(p.N, p.sol, red) != 0) return false;
  const double dtau_rhs = p.rtau, dkappa_rhs = p.tau * p.kappa;
  if (!tau_elim(A, p, p.sol, p.sol + n, dtau_rhs, dkappa_rhs, &p.dtaua, red)) return false;
  vlin2(n, 1.0, p.sol, p.dtaua, p.x1, p.dxa);
  vlin2(m, 1.0, p.sol + n, p.dtaua, p.z1, p.dza);
  neg_hs_plus(p, p.dza, p.s, p.dsa);
  p.dkappaa = -(dkappa_rhs + p.kappa * p.dtaua) / p.tau;
  return isfinite(p.dtaua) && isfinite(p.dkappaa) &&
         count_nonfinite(n, p.dxa, red) + count_nonfinite(m, p.dza, red) + count_nonfinite(m, p.dsa, red) == 0;
}

__device__ bool combined_step(CArgs& A, PV& p, double sigma, double mu, double* colk, double* red) {
  const int n = p.n, m = p.m;  // solver_impl.cpp:233-286
  combined_rhs(p, p.z, p.dsa, p.dza, p.rz, sigma * mu, 1.0 - sigma, p.dsk, p.rhs + n);
  vlin2(n, 1.0 - sigma, p.rx, 0.0, p.rx, p.rhs);
  kkt_solve(A, p, p.rhs, p.sol, colk, red, A.refine_tol, A.refine_max_iter);
  if (count_nonfinite(p.N, p.sol, red) != 0) return false;
  const double dtau_rhs = (1.0 - sigma) * p.rtau;
  const double dkappa_rhs = p.tau * p.kappa + p.dtaua * p.dkappaa - sigma * mu;
  if (!tau_elim(A, p, p.sol, p.sol + n, dtau_rhs, dkappa_rhs, &p.dtau, red)) return false;
  vlin2(n, 1.0, p.sol, p.dtau, p.x1, p.dx);
  vlin2(m, 1.0, p.sol + n, p.dtau, p.z1, p.dz);
  neg_hs_plus(p, p.dz, p.dsk, p.ds);
  p.dkappa = -(dkappa_rhs + p.kappa * p.dtau) / p.tau;
  return isfinite(p.dtau) && isfinite(p.dkappa) &&
         count_nonfinite(n, p.dx, red) + count_nonfinite(m, p.dz, red) + count_nonfinite(m, p.ds, red) == 0;
}

__device__ double step_length(const PV& p, const double* dsv, const double* dzv, double dtau, double dkappa,
                              double* red) {  // solver_impl.cpp:288-296
  double alpha = 1.0;
  if (dtau < 0) alpha = dmin(alpha, -p.tau / dtau);
  if (dkappa < 0) alpha = dmin(alpha, -p.kappa / dkappa);
  alpha = dmin(alpha, cone_max_step(p, p.z, dzv, red));
  alpha = dmin(alpha, cone_max_step(p, p.s, dsv, red));
  return alpha;
}

// ---------------------------------------------------------------------------
// one problem, start to finish
// ---------------------------------------------------------------------------
// Per-problem view: sizes, bases, LDS/global placement of the factor and of
// the vectors.  global_M forces the factor into this problem's global slot
// (the backward pass reads the factor its forward pass left there).
__device__ void setup_view(CArgs& A, const BatchDesc& D, PV& p, double* lds, bool global_M) {
  p.n = D.n, p.m = D.m, p.nz = D.nz, p.nn = D.nn, p.nsoc = D.nsoc, p.nk = D.nk;
  p.N = D.n + D.m, p.nr = D.nr, p.pnnz = D.pnnz;
  p.co = D.co, p.so = D.so, p.ko = D.ko, p.pfp = D.pfp, p.acp = D.acp, p.arp = D.arp;
  p.cur_a = 0.0;
  p.nfact = p.nsolve = p.nkkt = 0;
  p.t_fact = p.t_solve = p.t_res = p.t_kkt_last = 0;
  p.ec = D.ec;
  p.q = A.q + D.xo;
  p.b = A.b + D.co;
  const int n = D.n, m = D.m, N = n + m;
  // LDS placement: the vectors (touched by every step) before the factor M
  // (touched once per factorization and triangular solve); what does not fit
  // lives in this problem's global workspace
  const long V = 14L * n + 19L * m, PM = packed_size(D.nr), cap = A.lds_doubles;
  double* w;
  p.stage = nullptr;
  p.stage_cap = 0;
  if (V + PM <= cap && !global_M) {
    p.M = lds;
    w = lds + PM;
  } else if (V + (long)D.nr * PS <= cap) {  // vectors in LDS and room to stage every panel
    w = lds;
    p.M = A.Mg + D.mofs;
    p.stage = lds + V;
    p.stage_cap = cap - V;
  } else if (PM <= cap && !global_M) {
    p.M = lds;
    w = A.W + D.wofs;
  } else {
    p.M = A.Mg + D.mofs;
    w = A.W + D.wofs;
    p.stage = lds;
    p.stage_cap = cap;
  }
  p.m_lds = p.M == lds;
  p.v_lds = w == lds || w == lds + PM;
  p.x = w; w += n;
  p.Px = w; w += n;
  p.rx = w; w += n;
  p.x1 = w; w += n;
  p.Px1 = w; w += n;
  p.dxa = w; w += n;
  p.dx = w; w += n;
  p.wn = w; w += n;
  p.wn2 = w; w += n;
  p.s = w; w += m;
  p.z = w; w += m;
  p.rz = w; w += m;
  p.z1 = w; w += m;
  p.dza = w; w += m;
  p.dsa = w; w += m;
  p.dz = w; w += m;
  p.ds = w; w += m;
  p.dsk = w; w += m;
  p.wm = w; w += m;
  p.wm2 = w; w += m;
  double* cv = w; w += m;
  double* cw = w; w += m;
  double* chs = w; w += m;
  p.rhs = w; w += N;
  p.sol = w; w += N;
  p.res = w; w += N;
  p.tmp = w; w += N;
  p.dsol = w; w += N;
  p.C.zero_off = 0, p.C.zero_dim = D.nz;
  p.C.nn_off = D.nz, p.C.nn_dim = D.nn;
  p.C.nsoc = D.nsoc;
  p.C.soc_off = A.soc_off + D.so;
  p.C.soc_dim = A.soc_dim + D.so;
  p.C.small_list = nullptr, p.C.nsmall = 0, p.C.large_list = nullptr, p.C.nlarge = 0;
  p.C.beta = A.beta + D.so;
  p.C.vv = A.vv + D.so;
  p.C.ident = A.ident + D.so;
  p.C.v = cv, p.C.w = cw, p.C.hsd = chs;
  p.tau = p.kappa = 1.0;
  p.last_p = p.last_a = 0.0;
  p.rtau = p.dqx = p.dbz = p.dsz = p.dxPx = p.nrx2 = p.nrz2 = 0.0;
  p.dtaua = p.dkappaa = p.dtau = p.dkappa = 0.0;
  p.tcorr = nullptr;
  p.tdir = nullptr;
}


__device__ void solve_problem(CArgs& A, int pid, double* lds, double* colk, double* red) {
  const long long t_start = wall_clock64();
  const BatchDesc D = A.desc[pid];
  PV p;
  setup_view(A, D, p, lds, false);
  const int n = D.n, m = D.m;
  int status = CX_NUMERICAL_ERROR, iterations = 0;
  bool ok = true;
  const bool use_warm = A.warm_start && A.have_warm[pid];
  if (use_warm) {
    vcopy(n, A.wx + D.xo, p.x);
    vcopy(m, A.ws + D.co, p.s);
    vcopy(m, A.wz + D.co, p.z);
    vfill(p.nz, 0.0, p.s);
    ensure_interior_warm(p, p.s, red);
    ensure_interior_warm(p, p.z, red);
  } else {
    ok = initial_point(A, p, colk, red);
  }
  if (ok) {
    status = CX_MAX_ITERATIONS;
    const int deg = p.nn + p.nsoc;
    for (int iter = 0; iter < A.max_iter; ++iter) {
      iterations = iter + 1;
      cone_scaling(p, p.s, p.z);
      if (!factor_retry(A, p, colk, red) || !constant_solve(A, p, colk, red)) {
        status = CX_NUMERICAL_ERROR;
        break;
      }
      residuals(A, p, red);
      const double mu_now = (p.dsz + p.tau * p.kappa) / (double)(deg + 1);
      const double nrx = sqrt(p.nrx2) / p.tau, nrz = sqrt(p.nrz2) / p.tau;
      const double pobj = (0.5 * p.dxPx / p.tau + p.dqx) / p.tau;
      const double dobj = (-0.5 * p.dxPx / p.tau - p.dbz) / p.tau;
      const double gap_abs = fabs(pobj - dobj);
      const double gap_tol = A.tol_gap_abs + A.tol_gap_rel * dmax(fabs(pobj), fabs(dobj));
      if (nrx < A.tol_feas && nrz < A.tol_feas && gap_abs < gap_tol && p.tau > 1e-6) {
        status = CX_SOLVED;
        break;
      }
      const double ratio = p.tau / dmax(p.kappa, 1e-30);  // solver_impl.cpp:327-340
      if (ratio <= A.tol_infeas) {
        double v[2] = {0.0, 0.0};
        for (int i = threadIdx.x; i < m; i += BG) v[0] += p.z[i] * p.z[i];
        for (int i = threadIdx.x; i < n; i += BG) v[1] += p.x[i] * p.x[i];
        gsum<2>(v, red);
        if (p.dbz < -A.tol_infeas * dmax(1.0, sqrt(v[0]))) {
          status = CX_PRIMAL_INFEASIBLE;
          break;
        }
        if (p.dqx < -A.tol_infeas * dmax(1.0, sqrt(v[1]))) {
          status = CX_DUAL_INFEASIBLE;
          break;
        }
      }
      if (!affine_step(A, p, colk, red)) {
        status = CX_NUMERICAL_ERROR;
        break;
      }
      const double aa = step_length(p, p.dsa, p.dza, p.dtaua, p.dkappaa, red);
      double saz = 0.0;
      for (int i = threadIdx.x; i < m; i += BG) saz += (p.s[i] + aa * p.dsa[i]) * (p.z[i] + aa * p.dza[i]);
      saz = gsum1(saz, red);
      const double tau_a = p.tau + aa * p.dtaua, kap_a = p.kappa + aa * p.dkappaa;
      const double mu_aff = (saz + tau_a * kap_a) / (double)(deg + 1);
      double sigma;
      if (mu_now > 0) {
        const double r = dmax(mu_aff / mu_now, 0.0);
        sigma = dmin(dmax(r * r * r, 0.0), 1.0);
      } else {
        sigma = 1.0;
      }
      if (!combined_step(A, p, sigma, mu_now, colk, red)) {
        status = CX_NUMERICAL_ERROR;
        break;
      }
      const double alpha = step_length(p, p.ds, p.dz, p.dtau, p.dkappa, red) * A.step_frac;
      if (!isfinite(alpha) || alpha <= 0.0) {
        status = CX_NUMERICAL_ERROR;
        break;
      }
      vaxpy(n, alpha, p.dx, p.x);
      vaxpy(m, alpha, p.ds, p.s);
      vaxpy(m, alpha, p.dz, p.z);
      p.tau += alpha * p.dtau;
      p.kappa += alpha * p.dkappa;
      const double sc = dmax(p.tau, p.kappa);
      if (sc > 1e8 || sc < 1e-8) {
        const double inv = 1.0 / sc;
        vscale(n, inv, p.x);
        vscale(m, inv, p.s);
        vscale(m, inv, p.z);
        p.tau *= inv;
        p.kappa *= inv;
      }
    }
  }
  // finalizeSolution (solver_impl.cpp:342-365)
  const bool inf = status == CX_PRIMAL_INFEASIBLE || status == CX_DUAL_INFEASIBLE;
  const double scale = inf ? 1.0 / dmax(p.kappa, 1e-30) : 1.0 / dmax(p.tau, 1e-30);
  const double* ed = A.ed + D.xo;
  const double* ee = A.ee + D.co;
  for (int i = threadIdx.x; i < n; i += BG) A.xo[D.xo + i] = p.x[i] * scale * ed[i];
  for (int i = threadIdx.x; i < m; i += BG) {
    A.zo[D.co + i] = p.z[i] * scale * ee[i] / p.ec;
    A.so[D.co + i] = p.s[i] * scale / ee[i];
  }
  if (status == CX_SOLVED && A.warm_start) {
    for (int i = threadIdx.x; i < n; i += BG) A.wx[D.xo + i] = p.x[i];
    for (int i = threadIdx.x; i < m; i += BG) A.ws[D.co + i] = p.s[i], A.wz[D.co + i] = p.z[i];
  }
  // differentiation: the adjoint matrix (final NT scaling, SOC tangent
  // correction) is factored here, at the end of the forward pass, and left in
  // the problem's global slot with the normalized optimum, so that the
  // backward pass is one KKT solve
  int dok = 0;
  if (A.diff && status == CX_SOLVED) {
    double* tc = A.tcorr + D.so;
    double* td = A.tdir + D.co;
    soc_tangent(p, p.s, p.z, tc, td);
    p.tcorr = tc;
    p.tdir = td;
    const long long t_adj0 = wall_clock64();
    dok = factor_retry(A, p, colk, red) ? 1 : 0;
    double* Mg = A.Mg + D.mofs;
    if (p.M != Mg) {
      const long tot = packed_size(p.nr);
      for (long e = threadIdx.x; e < tot; e += BG) Mg[e] = p.M[e];
    }
    __syncthreads();
    if (threadIdx.x == 0) A.info[(long)INFO_STRIDE * pid + 16] = (double)(wall_clock64() - t_adj0);
    const double it = 1.0 / dmax(p.tau, 1e-30);
    for (int i = threadIdx.x; i < n; i += BG) A.xh[D.xo + i] = p.x[i] * it;
    for (int i = threadIdx.x; i < m; i += BG) A.sh[D.co + i] = p.s[i] * it, A.zh[D.co + i] = p.z[i] * it;
  }
  if (threadIdx.x == 0) {
    if (A.diff) {
      A.diff_ok[pid] = dok;
      A.diff_a[pid] = p.cur_a;
    }
    const double pobj = (0.5 * p.dxPx / p.tau + p.dqx) / p.tau;
    double* inf8 = A.info + (long)INFO_STRIDE * pid;
    inf8[0] = sqrt(p.nrz2) / dmax(p.tau, 1e-30);  // primal residual
    inf8[1] = sqrt(p.nrx2) / dmax(p.tau, 1e-30);  // dual residual
    inf8[2] = p.dsz / (p.tau * p.tau) / p.ec;     // gap
    inf8[3] = (p.dsz + p.tau * p.kappa) / (double)(p.nn + p.nsoc + 1);
    inf8[4] = pobj / p.ec;
    inf8[5] = p.nfact;
    inf8[6] = p.nsolve;
    inf8[7] = (double)(wall_clock64() - t_start);  // constant-rate device clock ticks
    inf8[8] = (double)p.t_fact;
    inf8[9] = (double)p.t_solve;
    inf8[10] = (double)p.t_res;
    inf8[11] = p.nkkt;
    inf8[15] = (double)p.t_kkt_last;  // the final iteration's KKT solve
    if (!(A.diff && status == CX_SOLVED)) inf8[16] = 0.0;  // no adjoint factorization
    A.status[pid] = status;
    A.iters[pid] = iterations;
    if (status == CX_SOLVED && A.warm_start) A.have_warm[pid] = 1;
  }
  __syncthreads();
}

__global__ __launch_bounds__(BG) void k_batch_solve(BatchArgs) {
  CArgs& A = kernel_args();
  extern __shared__ double lds[];
  __shared__ double colk[COLK];  // Wt of the diagonal block, pivots, their reciprocals
  __shared__ double red[64];
  __shared__ int slot;
  for (;;) {
    if (threadIdx.x == 0) slot = atomicAdd(A.counter, 1);
    __syncthreads();
    const int k = slot;
    __syncthreads();
    if (k >= A.nprob) break;
    solve_problem(A, A.order[k], lds, colk, red);
  }
}

// Backward pass of one problem whose forward pass left the adjoint factor in
// its global slot: the cone scaling is recomputed from the stored optimum (it
// is invariant under the common 1/tau scaling), then
//   K [u; v] = [D wx; E wz / c - H (ws / E)]     (refined against the same K)
//   dL/dq = -c D u,  dL/db = E v,  dL/dA_ij = -(z_i u_j + v_i x_j) E_i D_j,
//   dL/dP_ij (upper) = -(u_i x_j + u_j x_i) c D_i D_j, diagonal -u_i x_i c D_i^2,
// with x, z the normalized optimum in the equilibrated space.
__device__ void backward_problem(CArgs& A, int pid, double* lds, double* colk, double* red) {
  const long long t_start = wall_clock64();
  const BatchDesc D = A.desc[pid];
  PV p;
  setup_view(A, D, p, lds, false);
  const int n = D.n, m = D.m;
  // the forward pass left the adjoint factor in the global slot: bring it into
  // LDS when this problem's placement keeps M there
  double* Mg = A.Mg + D.mofs;
  if (p.M != Mg && A.diff_ok[pid]) {
    const long tot = packed_size(p.nr);
    for (long e = threadIdx.x; e < tot; e += BG) p.M[e] = Mg[e];
    __syncthreads();
  }
  if (!A.diff_ok[pid]) {
    for (int j = threadIdx.x; j < n; j += BG) A.gq[D.xo + j] = 0.0;
    for (int i = threadIdx.x; i < m; i += BG) A.gb[D.co + i] = 0.0;
    __syncthreads();
    return;
  }
  const double* ed = A.ed + D.xo;
  const double* ee = A.ee + D.co;
  vcopy(n, A.xh + D.xo, p.x);
  vcopy(m, A.sh + D.co, p.s);
  vcopy(m, A.zh + D.co, p.z);
  cone_scaling(p, p.s, p.z);
  p.tcorr = A.tcorr + D.so;
  p.tdir = A.tdir + D.co;
  p.cur_a = A.diff_a[pid];
  for (int j = threadIdx.x; j < n; j += BG) p.rhs[j] = A.gwx[D.xo + j] * ed[j];
  for (int i = threadIdx.x; i < m; i += BG) {
    p.wm[i] = A.gws[D.co + i] / ee[i];
    p.wm2[i] = -A.gwz[D.co + i] * ee[i] / p.ec;
  }
  __syncthreads();
  neg_hs_plus(p, p.wm, p.wm2, p.rhs + n);
  soc_tangent_apply(p, p.wm, p.rhs + n);
  // the adjoint stops at its own target or at the first step that stalls
  kkt_solve(A, p, p.rhs, p.sol, colk, red, A.adj_tol, 1);
  const double* u = p.sol;
  const double* v = p.sol + n;
  // gradient assembly: one wave per column of A / row of P, lanes along its entries
  for (int j = threadIdx.x; j < n; j += BG) A.gq[D.xo + j] = -p.ec * ed[j] * u[j];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int j = wv; j < n; j += BG / 64) {
    const double uj = u[j], xj = p.x[j], dj = ed[j];
    for (int e = A.ac_ptr[p.acp + j] + lane; e < A.ac_ptr[p.acp + j + 1]; e += 64) {
      const int i = A.ac_row[e];
      A.gA[e] = -(p.z[i] * uj + v[i] * xj) * ee[i] * dj;
    }
    for (int e = A.pf_ptr[p.pfp + j] + lane; e < A.pf_ptr[p.pfp + j + 1]; e += 64) {
      const int c = A.pf_col[e];
      if (c < j) continue;  // each stored upper entry once, from its row
      A.gP[A.pf_src[e]] = (c == j) ? -uj * xj * p.ec * dj * dj
                                   : -(uj * p.x[c] + u[c] * xj) * p.ec * dj * ed[c];
    }
  }
  for (int i = threadIdx.x; i < m; i += BG) A.gb[D.co + i] = ee[i] * v[i];
  if (threadIdx.x == 0) {
    A.info[(long)INFO_STRIDE * pid + 12] = (double)(wall_clock64() - t_start);
    A.info[(long)INFO_STRIDE * pid + 13] = (double)(p.t_solve + p.t_res);
    A.info[(long)INFO_STRIDE * pid + 14] = p.nsolve;
  }
  __syncthreads();
}

__global__ __launch_bounds__(BG) void k_batch_backward(BatchArgs) {
  CArgs& A = kernel_args();
  extern __shared__ double lds[];
  __shared__ double colk[COLK];
  __shared__ double red[64];
  __shared__ int slot;
  for (;;) {
    if (threadIdx.x == 0) slot = atomicAdd(A.counter2, 1);
    __syncthreads();
    const int k = slot;
    __syncthreads();
    if (k >= A.nprob) break;
    backward_problem(A, A.order[k], lds, colk, red);
  }
}

// q_scaled = c D q_raw, b_scaled = E b_raw per problem; counter reset
__global__ void k_batch_prepare(int total_n, int total_m, const int* xprob, const BatchDesc* desc,
                                const double* qraw, const double* braw, const double* ed,
                                const double* ee, double* q, double* b, int* counter) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) counter[0] = 0;
  if (i < total_n) q[i] = qraw[i] * desc[xprob[i]].ec * ed[i];
  if (i < total_m) b[i] = braw[i] * ee[i];
}

// P / A values of an update: scaled = raw * (per-entry equilibration factor)
__global__ void k_batch_scale_values(long n, const double* __restrict__ raw, const double* __restrict__ scale,
                                     double* __restrict__ out) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = raw[i] * scale[i];
}

double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

bool valid_csc(int rows, int cols, const int* p, const int* i) {
  if (!p || p[0] != 0) return false;
  for (int c = 0; c < cols; ++c) {
    if (p[c + 1] < p[c]) return false;
    for (int k = p[c]; k < p[c + 1]; ++k) {
      if (!i || i[k] < 0 || i[k] >= rows) return false;
      if (k > p[c] && i[k] <= i[k - 1]) return false;
    }
  }
  return true;
}

constexpr int LDS_MAX_BYTES = 160 * 1024 - COLK * 8 - 64 * 8 - 64;  // minus the static LDS of the kernel

}  // namespace

BatchSolver::BatchSolver() {
  CX_HIP(hipStreamCreateWithFlags(&st_, hipStreamNonBlocking));
  CX_HIP(hipEventCreate(&ev0_));
  CX_HIP(hipEventCreate(&ev1_));
  cx_default_settings(&set_);
}

BatchSolver::~BatchSolver() {
  if (gexec_) (void)hipGraphExecDestroy(gexec_);
  if (graph_) (void)hipGraphDestroy(graph_);
  if (stage_q_) (void)hipHostFree(stage_q_);
  if (stage_b_) (void)hipHostFree(stage_b_);
  if (stage_P_) (void)hipHostFree(stage_P_);
  if (stage_A_) (void)hipHostFree(stage_A_);
  if (ev0_) (void)hipEventDestroy(ev0_);
  if (ev1_) (void)hipEventDestroy(ev1_);
  if (st_) (void)hipStreamDestroy(st_);
}

void BatchSolver::set_settings(const cx_settings& st) {
  set_ = st;
  if (gexec_) (void)hipGraphExecDestroy(gexec_), gexec_ = nullptr;
  if (graph_) (void)hipGraphDestroy(graph_), graph_ = nullptr;
}

bool BatchSolver::setup(int nprob, const int* n, const int* m, const int* const* Pp, const int* const* Pi,
                        const double* const* Px, const double* const* q, const int* const* Ap,
                        const int* const* Ai, const double* const* Ax, const double* const* b,
                        const int* zero_dim, const int* nonneg_dim, const int* nsoc,
                        const int* const* soc_dims, const cx_settings& st) {
  ready_ = false;
  if (gexec_) (void)hipGraphExecDestroy(gexec_), gexec_ = nullptr;
  if (graph_) (void)hipGraphDestroy(graph_), graph_ = nullptr;
  if (nprob <= 0) return false;
  set_ = st;
  np_ = nprob;
  n_.assign(n, n + nprob);
  m_.assign(m, m + nprob);
  nz_.assign(zero_dim, zero_dim + nprob);
  nn_.assign(nonneg_dim, nonneg_dim + nprob);
  socd_.assign(nprob, {});
  // validation (solver_impl.cpp:11-16 per problem)
  for (int k = 0; k < nprob; ++k) {
    if (n[k] < 0 || m[k] < 0 || zero_dim[k] < 0 || nonneg_dim[k] < 0 || nsoc[k] < 0) return false;
    long tot = (long)zero_dim[k] + nonneg_dim[k];
    for (int t = 0; t < nsoc[k]; ++t) {
      if (soc_dims[k][t] < 1) return false;
      tot += soc_dims[k][t];
      socd_[k].push_back(soc_dims[k][t]);
    }
    if (tot != m[k]) return false;
    if (!valid_csc(n[k], n[k], Pp[k], Pi[k]) || !valid_csc(m[k], n[k], Ap[k], Ai[k])) return false;
    if ((n[k] > 0 && !q[k]) || (m[k] > 0 && !b[k])) return false;
  }
  // host structure
  desc_h_.assign(nprob, BatchDesc{});
  std::vector<int> pf_ptr, pf_col, pf_src, ac_ptr, ac_row, ar_ptr, ar_col, ar_src, soc_off, soc_dim;
  std::vector<int> rk, kr, xprob;
  std::vector<double> Pv, Av, qv, bv, sP, sA;
  Pp_h_.assign(nprob, {}), Pi_h_.assign(nprob, {}), Ap_h_.assign(nprob, {}), Ai_h_.assign(nprob, {});
  Praw_h_.clear(), Araw_h_.clear();
  ed_h_.clear();
  ee_h_.clear();
  long wofs = 0, mofs = 0;
  int xo = 0, co = 0;
  max_packed_ = 0;
  max_lds_need_ = 0;
  pv_base_.clear();
  av_base_.clear();
  av_len_.clear();
  for (int k = 0; k < nprob; ++k) {
    const int nk = n[k], mk = m[k], nzk = zero_dim[k], nnk = nonneg_dim[k], nsk = nsoc[k];
    BatchDesc& D = desc_h_[k];
    D.n = nk, D.m = mk, D.nz = nzk, D.nn = nnk, D.nsoc = nsk;
    D.xo = xo, D.co = co, D.so = (int)soc_off.size(), D.ko = (int)kr.size();
    const int nP = Pp[k][nk], nA = Ap[k][nk];
    // equilibration
    std::vector<double> pv(Px[k], Px[k] + nP), av(Ax[k], Ax[k] + nA);
    std::vector<int> so_k(nsk);
    {
      int off = nzk + nnk;
      for (int t = 0; t < nsk; ++t) so_k[t] = off, off += socd_[k][t];
    }
    std::vector<double> edk, eek;
    double eck = 1.0;
    ruiz_equilibrate(nk, mk, Pp[k], Pi[k], pv, Ap[k], Ai[k], av, so_k, socd_[k], st, edk, eek, eck);
    D.ec = eck;
    // setup's pattern and the per-entry scale of later value updates
    // (Equilibration::scaleP / scaleA, equilibration.cpp:102-115)
    Pp_h_[k].assign(Pp[k], Pp[k] + nk + 1), Pi_h_[k].assign(Pi[k], Pi[k] + nP);
    Ap_h_[k].assign(Ap[k], Ap[k] + nk + 1), Ai_h_[k].assign(Ai[k], Ai[k] + nA);
    Praw_h_.insert(Praw_h_.end(), Px[k], Px[k] + nP), Araw_h_.insert(Araw_h_.end(), Ax[k], Ax[k] + nA);
    for (int c = 0; c < nk; ++c) {
      for (int e = Pp[k][c]; e < Pp[k][c + 1]; ++e) sP.push_back(eck * edk[Pi[k][e]] * edk[c]);
      for (int e = Ap[k][c]; e < Ap[k][c + 1]; ++e) sA.push_back(eek[Ai[k][e]] * edk[c]);
    }
    ed_h_.insert(ed_h_.end(), edk.begin(), edk.end());
    ee_h_.insert(ee_h_.end(), eek.begin(), eek.end());
    const int pbase = (int)Pv.size(), abase = (int)Av.size();
    pv_base_.push_back(pbase);
    av_base_.push_back(abase);
    av_len_.push_back(nA);
    Pv.insert(Pv.end(), pv.begin(), pv.end());
    Av.insert(Av.end(), av.begin(), av.end());
    D.pnnz = nP;
    qv.insert(qv.end(), q[k], q[k] + nk);
    bv.insert(bv.end(), b[k], b[k] + mk);
    for (int i = 0; i < nk; ++i) xprob.push_back(k);
    // full symmetric P by rows (upper storage; entries below the diagonal ignored)
    {
      std::vector<std::vector<std::pair<int, int>>> rows(nk);
      for (int c = 0; c < nk; ++c)
        for (int e = Pp[k][c]; e < Pp[k][c + 1]; ++e) {
          const int r = Pi[k][e];
          if (r > c) continue;
          rows[r].push_back({c, pbase + e});
          if (r != c) rows[c].push_back({r, pbase + e});
        }
      D.pfp = (int)pf_ptr.size();
      for (int r = 0; r < nk; ++r) {
        pf_ptr.push_back((int)pf_col.size());
        for (auto& ce : rows[r]) pf_col.push_back(ce.first), pf_src.push_back(ce.second);
      }
      pf_ptr.push_back((int)pf_col.size());
    }
    // A by columns (entry e <-> Av[abase + e]) and by rows
    D.acp = (int)ac_ptr.size();
    for (int c = 0; c <= nk; ++c) ac_ptr.push_back(abase + Ap[k][c]);
    for (int e = 0; e < nA; ++e) ac_row.push_back(Ai[k][e]);
    // ac_row is indexed globally by the entry: keep it aligned with Av
    {
      std::vector<std::vector<std::pair<int, int>>> rows(mk);
      for (int c = 0; c < nk; ++c)
        for (int e = Ap[k][c]; e < Ap[k][c + 1]; ++e) rows[Ai[k][e]].push_back({c, abase + e});
      D.arp = (int)ar_ptr.size();
      for (int r = 0; r < mk; ++r) {
        ar_ptr.push_back((int)ar_col.size());
        for (auto& ce : rows[r]) ar_col.push_back(ce.first), ar_src.push_back(ce.second);
      }
      ar_ptr.push_back((int)ar_col.size());
      // bound rows (nonnegative rows with one entry) are folded into the P
      // block; every other row is kept in M after the variables
      int nkept = 0;
      for (int i = 0; i < mk; ++i) {
        const bool bound = i >= nzk && i < nzk + nnk && rows[i].size() == 1;
        if (bound) {
          rk.push_back(-1);
        } else {
          rk.push_back(nkept++);
          kr.push_back(i);
        }
      }
      D.nk = nkept;
      D.nr = nk + nkept;
    }
    // the device addresses a problem's packed factor with 32-bit offsets
    if (D.nr > 65535) throw std::runtime_error("batch: reduced KKT order above 65535 (use the single-problem solver)");
    for (int t = 0; t < nsk; ++t) soc_off.push_back(so_k[t]), soc_dim.push_back(socd_[k][t]);
    // workspace: 9 n-vectors, 14 m-vectors, 5 N-vectors
    D.wofs = wofs;
    wofs += 9L * nk + 14L * mk + 5L * (nk + mk);
    const long pk = (long)D.nr * (D.nr + 1) / 2;
    max_packed_ = std::max<long>(max_packed_, pk);
    max_lds_need_ = std::max<long>(max_lds_need_, pk + 14L * nk + 19L * mk);
    D.mofs = mofs;
    mofs += pk;
    xo += nk;
    co += mk;
  }
  sum_n_ = xo, sum_m_ = co;
  nnzP_total_ = (long)Pv.size(), nnzA_total_ = (long)Av.size(), nsoc_total_ = soc_off.size();
  if (ac_row.size() != Av.size()) throw std::runtime_error("batch: A entry maps out of step");
  // problems largest first
  std::vector<int> order(nprob);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b2) { return desc_h_[a].nr > desc_h_[b2].nr; });

  auto up_i = [&](DevBuf<int>& d, const std::vector<int>& v) { d.upload(v.empty() ? std::vector<int>{0} : v, st_); };
  auto up_d = [&](DevBuf<double>& d, const std::vector<double>& v) {
    d.upload(v.empty() ? std::vector<double>{0.0} : v, st_);
  };
  desc_.upload(desc_h_, st_);
  up_i(order_, order);
  counter_.alloc(2);  // [0] forward, [1] backward work counters
  up_i(pf_ptr_, pf_ptr), up_i(pf_col_, pf_col), up_i(pf_src_, pf_src);
  up_i(ac_ptr_, ac_ptr), up_i(ac_row_, ac_row);
  up_i(ar_ptr_, ar_ptr), up_i(ar_col_, ar_col), up_i(ar_src_, ar_src);
  up_i(soc_off_, soc_off), up_i(soc_dim_, soc_dim);
  up_i(rk_, rk), up_i(kr_, kr);
  up_d(Pv_, Pv), up_d(Av_, Av), up_d(qraw_, qv), up_d(braw_, bv), up_d(ed_, ed_h_), up_d(ee_, ee_h_);
  if (sP.size() != Pv.size() || sA.size() != Av.size()) throw std::runtime_error("batch: value scales out of step");
  up_d(sP_, sP), up_d(sA_, sA);
  P_dirty_ = A_dirty_ = false;
  if (stage_P_) (void)hipHostFree(stage_P_), stage_P_ = nullptr;
  if (stage_A_) (void)hipHostFree(stage_A_), stage_A_ = nullptr;
  q_.alloc(std::max<long>(1, sum_n_));
  b_.alloc(std::max<long>(1, sum_m_));
  {
    DevBuf<int> xp;
    up_i(xp, xprob);
    hipLaunchKernelGGL(k_batch_prepare, dim3(ceil_div(std::max(sum_n_, sum_m_) + 1, 256)), dim3(256), 0, st_,
                       (int)sum_n_, (int)sum_m_, xp.p, desc_.p, qraw_.p, braw_.p, ed_.p, ee_.p, q_.p, b_.p,
                       counter_.p);
    CX_LAUNCH_CHECK();
    CX_HIP(hipStreamSynchronize(st_));
    xprob_.upload(xprob, st_);
  }
  W_.alloc(std::max<long>(1, wofs));
  W_.zero(st_);
  Mg_.alloc(std::max<long>(1, mofs));
  const size_t nsoc_tot = std::max<size_t>(1, soc_off.size());
  beta_.alloc(nsoc_tot), vv_.alloc(nsoc_tot), ident_.alloc(nsoc_tot);
  wxs_.alloc(std::max<long>(1, sum_n_)), wss_.alloc(std::max<long>(1, sum_m_)), wzs_.alloc(std::max<long>(1, sum_m_));
  xo_.alloc(std::max<long>(1, sum_n_)), so_.alloc(std::max<long>(1, sum_m_)), zo_.alloc(std::max<long>(1, sum_m_));
  info_.alloc((long)INFO_STRIDE * nprob);
  status_.alloc(nprob), iters_.alloc(nprob), have_warm_.alloc(nprob);
  have_warm_.zero(st_);
  if (stage_q_) (void)hipHostFree(stage_q_), stage_q_ = nullptr;
  if (stage_b_) (void)hipHostFree(stage_b_), stage_b_ = nullptr;
  CX_HIP(hipHostMalloc((void**)&stage_q_, std::max<long>(1, sum_n_) * sizeof(double), hipHostMallocDefault));
  CX_HIP(hipHostMalloc((void**)&stage_b_, std::max<long>(1, sum_m_) * sizeof(double), hipHostMallocDefault));
  std::memcpy(stage_q_, qv.data(), sum_n_ * sizeof(double));
  std::memcpy(stage_b_, bv.data(), sum_m_ * sizeof(double));
  // default tuning: every M that fits in LDS goes there, one workgroup per CU slot
  int dev = 0, ncu = 0;
  CX_HIP(hipGetDevice(&dev));
  CX_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  tune_.lds_doubles = std::min<long>(max_lds_need_, LDS_MAX_BYTES / 8);
  tune_.grid = std::min(nprob, ncu * std::max(1, (LDS_MAX_BYTES / 8) / std::max(1, tune_.lds_doubles)));
  tune_.grid = std::max(1, std::min(tune_.grid, ncu * 8));
  {
    int khz = 0;
    CX_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
    clock_khz_ = khz > 0 ? khz : 100000;
  }
  CX_HIP(hipFuncSetAttribute((const void*)k_batch_solve, hipFuncAttributeMaxDynamicSharedMemorySize,
                             LDS_MAX_BYTES));
  CX_HIP(hipStreamSynchronize(st_));
  ready_ = true;
  fetched_ = false;
  if (diff_) {
    diff_ = false;
    set_differentiable(true);
  }
  return true;
}

bool BatchSolver::update_data(const int* const* Pp, const int* const* Pi, const double* const* Px,
                              const double* const* q, const int* const* Ap, const int* const* Ai,
                              const double* const* Ax, const double* const* b) {
  if (!ready_) return false;
  // every problem's pattern is checked before anything changes
  // (KktSystem::sameSparsityPattern, kkt_system.cpp:10-19)
  auto same = [](const std::vector<int>& p, const std::vector<int>& i, int ncol, const int* pp, const int* ii) {
    if (!pp || !ii) return false;
    if (!std::equal(p.begin(), p.end(), pp)) return false;
    return std::equal(i.begin(), i.end(), ii) && pp[ncol] == (int)i.size();
  };
  for (int k = 0; k < np_; ++k) {
    const int nk = desc_h_[k].n;
    if (Px && Px[k] && !same(Pp_h_[k], Pi_h_[k], nk, Pp ? Pp[k] : nullptr, Pi ? Pi[k] : nullptr)) return false;
    if (Ax && Ax[k] && !same(Ap_h_[k], Ai_h_[k], nk, Ap ? Ap[k] : nullptr, Ai ? Ai[k] : nullptr)) return false;
  }
  bool anyP = false, anyA = false;
  for (int k = 0; k < np_; ++k) anyP |= Px && Px[k], anyA |= Ax && Ax[k];
  if (anyP && !stage_P_) {
    CX_HIP(hipHostMalloc((void**)&stage_P_, std::max<long>(1, nnzP_total_) * sizeof(double), hipHostMallocDefault));
    Praw_.alloc(std::max<long>(1, nnzP_total_));
    // problems not passed keep their values: the staging starts from setup's
    std::memcpy(stage_P_, Praw_h_.data(), Praw_h_.size() * sizeof(double));
  }
  if (anyA && !stage_A_) {
    CX_HIP(hipHostMalloc((void**)&stage_A_, std::max<long>(1, nnzA_total_) * sizeof(double), hipHostMallocDefault));
    Araw_.alloc(std::max<long>(1, nnzA_total_));
    std::memcpy(stage_A_, Araw_h_.data(), Araw_h_.size() * sizeof(double));
  }
  for (int k = 0; k < np_; ++k) {
    if (Px && Px[k]) std::memcpy(stage_P_ + pv_base_[k], Px[k], Pi_h_[k].size() * sizeof(double));
    if (Ax && Ax[k]) std::memcpy(stage_A_ + av_base_[k], Ax[k], Ai_h_[k].size() * sizeof(double));
  }
  P_dirty_ |= anyP, A_dirty_ |= anyA;
  return update_data(q, b);
}

bool BatchSolver::update_data(const double* const* q, const double* const* b) {
  if (!ready_) return false;
  for (int k = 0; k < np_; ++k) {
    const BatchDesc& D = desc_h_[k];
    if (q && q[k]) std::memcpy(stage_q_ + D.xo, q[k], D.n * sizeof(double));
    if (b && b[k]) std::memcpy(stage_b_ + D.co, b[k], D.m * sizeof(double));
  }
  return true;
}

void BatchSolver::build_args(void* argp) {
  BatchArgs& A = *static_cast<BatchArgs*>(argp);
  A.nprob = np_;
  A.desc = desc_.p;
  A.order = order_.p;
  A.counter = counter_.p;
  A.tol_feas = set_.tol_feas, A.tol_gap_abs = set_.tol_gap_abs, A.tol_gap_rel = set_.tol_gap_rel;
  A.tol_infeas = set_.tol_infeas, A.sreg_p = set_.static_reg_P, A.sreg_a = set_.static_reg_A;
  A.dyn_eps = set_.dynamic_reg_eps, A.dyn_delta = set_.dynamic_reg_delta;
  A.refine_tol = set_.refine_tol, A.step_frac = set_.max_step_fraction;
  // The adjoint solve differentiates a solution known to the IPM tolerances:
  // its residual target sits two orders below them.
  A.adj_tol = std::max(set_.refine_tol, 1e-2 * std::min(set_.tol_feas, set_.tol_gap_rel));
  A.max_iter = set_.max_iter, A.refine_max_iter = set_.refine_max_iter;
  A.mode = set_.linear_solver == CX_LS_REGULARIZED ? PV_REGULARIZE : PV_CHECK;
  A.warm_start = set_.warm_start;
  A.lds_doubles = tune_.lds_doubles;
  A.pf_ptr = pf_ptr_.p, A.pf_col = pf_col_.p, A.pf_src = pf_src_.p;
  A.ac_ptr = ac_ptr_.p, A.ac_row = ac_row_.p;
  A.ar_ptr = ar_ptr_.p, A.ar_col = ar_col_.p, A.ar_src = ar_src_.p;
  A.soc_off = soc_off_.p, A.soc_dim = soc_dim_.p;
  A.rk = rk_.p, A.kr = kr_.p;
  A.Pv = Pv_.p, A.Av = Av_.p, A.q = q_.p, A.b = b_.p, A.ed = ed_.p, A.ee = ee_.p;
  A.W = W_.p, A.Mg = Mg_.p;
  A.beta = beta_.p, A.vv = vv_.p, A.ident = ident_.p;
  A.wx = wxs_.p, A.ws = wss_.p, A.wz = wzs_.p, A.have_warm = have_warm_.p;
  A.xo = xo_.p, A.so = so_.p, A.zo = zo_.p, A.info = info_.p;
  A.status = status_.p, A.iters = iters_.p;
  fill_diff_args(&A);
}

void BatchSolver::launch_all() {
  BatchArgs A{};
  build_args(&A);
  const size_t lds = (size_t)std::max(1, tune_.lds_doubles) * sizeof(double);
  CX_HIP(hipMemcpyAsync(qraw_.p, stage_q_, sum_n_ * sizeof(double), hipMemcpyHostToDevice, st_));
  CX_HIP(hipMemcpyAsync(braw_.p, stage_b_, sum_m_ * sizeof(double), hipMemcpyHostToDevice, st_));
  hipLaunchKernelGGL(k_batch_prepare, dim3(ceil_div(std::max(sum_n_, sum_m_) + 1, 256)), dim3(256), 0, st_,
                     (int)sum_n_, (int)sum_m_, xprob_.p, desc_.p, qraw_.p, braw_.p, ed_.p, ee_.p, q_.p, b_.p,
                     counter_.p);
  CX_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_batch_solve, dim3(tune_.grid), dim3(BG), lds, st_, A);
  CX_LAUNCH_CHECK();
}

void BatchSolver::fill_diff_args(void* argp) {
  BatchArgs& A = *static_cast<BatchArgs*>(argp);
  A.diff = diff_ ? 1 : 0;
  A.tcorr = tcorr_.p, A.tdir = tdir_.p, A.xh = xh_.p, A.sh = sh_.p, A.zh = zh_.p;
  A.diff_a = diff_a_.p, A.diff_ok = diff_ok_.p;
  A.gwx = gwx_.p, A.gws = gws_.p, A.gwz = gwz_.p;
  A.gP = gP_.p, A.gq = gq_.p, A.gA = gA_.p, A.gb = gb_.p;
  A.counter2 = counter_.p + 1;
}

void BatchSolver::set_differentiable(bool on) {
  if (on == diff_) return;
  diff_ = on;
  if (on && ready_) {
    tcorr_.alloc(std::max<size_t>(1, nsoc_total_));
    tdir_.alloc(std::max<long>(1, sum_m_));
    xh_.alloc(std::max<long>(1, sum_n_));
    sh_.alloc(std::max<long>(1, sum_m_));
    zh_.alloc(std::max<long>(1, sum_m_));
    diff_a_.alloc(np_);
    diff_ok_.alloc(np_);
    diff_ok_.zero(st_);
    gwx_.alloc(std::max<long>(1, sum_n_));
    gws_.alloc(std::max<long>(1, sum_m_));
    gwz_.alloc(std::max<long>(1, sum_m_));
    gP_.alloc(std::max<long>(1, nnzP_total_));
    gA_.alloc(std::max<long>(1, nnzA_total_));
    gq_.alloc(std::max<long>(1, sum_n_));
    gb_.alloc(std::max<long>(1, sum_m_));
    CX_HIP(hipStreamSynchronize(st_));
  }
  if (gexec_) (void)hipGraphExecDestroy(gexec_), gexec_ = nullptr;
  if (graph_) (void)hipGraphDestroy(graph_), graph_ = nullptr;
}

int BatchSolver::backward(const double* const* wx, const double* const* ws, const double* const* wz,
                          double* const* gP, double* const* gq, double* const* gA, double* const* gb) {
  if (!ready_ || !diff_) return -1;
  std::vector<double> hx(std::max<long>(1, sum_n_), 0.0), hs(std::max<long>(1, sum_m_), 0.0),
      hz(std::max<long>(1, sum_m_), 0.0);
  for (int k = 0; k < np_; ++k) {
    const BatchDesc& D = desc_h_[k];
    if (wx && wx[k]) std::memcpy(hx.data() + D.xo, wx[k], D.n * sizeof(double));
    if (ws && ws[k]) std::memcpy(hs.data() + D.co, ws[k], D.m * sizeof(double));
    if (wz && wz[k]) std::memcpy(hz.data() + D.co, wz[k], D.m * sizeof(double));
  }
  gwx_.upload(hx, st_);
  gws_.upload(hs, st_);
  gwz_.upload(hz, st_);
  CX_HIP(hipMemsetAsync(gP_.p, 0, gP_.n * sizeof(double), st_));
  CX_HIP(hipMemsetAsync(gA_.p, 0, gA_.n * sizeof(double), st_));
  CX_HIP(hipMemsetAsync(counter_.p + 1, 0, sizeof(int), st_));
  BatchArgs A{};
  build_args(&A);
  const size_t lds = (size_t)std::max(1, tune_.lds_doubles) * sizeof(double);
  CX_HIP(hipEventRecord(ev0_, st_));
  hipLaunchKernelGGL(k_batch_backward, dim3(tune_.grid), dim3(BG), lds, st_, A);
  CX_LAUNCH_CHECK();
  CX_HIP(hipEventRecord(ev1_, st_));
  std::vector<double> hgP(gP_.n), hgA(gA_.n), hgq(gq_.n), hgb(gb_.n);
  std::vector<int> ok(np_);
  gP_.download(hgP.data(), gP_.n, st_);
  gA_.download(hgA.data(), gA_.n, st_);
  gq_.download(hgq.data(), gq_.n, st_);
  gb_.download(hgb.data(), gb_.n, st_);
  diff_ok_.download(ok.data(), np_, st_);
  CX_HIP(hipStreamSynchronize(st_));
  float ms = 0;
  CX_HIP(hipEventElapsedTime(&ms, ev0_, ev1_));
  last_bw_ms_ = ms;
  fetched_ = false;
  int nok = 0;
  for (int k = 0; k < np_; ++k) {
    const BatchDesc& D = desc_h_[k];
    nok += ok[k] ? 1 : 0;
    if (gP && gP[k]) std::memcpy(gP[k], hgP.data() + pv_base_[k], D.pnnz * sizeof(double));
    if (gA && gA[k]) std::memcpy(gA[k], hgA.data() + av_base_[k], av_len_[k] * sizeof(double));
    if (gq && gq[k]) std::memcpy(gq[k], hgq.data() + D.xo, D.n * sizeof(double));
    if (gb && gb[k]) std::memcpy(gb[k], hgb.data() + D.co, D.m * sizeof(double));
  }
  return nok;
}

// P / A values staged by update_data: uploaded and scaled on the solve's
// stream ahead of the graph (q and b travel inside the graph, every solve)
void BatchSolver::upload_values() {
  if (P_dirty_ && nnzP_total_ > 0) {
    CX_HIP(hipMemcpyAsync(Praw_.p, stage_P_, nnzP_total_ * sizeof(double), hipMemcpyHostToDevice, st_));
    hipLaunchKernelGGL(k_batch_scale_values, dim3(ceil_div(nnzP_total_, 256)), dim3(256), 0, st_, nnzP_total_,
                       Praw_.p, sP_.p, Pv_.p);
    CX_LAUNCH_CHECK();
  }
  if (A_dirty_ && nnzA_total_ > 0) {
    CX_HIP(hipMemcpyAsync(Araw_.p, stage_A_, nnzA_total_ * sizeof(double), hipMemcpyHostToDevice, st_));
    hipLaunchKernelGGL(k_batch_scale_values, dim3(ceil_div(nnzA_total_, 256)), dim3(256), 0, st_, nnzA_total_,
                       Araw_.p, sA_.p, Av_.p);
    CX_LAUNCH_CHECK();
  }
  P_dirty_ = A_dirty_ = false;
}

int BatchSolver::solve(int* status_out) {
  if (!ready_) return -1;
  if (!gexec_) {
    CX_HIP(hipStreamBeginCapture(st_, hipStreamCaptureModeThreadLocal));
    launch_all();
    CX_HIP(hipStreamEndCapture(st_, &graph_));
    CX_HIP(hipGraphInstantiate(&gexec_, graph_, nullptr, nullptr, 0));
  }
  CX_HIP(hipEventRecord(ev0_, st_));
  upload_values();
  CX_HIP(hipGraphLaunch(gexec_, st_));
  CX_HIP(hipEventRecord(ev1_, st_));
  CX_HIP(hipStreamSynchronize(st_));
  float ms = 0;
  CX_HIP(hipEventElapsedTime(&ms, ev0_, ev1_));
  last_ms_ = ms;
  fetched_ = false;
  fetch();
  int solved = 0;
  for (int k = 0; k < np_; ++k) {
    if (status_out) status_out[k] = hstatus_[k];
    solved += hstatus_[k] == CX_SOLVED;
  }
  return solved;
}

void BatchSolver::fetch() {
  if (fetched_) return;
  hx_.resize(std::max<long>(1, sum_n_));
  hs_.resize(std::max<long>(1, sum_m_));
  hz_.resize(std::max<long>(1, sum_m_));
  hinfo_.resize((long)INFO_STRIDE * np_);
  hstatus_.resize(np_);
  hiters_.resize(np_);
  xo_.download(hx_.data(), sum_n_, st_);
  so_.download(hs_.data(), sum_m_, st_);
  zo_.download(hz_.data(), sum_m_, st_);
  info_.download(hinfo_.data(), (long)INFO_STRIDE * np_, st_);
  status_.download(hstatus_.data(), np_, st_);
  iters_.download(hiters_.data(), np_, st_);
  CX_HIP(hipStreamSynchronize(st_));
  fetched_ = true;
}

int BatchSolver::get_solution(int k, double* x, double* s, double* z, double* obj, cx_info* info) {
  if (!ready_ || k < 0 || k >= np_) return CX_UNSOLVED;
  fetch();
  const BatchDesc& D = desc_h_[k];
  if (x) std::memcpy(x, hx_.data() + D.xo, D.n * sizeof(double));
  if (s) std::memcpy(s, hs_.data() + D.co, D.m * sizeof(double));
  if (z) std::memcpy(z, hz_.data() + D.co, D.m * sizeof(double));
  if (obj) *obj = hinfo_[(long)INFO_STRIDE * k + 4];
  if (info) {
    info->iterations = hiters_[k];
    info->primal_residual = hinfo_[(long)INFO_STRIDE * k + 0];
    info->dual_residual = hinfo_[(long)INFO_STRIDE * k + 1];
    info->duality_gap = hinfo_[(long)INFO_STRIDE * k + 2];
    info->mu = hinfo_[(long)INFO_STRIDE * k + 3];
    info->setup_time_s = 0;
    info->solve_time_s = last_ms_ * 1e-3;
  }
  return hstatus_[k];
}

// Launch-configuration autotuner.  A candidate is an LDS budget per
// workgroup (all of it: every factor that fits kept on chip; or a half, third,
// quarter: larger factors in global memory, some unstaged) and the matching
// persistent grid (resident slots, then twice that).  A smaller budget is a
// candidate only when the runtime's occupancy calculator gives it more
// resident workgroups per CU than every larger budget (registers can cap the
// residency below what the LDS budget would allow).  Each candidate solves
// this batch reps times from the same starting state (the warm-start flags
// are restored before every trial and at the end); the fastest device time
// wins.  Solutions left in place are those of the last trial.
bool BatchSolver::set_tuning(int lds_doubles, int grid) {
  if (!ready_ || lds_doubles < 0 || lds_doubles > LDS_MAX_BYTES / 8 || grid < 1) return false;
  tune_.lds_doubles = lds_doubles;
  tune_.grid = grid;
  tuned_ = true;
  if (gexec_) (void)hipGraphExecDestroy(gexec_), gexec_ = nullptr;
  if (graph_) (void)hipGraphDestroy(graph_), graph_ = nullptr;
  return true;
}

int BatchSolver::autotune(int reps) {
  if (!ready_) return -1;
  reps = std::max(1, reps);
  int dev = 0, ncu = 0;
  CX_HIP(hipGetDevice(&dev));
  CX_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  std::vector<int> warm(np_);
  have_warm_.download(warm.data(), np_, st_);
  CX_HIP(hipStreamSynchronize(st_));
  const long full = std::min<long>(max_lds_need_, LDS_MAX_BYTES / 8);
  std::vector<BatchTuning> cand;
  int prev_res = 0;
  for (int per_cu = 1; per_cu <= 4; ++per_cu) {
    const int lds = (int)std::min<long>(full, (LDS_MAX_BYTES / 8) / per_cu);
    int res = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&res, reinterpret_cast<const void*>(&k_batch_solve), BG,
                                                     (size_t)lds * sizeof(double)) != hipSuccess ||
        res < 1)
      res = 1;
    if (res <= prev_res) continue;
    prev_res = res;
    const int slots = ncu * res;
    for (int over = 1; over <= 2; ++over) {
      BatchTuning t;
      t.lds_doubles = lds;
      t.grid = std::max(1, std::min(np_, slots * over));
      bool dup = false;
      for (const BatchTuning& c : cand) dup |= c.lds_doubles == t.lds_doubles && c.grid == t.grid;
      if (!dup) cand.push_back(t);
    }
  }
  const BatchTuning keep = tune_;
  double best_ms = 1e300;
  BatchTuning best = keep;
  for (const BatchTuning& t : cand) {
    tune_ = t;
    if (gexec_) (void)hipGraphExecDestroy(gexec_), gexec_ = nullptr;
    if (graph_) (void)hipGraphDestroy(graph_), graph_ = nullptr;
    double ms = 1e300;
    for (int r = 0; r < reps; ++r) {
      have_warm_.upload(warm, st_);
      if (solve(nullptr) < 0) break;
      ms = std::min(ms, last_ms_);
    }
    if (ms < best_ms) best_ms = ms, best = t;
  }
  tune_ = best;
  tuned_ = true;
  if (gexec_) (void)hipGraphExecDestroy(gexec_), gexec_ = nullptr;
  if (graph_) (void)hipGraphDestroy(graph_), graph_ = nullptr;
  have_warm_.upload(warm, st_);
  CX_HIP(hipStreamSynchronize(st_));
  return (int)cand.size();
}

void BatchSolver::stats(int k, double* out13) {
  for (int i = 0; i < 13; ++i) out13[i] = 0.0;
  if (!ready_ || k < 0 || k >= np_) return;
  fetch();
  out13[0] = hiters_[k];
  out13[1] = hinfo_[(long)INFO_STRIDE * k + 5];
  out13[2] = hinfo_[(long)INFO_STRIDE * k + 6];
  out13[3] = hinfo_[(long)INFO_STRIDE * k + 7] / (clock_khz_ * 1e-3);  // microseconds
  out13[4] = hinfo_[(long)INFO_STRIDE * k + 8] / (clock_khz_ * 1e-3);
  out13[5] = hinfo_[(long)INFO_STRIDE * k + 9] / (clock_khz_ * 1e-3);
  out13[6] = hinfo_[(long)INFO_STRIDE * k + 10] / (clock_khz_ * 1e-3);
  out13[7] = hinfo_[(long)INFO_STRIDE * k + 11];
  out13[8] = hinfo_[(long)INFO_STRIDE * k + 12] / (clock_khz_ * 1e-3);
  out13[9] = hinfo_[(long)INFO_STRIDE * k + 13] / (clock_khz_ * 1e-3);
  out13[10] = hinfo_[(long)INFO_STRIDE * k + 14];
  out13[11] = hinfo_[(long)INFO_STRIDE * k + 15] / (clock_khz_ * 1e-3);
  out13[12] = hinfo_[(long)INFO_STRIDE * k + 16] / (clock_khz_ * 1e-3);
}

void BatchSolver::tuning(double* out4) const {
  out4[0] = tune_.lds_doubles;
  out4[1] = tune_.grid;
  out4[2] = max_packed_;
  out4[3] = tuned_ ? 1 : 0;
}

}  // namespace cx
