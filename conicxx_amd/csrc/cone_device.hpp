// Device cone algebra shared by the per-problem kernels (ipm_kernels.hip) and
// the batched persistent solver (batch.hip).
//
// Follows src/conicxx/cones/second_order_cone.cpp and nonnegative_cone.cpp.
// The SOC Nesterov-Todd scaling is kept in its rank-one form W = beta (2 v v' - J)
// (the matrix the reference builds densely at second_order_cone.cpp:133-136):
//   W x      = beta (2 v (v'x) - J x)
//   W^-1 x   = (1/beta) (2 Jv ((Jv)'x) - J x)       (v'Jv = 1)
//   Hs x     = beta^2 (4 (v'v) v (v'x) - 2 v ((Jv)'x) - 2 Jv (v'x) + x)
// with the reference's identity fallback when W is non-finite or
// max|W_ij| > 1e8 (second_order_cone.cpp:152-156).  Template parameter T is the
// number of lanes sharing one block (1: a thread per block, 64: a wave).
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>

#include "ipm_kernels.hpp"

namespace cx {
namespace conedev {

constexpr double kTiny = 1e-12;
constexpr double kAbsFloor = DBL_MIN * 1e4;
constexpr double kStepEps = DBL_MIN * 1e4;

template <int T>
__device__ inline double tsum(double v) {
  if constexpr (T > 1) {
#pragma unroll
    for (int o = T / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  }
  return v;
}
template <int T>
__device__ inline double tmin(double v) {
  if constexpr (T > 1) {
#pragma unroll
    for (int o = T / 2; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
  }
  return v;
}

// ---------------------------------------------------------------------------
// SOC helpers with team size T (1: one thread per block, 64: one wave per block)
// ---------------------------------------------------------------------------
struct SocState {
  double beta, vv;
  int ident;
};

__device__ inline double jsign(int i) { return i == 0 ? 1.0 : -1.0; }

// (W x)_i given v'x
__device__ inline double w_apply(const SocState& S, const double* v, int i, double xi, double vx) {
  if (S.ident) return xi;
  return S.beta * (2.0 * v[i] * vx - jsign(i) * xi);
}
// (W^-1 x)_i given (Jv)'x
__device__ inline double winv_apply(const SocState& S, const double* v, int i, double xi, double jvx) {
  if (S.ident) return xi;
  return (2.0 * jsign(i) * v[i] * jvx - jsign(i) * xi) / S.beta;
}

template <int T>
__device__ inline void soc_scaling(int lane, int off, int d, const double* s, const double* z, double* v,
                            double* beta_o, double* vv_o, int* ident_o) {
  const double s0 = s[off], z0 = z[off];
  double ps1 = 0.0, pz1 = 0.0, psz = 0.0;
  for (int i = 1 + lane; i < d; i += T) {
    const double si = s[off + i], zi = z[off + i];
    ps1 += si * si;
    pz1 += zi * zi;
    psz += si * zi;
  }
  ps1 = tsum<T>(ps1);
  pz1 = tsum<T>(pz1);
  psz = tsum<T>(psz);
  const double a = sqrt(fmax(fmax(s0 * s0 - ps1, 1e-12 * s0 * s0), kAbsFloor));
  const double b = sqrt(fmax(fmax(z0 * z0 - pz1, 1e-12 * z0 * z0), kAbsFloor));
  const double beta = fmin(fmax(sqrt(a / b), 1e-8), 1e8);
  const double sz = s0 * z0 + psz;
  const double c = sqrt(fmax((sz / (a * b) + 1.0) / 2.0, kTiny));
  double v0 = (z0 / b + s0 / a) / (2.0 * c);
  v0 += 1.0;
  v0 = fmax(v0, kTiny);
  const double sc = sqrt(2.0 * v0);
  const double v0n = v0 / sc;
  // v tail, v'v, finiteness, two largest |v_i| (i >= 1)
  double pvv = 0.0, m1 = 0.0, m2 = 0.0, dmax = 0.0;
  bool fin = true;
  for (int i = 1 + lane; i < d; i += T) {
    double vi = ((-z[off + i] / b) + s[off + i] / a) / (2.0 * c);
    vi = vi / sc;
    v[off + i] = vi;
    pvv += vi * vi;
    const double av = fabs(vi);
    if (!isfinite(vi)) fin = false;
    if (av > m1) m2 = m1, m1 = av;
    else if (av > m2) m2 = av;
    dmax = fmax(dmax, fabs(2.0 * vi * vi + 1.0));
  }
  pvv = tsum<T>(pvv);
  dmax = fmax(-tmin<T>(-dmax), 0.0);
  // merge top-two across the team
  if constexpr (T > 1) {
#pragma unroll
    for (int o = T / 2; o > 0; o >>= 1) {
      double b1 = __shfl_xor(m1, o, 64), b2 = __shfl_xor(m2, o, 64);
      double hi = fmax(m1, b1);
      double lo = fmax(fmin(m1, b1), fmax(m2, b2));
      m1 = hi;
      m2 = lo;
    }
  }
  int nonfin = fin ? 0 : 1;
  if constexpr (T > 1) {
#pragma unroll
    for (int o = T / 2; o > 0; o >>= 1) nonfin |= __shfl_xor(nonfin, o, 64);
  }
  const double vv = v0n * v0n + pvv;
  // max |W_ij| / beta over diagonal and off-diagonal entries
  double wdiag = fmax(fabs(2.0 * v0n * v0n - 1.0), dmax);
  double off01 = 2.0 * fabs(v0n) * m1;  // (0, i) entries
  double offii = 2.0 * m1 * m2;         // (i, j) entries, i != j >= 1
  double wmax = beta * fmax(wdiag, fmax(off01, offii));
  bool bad = nonfin || !isfinite(v0n) || !isfinite(beta) || !isfinite(wmax) || wmax > 1e8;
  if (lane == 0) {
    v[off] = v0n;
    *beta_o = beta;
    *vv_o = vv;
    *ident_o = bad ? 1 : 0;
  }
}

// Entry (a, b), a >= b, of one SOC block's Hs = W W = beta^2 (2 v v' - J)^2 as
// the KKT matrix stores it (the reference writes the dense Hs block into K,
// kkt_system.cpp:139-150).  Every product that must agree with the factored
// matrix (the refinement residual) uses these entries: near the cone boundary
// |v| is large and this entry-wise form and the rank-one product below differ
// by far more than a rounding of the solution.
__device__ inline double soc_hs_entry(double beta, double vv, const double* v, int a, int b) {
  const double va = v[a], vb = v[b];
  return beta * beta *
         (4.0 * vv * va * vb - 2.0 * va * (jsign(b) * vb) - 2.0 * (jsign(a) * va) * vb + (a == b ? 1.0 : 0.0));
}

// out = -(Hs x + y)
template <int T>
__device__ inline void soc_neg_hs_plus(int lane, const ConeDev& C, int k, const double* x, const double* y,
                                double* out) {
  const int off = C.soc_off[k], d = C.soc_dim[k];
  SocState S{C.beta[k], C.vv[k], C.ident[k]};
  const double* v = C.v + off;
  double vx = 0.0, jvx = 0.0;
  for (int i = lane; i < d; i += T) {
    vx += v[i] * x[off + i];
    jvx += jsign(i) * v[i] * x[off + i];
  }
  vx = tsum<T>(vx);
  jvx = tsum<T>(jvx);
  const double b2 = S.beta * S.beta;
  for (int i = lane; i < d; i += T) {
    double h;
    if (S.ident)
      h = x[off + i];
    else
      h = b2 * (4.0 * S.vv * v[i] * vx - 2.0 * v[i] * jvx - 2.0 * jsign(i) * v[i] * vx + x[off + i]);
    out[off + i] = -(h + y[off + i]);
  }
}

// combined-step cone rhs (solver_impl.cpp:233-251) for one SOC block
template <int T>
__device__ inline void soc_combined(int lane, const ConeDev& C, int k, const double* z, const double* dsa,
                             const double* dza, const double* rz, double smu, double oms,
                             double* dsc_out, double* tail) {
  const int off = C.soc_off[k], d = C.soc_dim[k];
  SocState S{C.beta[k], C.vv[k], C.ident[k]};
  const double* v = C.v + off;
  // pass 1: v'z, (Jv)'ds_aff, v'dz_aff
  double vz = 0.0, jvds = 0.0, vdz = 0.0;
  for (int i = lane; i < d; i += T) {
    vz += v[i] * z[off + i];
    jvds += jsign(i) * v[i] * dsa[off + i];
    vdz += v[i] * dza[off + i];
  }
  vz = tsum<T>(vz);
  jvds = tsum<T>(jvds);
  vdz = tsum<T>(vdz);
  // element values: lambda = W z, a1 = W^-1 ds_aff, a2 = W dz_aff
  const double l0 = w_apply(S, v, 0, z[off], vz);
  const double a10 = winv_apply(S, v, 0, dsa[off], jvds);
  const double a20 = w_apply(S, v, 0, dza[off], vdz);
  // pass 2: lambda'lambda, a1'a2, ||lambda_1||^2
  double lam2 = 0.0, a12 = 0.0, r1 = 0.0;
  for (int i = lane; i < d; i += T) {
    double li = w_apply(S, v, i, z[off + i], vz);
    double a1 = winv_apply(S, v, i, dsa[off + i], jvds);
    double a2 = w_apply(S, v, i, dza[off + i], vdz);
    lam2 += li * li;
    a12 += a1 * a2;
    if (i > 0) r1 += li * li;
  }
  lam2 = tsum<T>(lam2);
  a12 = tsum<T>(a12);
  r1 = tsum<T>(r1);
  const double dsc0 = lam2 + a12 - smu;
  // pass 3: nu = lambda_1' dsc_1
  double nu = 0.0;
  for (int i = 1 + lane; i < d; i += T) {
    double li = w_apply(S, v, i, z[off + i], vz);
    double a1 = winv_apply(S, v, i, dsa[off + i], jvds);
    double a2 = w_apply(S, v, i, dza[off + i], vdz);
    double dsci = (l0 * li + li * l0) + (a10 * a2 + a20 * a1);
    nu += li * dsci;
  }
  nu = tsum<T>(nu);
  // inverse product lambda \ dsc (second_order_cone.cpp:50-68)
  double rho = l0 * l0 - r1;
  if (fabs(rho) < kTiny) rho = (rho >= 0 ? kTiny : -kTiny);
  double u0 = l0;
  if (fabs(u0) < kTiny) u0 = kTiny;
  const double lid0 = (u0 * dsc0 - nu) / rho;
  // pass 4: v' lid  (W lid needs it)
  double vlid = 0.0;
  for (int i = lane; i < d; i += T) {
    double lidi;
    if (i == 0) {
      lidi = lid0;
    } else {
      double li = w_apply(S, v, i, z[off + i], vz);
      double a1 = winv_apply(S, v, i, dsa[off + i], jvds);
      double a2 = w_apply(S, v, i, dza[off + i], vdz);
      double dsci = (l0 * li + li * l0) + (a10 * a2 + a20 * a1);
      lidi = ((nu / u0 - dsc0) * li + (rho / u0) * dsci) / rho;
    }
    vlid += v[i] * lidi;
  }
  vlid = tsum<T>(vlid);
  // pass 5: ds_const = W lid
  for (int i = lane; i < d; i += T) {
    double lidi;
    if (i == 0) {
      lidi = lid0;
    } else {
      double li = w_apply(S, v, i, z[off + i], vz);
      double a1 = winv_apply(S, v, i, dsa[off + i], jvds);
      double a2 = w_apply(S, v, i, dza[off + i], vdz);
      double dsci = (l0 * li + li * l0) + (a10 * a2 + a20 * a1);
      lidi = ((nu / u0 - dsc0) * li + (rho / u0) * dsci) / rho;
    }
    double dsi = w_apply(S, v, i, lidi, vlid);
    dsc_out[off + i] = dsi;
    tail[off + i] = dsi - oms * rz[off + i];
  }
}

// SecondOrderCone::maxStep (second_order_cone.cpp:174-215), alpha_max = +inf
template <int T>
__device__ inline double soc_max_step(int lane, int off, int d, const double* x, const double* dx) {
  double pd = 0.0, dd = 0.0, pp = 0.0;
  for (int i = 1 + lane; i < d; i += T) {
    pd += x[off + i] * dx[off + i];
    dd += dx[off + i] * dx[off + i];
    pp += x[off + i] * x[off + i];
  }
  pd = tsum<T>(pd);
  dd = tsum<T>(dd);
  pp = tsum<T>(pp);
  const double p0 = x[off], d0 = dx[off];
  const double a = d0 * d0 - dd;
  const double b = 2.0 * (p0 * d0 - pd);
  const double c = p0 * p0 - pp;
  double alpha = INFINITY;
  auto feas = [&](double r) { return r > 0 && (p0 + r * d0) >= 0; };
  if (fabs(a) < kStepEps) {
    if (fabs(b) > kStepEps) {
      double r = -c / b;
      if (feas(r)) alpha = fmin(alpha, r);
    }
    return alpha;
  }
  double disc = b * b - 4.0 * a * c;
  if (disc < 0) return alpha;
  double sq = sqrt(disc);
  double q = (b >= 0) ? -0.5 * (b + sq) : -0.5 * (b - sq);
  double r1, r2;
  if (fabs(q) > kStepEps) {
    r1 = q / a;
    r2 = c / q;
  } else {
    r1 = r2 = -b / (2.0 * a);
  }
  if (feas(r1)) alpha = fmin(alpha, r1);
  if (feas(r2)) alpha = fmin(alpha, r2);
  return alpha;
}


}  // namespace conedev
}  // namespace cx
