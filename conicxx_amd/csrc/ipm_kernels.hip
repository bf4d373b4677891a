// Device kernels of the interior-point iteration (one problem per launch).
//
// Cone algebra follows src/conicxx/cones/*.cpp through the shared device
// functions of cone_device.hpp.  Small SOC blocks are handled one per thread,
// large ones one per 64-lane wave with shuffle reductions.  Reductions are
// two-stage and deterministic.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>

#include "common.hpp"
#include "cone_device.hpp"
#include "ipm_kernels.hpp"

namespace cx {

namespace {

using namespace conedev;

// ---------------------------------------------------------------------------
// block-level reduction helpers (256 threads)
// ---------------------------------------------------------------------------
__device__ inline double block_sum(double v, double* sh) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[wv] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) r = (sh[0] + sh[1]) + (sh[2] + sh[3]);
  return r;
}
__device__ inline double block_min(double v, double* sh) {
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[wv] = v;
  __syncthreads();
  double r = INFINITY;
  if (threadIdx.x == 0) r = fmin(fmin(sh[0], sh[1]), fmin(sh[2], sh[3]));
  return r;
}

// ---------------------------------------------------------------------------
// multi-term reductions
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_reduce_partial(RedBatch B, double* partial) {
  __shared__ double sh[4];
  for (int o = 0; o < B.nops; ++o) {
    const RedOp op = B.op[o];
    double acc = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < op.len; i += gridDim.x * blockDim.x) {
      if (op.kind == RED_DOT) {
        acc += op.a[i] * op.b[i];
      } else if (op.kind == RED_DOT_AXPY) {
        acc += (op.a[i] + op.alpha * op.c[i]) * (op.b[i] + op.alpha * op.d[i]);
      } else {
        acc += isfinite(op.a[i]) ? 0.0 : 1.0;
      }
    }
    double r = block_sum(acc, sh);
    if (threadIdx.x == 0) partial[o * RED_BLOCKS + blockIdx.x] = r;
  }
}

__global__ __launch_bounds__(256) void k_reduce_final(const double* partial, int nops, double* out) {
  __shared__ double sh[4];
  for (int o = 0; o < nops; ++o) {
    double v = (threadIdx.x < RED_BLOCKS) ? partial[o * RED_BLOCKS + threadIdx.x] : 0.0;
    double r = block_sum(v, sh);
    if (threadIdx.x == 0) out[o] = r;
  }
}

__global__ __launch_bounds__(256) void k_soc_scaling_small(ConeDev C, const double* s, const double* z) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= C.nsmall) return;
  int k = C.small_list[t];
  soc_scaling<1>(0, C.soc_off[k], C.soc_dim[k], s, z, C.v, &C.beta[k], &C.vv[k], &C.ident[k]);
}
__global__ __launch_bounds__(256) void k_soc_scaling_large(ConeDev C, const double* s, const double* z) {
  int t = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  int lane = threadIdx.x & 63;
  if (t >= C.nlarge) return;
  int k = C.large_list[t];
  soc_scaling<64>(lane, C.soc_off[k], C.soc_dim[k], s, z, C.v, &C.beta[k], &C.vv[k], &C.ident[k]);
}
__global__ void k_nn_scaling(ConeDev C, const double* s, const double* z) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= C.nn_dim) return;
  int r = C.nn_off + i;
  double w = sqrt(s[r] / z[r]);
  C.w[r] = w;
  C.hsd[r] = w * w;
}

__global__ __launch_bounds__(256) void k_neg_hs_plus(ConeDev C, int m, const double* x,
                                                     const double* y, double* out) {
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  // zero and nonnegative rows elementwise
  for (int i = tid; i < C.zero_dim + C.nn_dim; i += gridDim.x * blockDim.x) {
    if (i < C.zero_dim) {
      int r = C.zero_off + i;
      out[r] = -(0.0 + y[r]);
    } else {
      int r = C.nn_off + (i - C.zero_dim);
      out[r] = -(C.hsd[r] * x[r] + y[r]);
    }
  }
  for (int t = tid; t < C.nsmall; t += gridDim.x * blockDim.x) soc_neg_hs_plus<1>(0, C, C.small_list[t], x, y, out);
  const int wave = tid >> 6, lane = threadIdx.x & 63, nw = (gridDim.x * blockDim.x) >> 6;
  for (int t = wave; t < C.nlarge; t += nw) soc_neg_hs_plus<64>(lane, C, C.large_list[t], x, y, out);
}

__global__ __launch_bounds__(256) void k_combined_rhs(ConeDev C, const double* z, const double* dsa,
                                                      const double* dza, const double* rz,
                                                      double smu, double oms, double* dsc,
                                                      double* tail) {
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  for (int i = tid; i < C.zero_dim + C.nn_dim; i += gridDim.x * blockDim.x) {
    if (i < C.zero_dim) {
      int r = C.zero_off + i;
      dsc[r] = 0.0;
      tail[r] = 0.0 - oms * rz[r];
    } else {
      int r = C.nn_off + (i - C.zero_dim);
      double w = C.w[r];
      double lam = w * z[r];
      double a1 = dsa[r] / w;
      double a2 = w * dza[r];
      double dc = lam * lam + a1 * a2 - smu;
      double lid = dc / lam;
      double ds = w * lid;
      dsc[r] = ds;
      tail[r] = ds - oms * rz[r];
    }
  }
  for (int t = tid; t < C.nsmall; t += gridDim.x * blockDim.x)
    soc_combined<1>(0, C, C.small_list[t], z, dsa, dza, rz, smu, oms, dsc, tail);
  const int wave = tid >> 6, lane = threadIdx.x & 63, nw = (gridDim.x * blockDim.x) >> 6;
  for (int t = wave; t < C.nlarge; t += nw)
    soc_combined<64>(lane, C, C.large_list[t], z, dsa, dza, rz, smu, oms, dsc, tail);
}

__global__ __launch_bounds__(256) void k_max_step(ConeDev C, const double* x, const double* dx,
                                                  double* partial) {
  __shared__ double sh[4];
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  double a = INFINITY;
  for (int i = tid; i < C.nn_dim; i += gridDim.x * blockDim.x) {
    int r = C.nn_off + i;
    if (dx[r] < 0) a = fmin(a, -x[r] / dx[r]);
  }
  for (int t = tid; t < C.nsmall; t += gridDim.x * blockDim.x) {
    int k = C.small_list[t];
    a = fmin(a, soc_max_step<1>(0, C.soc_off[k], C.soc_dim[k], x, dx));
  }
  const int wave = tid >> 6, lane = threadIdx.x & 63, nw = (gridDim.x * blockDim.x) >> 6;
  for (int t = wave; t < C.nlarge; t += nw) {
    int k = C.large_list[t];
    a = fmin(a, soc_max_step<64>(lane, C.soc_off[k], C.soc_dim[k], x, dx));
  }
  double r = block_min(a, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = r;
}

__global__ __launch_bounds__(256) void k_min_final(const double* partial, int n, double* out) {
  __shared__ double sh[4];
  double v = INFINITY;
  for (int i = threadIdx.x; i < n; i += blockDim.x) v = fmin(v, partial[i]);
  double r = block_min(v, sh);
  if (threadIdx.x == 0) out[0] = r;
}

// margins: partial[0..G) = nonneg min, [G..2G) = SOC min, [2G..3G) = SOC sum of positive parts
__global__ __launch_bounds__(256) void k_margins(ConeDev C, const double* x, double* partial) {
  __shared__ double sh[4];
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  double nmin = INFINITY, smin = INFINITY, spos = 0.0;
  for (int i = tid; i < C.nn_dim; i += gridDim.x * blockDim.x) nmin = fmin(nmin, x[C.nn_off + i]);
  for (int t = tid; t < C.nsmall; t += gridDim.x * blockDim.x) {
    int k = C.small_list[t], off = C.soc_off[k], d = C.soc_dim[k];
    double ss = 0.0;
    for (int i = 1; i < d; ++i) ss += x[off + i] * x[off + i];
    double mg = x[off] - sqrt(ss);
    if (isfinite(mg)) smin = fmin(smin, mg), spos += fmax(mg, 0.0);
  }
  const int wave = tid >> 6, lane = threadIdx.x & 63, nw = (gridDim.x * blockDim.x) >> 6;
  for (int t = wave; t < C.nlarge; t += nw) {
    int k = C.large_list[t], off = C.soc_off[k], d = C.soc_dim[k];
    double ss = 0.0;
    for (int i = 1 + lane; i < d; i += 64) ss += x[off + i] * x[off + i];
    ss = tsum<64>(ss);
    double mg = x[off] - sqrt(ss);
    if (lane == 0 && isfinite(mg)) smin = fmin(smin, mg), spos += fmax(mg, 0.0);
  }
  double a = block_min(nmin, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = a;
  double b = block_min(smin, sh);
  if (threadIdx.x == 0) partial[gridDim.x + blockIdx.x] = b;
  double c = block_sum(spos, sh);
  if (threadIdx.x == 0) partial[2 * gridDim.x + blockIdx.x] = c;
}

__global__ __launch_bounds__(256) void k_margins_final(const double* partial, int G, int has_nn,
                                                       double* out) {
  __shared__ double sh[4];
  double a = INFINITY, b = INFINITY, c = 0.0;
  for (int i = threadIdx.x; i < G; i += blockDim.x) {
    a = fmin(a, partial[i]);
    b = fmin(b, partial[G + i]);
    c += partial[2 * G + i];
  }
  double nmin = block_min(a, sh);
  double smin = block_min(b, sh);
  double spos = block_sum(c, sh);
  if (threadIdx.x == 0) {
    double mn = smin, ps = spos;
    if (has_nn && isfinite(nmin)) {
      mn = fmin(mn, nmin);
      ps += fmax(nmin, 0.0);
    }
    if (!isfinite(mn)) mn = 0.0;
    out[0] = mn;
    out[1] = ps;
  }
}

__global__ void k_unit_shift(ConeDev C, double* x, double alpha) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < C.nn_dim) x[C.nn_off + i] += alpha;
  if (i < C.nsoc) x[C.soc_off[i]] += alpha;
}

__global__ void k_fill(double* x, int n, double v) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = v;
}

__global__ __launch_bounds__(256) void k_csr_mv(int rows, const int* __restrict__ ptr,
                                                const int* __restrict__ col, const int* __restrict__ src,
                                                const double* __restrict__ val,
                                                const double* __restrict__ x, double* y) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= rows) return;
  double acc = 0.0;
  for (int e = ptr[i]; e < ptr[i + 1]; ++e) acc += val[src[e]] * x[col[e]];
  y[i] = acc;
}

__global__ void k_axpby(int n, double a, const double* x, double b, double* y) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = a * x[i] + b * y[i];
}
__global__ void k_lin2(int n, double a, const double* x, double b, const double* y, double* out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = a * x[i] + b * y[i];
}
__global__ void k_lin3(int n, double a, const double* x, double b, const double* y, double c,
                       const double* z, double* out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = a * x[i] + b * y[i] + c * z[i];
}
__global__ void k_scale(int n, double a, double* x) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] *= a;
}
__global__ void k_copy(int n, const double* x, double* y) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = x[i];
}
__global__ void k_scatter(int n, const int* slot, const int* src, const double* val, double add,
                          double* Kx) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int s = src[i];
  double v = (s >= 0) ? val[s] : 0.0;
  Kx[slot[i]] = v + add;
}

__global__ void k_hs_scatter(ConeDev C, int nhs, const int* slot, const int* kind, const int* blk,
                             const int* ia, const int* ib, const int* row, double sreg_a,
                             const double* tcorr, const double* tdir, double* Kx) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nhs) return;
  const int a = ia[t], b = ib[t];
  double h = 0.0;
  if (kind[t] == 1) {
    h = C.hsd[row[t]];
  } else if (kind[t] == 2) {
    const int k = blk[t];
    if (C.ident[k]) {
      h = (a == b) ? 1.0 : 0.0;
    } else {
      h = soc_hs_entry(C.beta[k], C.vv[k], C.v + C.soc_off[k], a, b);
      if (tcorr && a > 0 && b > 0) {
        const double* u = tdir + C.soc_off[k];
        h += tcorr[k] * ((a == b ? 1.0 : 0.0) - u[a] * u[b]);
      }
    }
  }
  double val = -h;
  if (a == b) val -= sreg_a;
  Kx[slot[t]] = val;
}

__global__ void k_scale_csc(int cols, const int* ptr, const int* row, const double* in, double c,
                            const double* drow, const double* dcol, double* out) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= cols) return;
  for (int k = ptr[j]; k < ptr[j + 1]; ++k) out[k] = in[k] * (c * drow[row[k]] * dcol[j]);
}
__global__ void k_vscale(int n, const double* d, double c, const double* in, double* out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = in[i] * (c * d[i]);
}
__global__ void k_vmuldiv(int n, const double* in, const double* mul, const double* div, double c,
                          double* out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    double v = in[i];
    if (mul) v = v * mul[i];
    if (div) v = v / div[i];
    out[i] = v * c;
  }
}

// Adjoint gradient on a CSC pattern, one thread per stored entry (column found
// by binary search).  u, v: adjoint solution (x part, cone part); xh, zh: the
// optimum; all in the equilibrated space, xh = xs * inv_tau, zh = zs * inv_tau.
// mode 0, A (m x n):      g = -(zh[i] u[j] + v[i] xh[j]) * drow[i] * dcol[j]
// mode 1, P upper (n x n): i < j: -(u[i] xh[j] + u[j] xh[i]) * c * drow[i] * dcol[j]
//                          i == j: -u[i] xh[i] * c * drow[i]^2;  i > j (ignored entry): 0
__global__ void k_grad_csc(int cols, int nnz, const int* __restrict__ ptr, const int* __restrict__ row,
                           int mode, const double* __restrict__ u, const double* __restrict__ v,
                           const double* __restrict__ xs, const double* __restrict__ zs, double inv_tau,
                           const double* __restrict__ drow, const double* __restrict__ dcol, double c,
                           double* __restrict__ g) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nnz) return;
  int lo = 0, hi = cols;  // largest j with ptr[j] <= k
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (ptr[mid] <= k) lo = mid; else hi = mid;
  }
  const int j = lo, i = row[k];
  double val;
  if (mode == 0) {
    val = -(zs[i] * inv_tau * u[j] + v[i] * xs[j] * inv_tau) * drow[i] * dcol[j];
  } else if (i < j) {
    val = -(u[i] * xs[j] + u[j] * xs[i]) * inv_tau * c * drow[i] * dcol[j];
  } else if (i == j) {
    val = -u[i] * xs[i] * inv_tau * c * drow[i] * dcol[j];
  } else {
    val = 0.0;
  }
  g[k] = val;
}

// Tangent correction of the SOC scaling at a (near-)complementary pair, for
// the adjoint system.  At s o z = 0 with both s and z on the cone boundary the
// exact linearization arw(z) ds + arw(s) dz = 0 gives ds_T = -(s0/z0) dz_T on
// the directions T = {(0, t): t orthogonal to z1}; the NT scaling W^2 has the
// eigenvalue beta^2 = sqrt((s0^2-|s1|^2)/(z0^2-|z1|^2)) there, which equals
// s0/z0 only on the central path.  Per block: tcorr = s0/z0 - beta^2 and
// tdir = (0, z1/|z1|); tcorr = 0 where |z1| = 0 or the scaling fell back to I.
__global__ void k_soc_tangent(ConeDev C, const double* __restrict__ s, const double* __restrict__ z,
                              double* tcorr, double* tdir) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= C.nsoc) return;
  const int o = C.soc_off[k], d = C.soc_dim[k];
  double zn = 0.0;
  for (int j = 1; j < d; ++j) zn += z[o + j] * z[o + j];
  zn = sqrt(zn);
  const bool ok = zn > 0.0 && z[o] > 0.0 && !C.ident[k];
  tdir[o] = 0.0;
  for (int j = 1; j < d; ++j) tdir[o + j] = ok ? z[o + j] / zn : 0.0;
  tcorr[k] = ok ? s[o] / z[o] - C.beta[k] * C.beta[k] : 0.0;
}

// out -= tcorr_k * P_T x on every SOC block (the correction term of k_soc_tangent
// applied to a vector; P_T = projection onto (0, t), t orthogonal to tdir).
__global__ void k_soc_tangent_apply(ConeDev C, const double* __restrict__ tcorr,
                                    const double* __restrict__ tdir, const double* __restrict__ x,
                                    double* out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= C.nsoc) return;
  const int o = C.soc_off[k], d = C.soc_dim[k];
  const double c = tcorr[k];
  if (c == 0.0) return;
  double ux = 0.0;
  for (int j = 1; j < d; ++j) ux += tdir[o + j] * x[o + j];
  for (int j = 1; j < d; ++j) out[o + j] -= c * (x[o + j] - tdir[o + j] * ux);
}

inline int g1(long n) { return ceil_div(n > 0 ? n : 1, 256); }
inline int gcap(long n, int cap = 1024) {
  int g = g1(n);
  return g < cap ? g : cap;
}

}  // namespace

void launch_reduce(const RedBatch& b, double* partial, double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_reduce_partial, dim3(RED_BLOCKS), dim3(256), 0, st, b, partial);
  hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(256), 0, st, partial, b.nops, out);
  CX_LAUNCH_CHECK();
}

void launch_cone_scaling(const ConeDev& C, const double* s, const double* z, hipStream_t st) {
  if (C.nn_dim) hipLaunchKernelGGL(k_nn_scaling, dim3(g1(C.nn_dim)), dim3(256), 0, st, C, s, z);
  if (C.nsmall) hipLaunchKernelGGL(k_soc_scaling_small, dim3(g1(C.nsmall)), dim3(256), 0, st, C, s, z);
  if (C.nlarge) hipLaunchKernelGGL(k_soc_scaling_large, dim3(ceil_div(C.nlarge, 4)), dim3(256), 0, st, C, s, z);
  CX_LAUNCH_CHECK();
}

void launch_neg_hs_plus(const ConeDev& C, int m, const double* x, const double* y, double* out,
                        hipStream_t st) {
  long work = (long)C.zero_dim + C.nn_dim + C.nsmall + 64L * C.nlarge;
  (void)m;
  hipLaunchKernelGGL(k_neg_hs_plus, dim3(gcap(work)), dim3(256), 0, st, C, m, x, y, out);
  CX_LAUNCH_CHECK();
}

void launch_combined_rhs(const ConeDev& C, int m, const double* z, const double* ds_aff,
                         const double* dz_aff, const double* rz, double sigma_mu, double one_m_sigma,
                         double* ds_const, double* rhs_tail, hipStream_t st) {
  (void)m;
  long work = (long)C.zero_dim + C.nn_dim + C.nsmall + 64L * C.nlarge;
  hipLaunchKernelGGL(k_combined_rhs, dim3(gcap(work)), dim3(256), 0, st, C, z, ds_aff, dz_aff, rz,
                     sigma_mu, one_m_sigma, ds_const, rhs_tail);
  CX_LAUNCH_CHECK();
}

void launch_max_step(const ConeDev& C, const double* x, const double* dx, double* partial,
                     double* out, hipStream_t st) {
  long work = (long)C.nn_dim + C.nsmall + 64L * C.nlarge;
  int G = gcap(work, 240);
  hipLaunchKernelGGL(k_max_step, dim3(G), dim3(256), 0, st, C, x, dx, partial);
  hipLaunchKernelGGL(k_min_final, dim3(1), dim3(256), 0, st, partial, G, out);
  CX_LAUNCH_CHECK();
}

void launch_margins(const ConeDev& C, const double* x, double* partial, double* out,
                    hipStream_t st) {
  long work = (long)C.nn_dim + C.nsmall + 64L * C.nlarge;
  int G = gcap(work, 240);
  hipLaunchKernelGGL(k_margins, dim3(G), dim3(256), 0, st, C, x, partial);
  hipLaunchKernelGGL(k_margins_final, dim3(1), dim3(256), 0, st, partial, G, C.nn_dim > 0 ? 1 : 0, out);
  CX_LAUNCH_CHECK();
}

void launch_unit_shift(const ConeDev& C, double* x, double alpha, hipStream_t st) {
  int n = C.nn_dim > C.nsoc ? C.nn_dim : C.nsoc;
  if (n == 0) return;
  hipLaunchKernelGGL(k_unit_shift, dim3(g1(n)), dim3(256), 0, st, C, x, alpha);
  CX_LAUNCH_CHECK();
}

void launch_fill(double* x, int n, double v, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fill, dim3(g1(n)), dim3(256), 0, st, x, n, v);
  CX_LAUNCH_CHECK();
}

void launch_csr_mv(int rows, const int* ptr, const int* col, const int* src, const double* val,
                   const double* x, double* y, hipStream_t st) {
  if (rows <= 0) return;
  hipLaunchKernelGGL(k_csr_mv, dim3(g1(rows)), dim3(256), 0, st, rows, ptr, col, src, val, x, y);
  CX_LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------
// Krylov helpers (GMRES on the KKT system with the FP32 factorization as the
// preconditioner): block inner products, Gram-Schmidt update, combination
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_mdot_partial(int n, int k, const double* __restrict__ V, long ld,
                                                      const double* __restrict__ w, double* partial) {
  __shared__ double red[4][KRY_MAX];
  double acc[KRY_MAX];
#pragma unroll
  for (int t = 0; t < KRY_MAX; ++t) acc[t] = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const double wi = w[i];
#pragma unroll
    for (int t = 0; t < KRY_MAX; ++t)
      if (t < k) acc[t] += V[t * ld + i] * wi;
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int t = 0; t < KRY_MAX; ++t) {
    if (t < k) {
      double v = acc[t];
      for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
      if (lane == 0) red[wv][t] = v;
    }
  }
  __syncthreads();
  if (threadIdx.x < k)
    partial[threadIdx.x * gridDim.x + blockIdx.x] =
        (red[0][threadIdx.x] + red[1][threadIdx.x]) + (red[2][threadIdx.x] + red[3][threadIdx.x]);
}

// one block per vector: h[t] = sum of its partials
__global__ __launch_bounds__(256) void k_mdot_final(int nb, const double* __restrict__ partial, double* h) {
  __shared__ double red[256];
  const int t = blockIdx.x;
  double v = 0.0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) v += partial[t * nb + b];
  red[threadIdx.x] = v;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) h[t] = red[0];
}

__global__ __launch_bounds__(256) void k_mgs_update(int n, int k, const double* __restrict__ V, long ld,
                                                    const double* __restrict__ h, double* w) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double acc = w[i];
  for (int t = 0; t < k; ++t) acc -= h[t] * V[t * ld + i];
  w[i] = acc;
}

__global__ __launch_bounds__(256) void k_multi_axpy(int n, int k, const double* __restrict__ Z, long ld,
                                                    KryCoef y, double* x) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double acc = x[i];
  for (int t = 0; t < k; ++t) acc += y.c[t] * Z[t * ld + i];
  x[i] = acc;
}

void launch_mdot(int n, int k, const double* V, long ld, const double* w, double* partial, double* h,
                 hipStream_t st) {
  if (k <= 0 || k > KRY_MAX) return;
  hipLaunchKernelGGL(k_mdot_partial, dim3(KRY_BLOCKS), dim3(256), 0, st, n, k, V, ld, w, partial);
  hipLaunchKernelGGL(k_mdot_final, dim3(k), dim3(256), 0, st, KRY_BLOCKS, partial, h);
  CX_LAUNCH_CHECK();
}
void launch_mgs_update(int n, int k, const double* V, long ld, const double* h, double* w, hipStream_t st) {
  if (n <= 0 || k <= 0) return;
  hipLaunchKernelGGL(k_mgs_update, dim3(g1(n)), dim3(256), 0, st, n, k, V, ld, h, w);
  CX_LAUNCH_CHECK();
}
void launch_multi_axpy(int n, int k, const double* Z, long ld, const KryCoef& y, double* x, hipStream_t st) {
  if (n <= 0 || k <= 0) return;
  hipLaunchKernelGGL(k_multi_axpy, dim3(g1(n)), dim3(256), 0, st, n, k, Z, ld, y, x);
  CX_LAUNCH_CHECK();
}

void launch_axpby(int n, double a, const double* x, double b, double* y, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_axpby, dim3(g1(n)), dim3(256), 0, st, n, a, x, b, y);
  CX_LAUNCH_CHECK();
}
void launch_lin2(int n, double a, const double* x, double b, const double* y, double* out,
                 hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_lin2, dim3(g1(n)), dim3(256), 0, st, n, a, x, b, y, out);
  CX_LAUNCH_CHECK();
}
void launch_lin3(int n, double a, const double* x, double b, const double* y, double c,
                 const double* z, double* out, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_lin3, dim3(g1(n)), dim3(256), 0, st, n, a, x, b, y, c, z, out);
  CX_LAUNCH_CHECK();
}
void launch_scale(int n, double a, double* x, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_scale, dim3(g1(n)), dim3(256), 0, st, n, a, x);
  CX_LAUNCH_CHECK();
}
void launch_copy(int n, const double* x, double* y, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_copy, dim3(g1(n)), dim3(256), 0, st, n, x, y);
  CX_LAUNCH_CHECK();
}
void launch_scatter(int n, const int* slot, const int* src, const double* val, double add,
                    double* Kx, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_scatter, dim3(g1(n)), dim3(256), 0, st, n, slot, src, val, add, Kx);
  CX_LAUNCH_CHECK();
}
void launch_hs_scatter(const ConeDev& C, int nhs, const int* slot, const int* kind, const int* blk,
                       const int* ia, const int* ib, const int* row, double sreg_a, double* Kx,
                       hipStream_t st, const double* tcorr, const double* tdir) {
  if (nhs <= 0) return;
  hipLaunchKernelGGL(k_hs_scatter, dim3(g1(nhs)), dim3(256), 0, st, C, nhs, slot, kind, blk, ia, ib,
                     row, sreg_a, tcorr, tdir, Kx);
  CX_LAUNCH_CHECK();
}
void launch_soc_tangent_apply(const ConeDev& C, const double* tcorr, const double* tdir,
                              const double* x, double* out, hipStream_t st) {
  if (C.nsoc <= 0) return;
  hipLaunchKernelGGL(k_soc_tangent_apply, dim3(g1(C.nsoc)), dim3(256), 0, st, C, tcorr, tdir, x, out);
  CX_LAUNCH_CHECK();
}
void launch_soc_tangent(const ConeDev& C, const double* s, const double* z, double* tcorr, double* tdir,
                        hipStream_t st) {
  if (C.nsoc <= 0) return;
  hipLaunchKernelGGL(k_soc_tangent, dim3(g1(C.nsoc)), dim3(256), 0, st, C, s, z, tcorr, tdir);
  CX_LAUNCH_CHECK();
}
void launch_scale_csc(int cols, const int* ptr, const int* row, const double* in, double c,
                      const double* drow, const double* dcol, double* out, hipStream_t st) {
  if (cols <= 0) return;
  hipLaunchKernelGGL(k_scale_csc, dim3(g1(cols)), dim3(256), 0, st, cols, ptr, row, in, c, drow, dcol,
                     out);
  CX_LAUNCH_CHECK();
}
void launch_vscale(int n, const double* d, double c, const double* in, double* out, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_vscale, dim3(g1(n)), dim3(256), 0, st, n, d, c, in, out);
  CX_LAUNCH_CHECK();
}
void launch_vmuldiv(int n, const double* in, const double* mul, const double* div, double c,
                    double* out, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_vmuldiv, dim3(g1(n)), dim3(256), 0, st, n, in, mul, div, c, out);
  CX_LAUNCH_CHECK();
}

void launch_grad_csc(int cols, int nnz, const int* ptr, const int* row, int mode, const double* u,
                     const double* v, const double* xs, const double* zs, double inv_tau,
                     const double* drow, const double* dcol, double c, double* g, hipStream_t st) {
  if (nnz <= 0 || cols <= 0) return;
  hipLaunchKernelGGL(k_grad_csc, dim3(g1(nnz)), dim3(256), 0, st, cols, nnz, ptr, row, mode, u, v, xs,
                     zs, inv_tau, drow, dcol, c, g);
  CX_LAUNCH_CHECK();
}

}  // namespace cx
