// Device kernels of the interior-point iteration (cones, residuals, reductions).
#pragma once

#include <hip/hip_runtime.h>

namespace cx {

// One reduction term.  kinds:
//   RED_DOT      sum a[i]*b[i]
//   RED_DOT_AXPY sum (a[i] + alpha c[i]) * (b[i] + alpha d[i])
//   RED_NONFIN   count of non-finite a[i]
enum RedKind { RED_DOT = 0, RED_DOT_AXPY = 1, RED_NONFIN = 2 };
struct RedOp {
  int kind;
  int len;
  const double* a;
  const double* b;
  const double* c;
  const double* d;
  double alpha;
};
constexpr int RED_MAX_OPS = 8;
constexpr int RED_BLOCKS = 240;
struct RedBatch {
  int nops;
  RedOp op[RED_MAX_OPS];
};

// cone layout for device kernels
struct ConeDev {
  int zero_off, zero_dim;
  int nn_off, nn_dim;
  int nsoc;
  const int* soc_off;   // nsoc
  const int* soc_dim;   // nsoc
  const int* small_list;  // SOC blocks handled one per thread
  int nsmall;
  const int* large_list;  // SOC blocks handled one per wave
  int nlarge;
  // NT scaling state
  double* beta;  // nsoc
  double* vv;    // nsoc: v'v
  int* ident;    // nsoc: identity fallback
  double* v;     // m: SOC scaling vectors at their offsets
  double* w;     // m: nonnegative w = sqrt(s/z)
  double* hsd;   // m: nonnegative Hs diagonal = w^2
};

void launch_reduce(const RedBatch& b, double* partial, double* out, hipStream_t st);

void launch_cone_scaling(const ConeDev& C, const double* s, const double* z, hipStream_t st);
// out = -(Hs x + y)
void launch_neg_hs_plus(const ConeDev& C, int m, const double* x, const double* y, double* out,
                        hipStream_t st);
// Mehrotra combined-step cone right-hand side: ds_const and rhs_tail = ds_const - (1-sigma) rz
void launch_combined_rhs(const ConeDev& C, int m, const double* z, const double* ds_aff,
                         const double* dz_aff, const double* rz, double sigma_mu, double one_m_sigma,
                         double* ds_const, double* rhs_tail, hipStream_t st);
// min over cone blocks of the step to the boundary (candidates only; alpha cap applied on host)
void launch_max_step(const ConeDev& C, const double* x, const double* dx, double* partial,
                     double* out, hipStream_t st);
// (min margin, sum of positive margins) over non-zero blocks; out[0..1]
void launch_margins(const ConeDev& C, const double* x, double* partial, double* out,
                    hipStream_t st);
void launch_unit_shift(const ConeDev& C, double* x, double alpha, hipStream_t st);
void launch_fill(double* x, int n, double v, hipStream_t st);

// Sparse kernels (CSR by rows; value arrays addressed through src maps)
void launch_csr_mv(int rows, const int* ptr, const int* col, const int* src, const double* val,
                   const double* x, double* y, hipStream_t st);
// P upper given in CSC: y = P x using a full symmetric CSR map
// generic elementwise helpers
// y = a*x + b*y  (len n)
void launch_axpby(int n, double a, const double* x, double b, double* y, hipStream_t st);
// out = a*x + b*y
void launch_lin2(int n, double a, const double* x, double b, const double* y, double* out,
                 hipStream_t st);
// out = a*x + b*y + c*z
void launch_lin3(int n, double a, const double* x, double b, const double* y, double c,
                 const double* z, double* out, hipStream_t st);
void launch_scale(int n, double a, double* x, hipStream_t st);
// Krylov helpers over k <= KRY_MAX vectors of length n stored ld apart (V + t*ld)
constexpr int KRY_MAX = 32;
constexpr int KRY_BLOCKS = 240;
struct KryCoef { double c[KRY_MAX]; };
// h[t] = V_t . w  (partial: KRY_MAX * KRY_BLOCKS scratch)
void launch_mdot(int n, int k, const double* V, long ld, const double* w, double* partial, double* h,
                 hipStream_t st);
// w -= sum_t h[t] V_t  (h on the device)
void launch_mgs_update(int n, int k, const double* V, long ld, const double* h, double* w, hipStream_t st);
// x += sum_t y.c[t] Z_t
void launch_multi_axpy(int n, int k, const double* Z, long ld, const KryCoef& y, double* x, hipStream_t st);
void launch_copy(int n, const double* x, double* y, hipStream_t st);
// Kx[slot[i]] = val[src[i]] + add (src < 0 -> add only)
void launch_scatter(int n, const int* slot, const int* src, const double* val, double add,
                    double* Kx, hipStream_t st);
// Hs slots of the KKT (2,2) block
void launch_hs_scatter(const ConeDev& C, int nhs, const int* slot, const int* kind, const int* blk,
                       const int* ia, const int* ib, const int* row, double sreg_a, double* Kx,
                       hipStream_t st, const double* tcorr = nullptr, const double* tdir = nullptr);
// out -= tcorr_k P_T x per SOC block (see k_soc_tangent_apply)
void launch_soc_tangent_apply(const ConeDev& C, const double* tcorr, const double* tdir,
                              const double* x, double* out, hipStream_t st);
// SOC tangent correction for the adjoint system (see k_soc_tangent)
void launch_soc_tangent(const ConeDev& C, const double* s, const double* z, double* tcorr, double* tdir,
                        hipStream_t st);
// equilibration of values on device: out[k] = in[k] * c * d[r[k]] * d[col(k)]
void launch_scale_csc(int cols, const int* ptr, const int* row, const double* in, double c,
                      const double* drow, const double* dcol, double* out, hipStream_t st);
// x[i] *= d[i]*c
void launch_vscale(int n, const double* d, double c, const double* in, double* out, hipStream_t st);
// out[i] = in[i] * mul[i] / div[i]
void launch_vmuldiv(int n, const double* in, const double* mul, const double* div, double c,
                    double* out, hipStream_t st);

// Adjoint gradients on a CSC pattern (mode 0: A, mode 1: P upper); see k_grad_csc.
void launch_grad_csc(int cols, int nnz, const int* ptr, const int* row, int mode, const double* u,
                     const double* v, const double* xs, const double* zs, double inv_tau,
                     const double* drow, const double* dcol, double c, double* g, hipStream_t st);

}  // namespace cx
