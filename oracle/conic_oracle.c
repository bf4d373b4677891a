/*
 * conic_oracle.c -- CPU oracle (test infrastructure only; see conic_oracle.h).
 *
 * Plain C restatement of the reference conic IPM.  Each function names the
 * reference file:line it restates.  Dense per-block SOC matrices and an LU
 * for W^-1 are kept exactly as the reference keeps them (PartialPivLU); the
 * nonnegative-orthant Hs is stored as its diagonal (the reference's dense
 * l x l matrix has exact-zero off-diagonals, so every product is the same).
 */
#include "conic_oracle.h"

#include <float.h>
#include <tgmath.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

enum { ST_UNSOLVED = 0, ST_SOLVED, ST_PRIMAL_INF, ST_DUAL_INF, ST_MAXIT, ST_NUMERR };
enum { CT_ZERO = 0, CT_NONNEG = 1, CT_SOC = 2 };
enum { LS_EIGEN = 0, LS_QDLDL = 1, LS_REG = 2 };

#define OC_INF (1.0 / 0.0)

static oc_real dmax(oc_real a, oc_real b) { return a > b ? a : b; }
static oc_real dmin(oc_real a, oc_real b) { return a < b ? a : b; }
static void* xcalloc(size_t n, size_t sz) { return calloc(n ? n : 1, sz); }

/* ------------------------------------------------------------------ */
/* Settings defaults: include/conicxx/settings.h:77-118                */
/* ------------------------------------------------------------------ */
void oc_default_settings(oc_settings* s) {
  s->tol_feas = 1e-8;
  s->tol_gap_abs = 1e-8;
  s->tol_gap_rel = 1e-8;
  s->tol_infeas = 1e-7;
  s->max_iter = 200;
  s->static_reg_P = 1e-8;
  s->static_reg_A = 1e-8;
  s->dynamic_reg_eps = 1e-14;
  s->dynamic_reg_delta = 1e-7;
  s->refine_max_iter = 3;
  s->refine_tol = 1e-12;
  s->max_step_fraction = 0.99;
  s->equilibrate = 1;
  s->equilibrate_max_iter = 10;
  s->equilibrate_min_scale = 1e-4;
  s->equilibrate_max_scale = 1e4;
  s->warm_start = 1;
  s->verbose = 0;
  s->record_timings = 0;
  s->validate_inputs = 1;
  s->linear_solver = LS_QDLDL;
}

/* ------------------------------------------------------------------ */
/* Small dense helpers                                                 */
/* ------------------------------------------------------------------ */
static oc_real vdot(int n, const oc_real* a, const oc_real* b) {
  oc_real s = 0.0;
  for (int i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}
static oc_real vnorm(int n, const oc_real* a) { return sqrt(vdot(n, a, a)); }
static int vfinite(int n, const oc_real* a) {
  for (int i = 0; i < n; ++i)
    if (!isfinite(a[i])) return 0;
  return 1;
}

/* LU with partial pivoting of a dense d x d column-major matrix (the role
 * Eigen::PartialPivLU plays for W at include/conicxx/cones/second_order_cone.h:59,
 * src/conicxx/cones/second_order_cone.cpp:158,165-167). */
static void lu_factor(int d, oc_real* LU, int* piv) {
  for (int k = 0; k < d; ++k) {
    int p = k;
    oc_real best = fabs(LU[k + k * d]);
    for (int i = k + 1; i < d; ++i)
      if (fabs(LU[i + k * d]) > best) best = fabs(LU[i + k * d]), p = i;
    piv[k] = p;
    if (p != k)
      for (int j = 0; j < d; ++j) {
        oc_real t = LU[k + j * d];
        LU[k + j * d] = LU[p + j * d];
        LU[p + j * d] = t;
      }
    oc_real pv = LU[k + k * d];
    if (pv != 0.0)
      for (int i = k + 1; i < d; ++i) LU[i + k * d] /= pv;
    for (int j = k + 1; j < d; ++j) {
      oc_real f = LU[k + j * d];
      if (f != 0.0)
        for (int i = k + 1; i < d; ++i) LU[i + j * d] -= LU[i + k * d] * f;
    }
  }
}
static void lu_solve(int d, const oc_real* LU, const int* piv, const oc_real* b, oc_real* x) {
  for (int i = 0; i < d; ++i) x[i] = b[i];
  for (int k = 0; k < d; ++k) {
    int p = piv[k];
    if (p != k) {
      oc_real t = x[k];
      x[k] = x[p];
      x[p] = t;
    }
  }
  for (int k = 0; k < d; ++k)
    for (int i = k + 1; i < d; ++i) x[i] -= LU[i + k * d] * x[k];
  for (int k = d - 1; k >= 0; --k) {
    x[k] /= LU[k + k * d];
    for (int i = 0; i < k; ++i) x[i] -= LU[i + k * d] * x[k];
  }
}

/* ------------------------------------------------------------------ */
/* Cones: include/conicxx/cones/ and src/conicxx/cones/ (all files)      */
/* ------------------------------------------------------------------ */
typedef struct {
  int type, off, dim;
  oc_real* W;   /* SOC: d*d */
  oc_real* Hs;  /* SOC: d*d ; NONNEG: d (diagonal) */
  oc_real* LU;  /* SOC: d*d */
  int* piv;    /* SOC: d */
  oc_real* w;   /* NONNEG: d */
} oc_block;

typedef struct {
  int nblocks, total, degree;
  oc_block* b;
  oc_real* identity;
} oc_cones;

/* ConeSet ctor: src/conicxx/cones/cone_set.cpp:12-32 (fixed order Zero, Nonneg, SOC...). */
static void cones_init(oc_cones* C, int zero_dim, int nonneg_dim, int n_soc, const int* soc_dims) {
  memset(C, 0, sizeof(*C));
  C->b = (oc_block*)xcalloc((size_t)(2 + n_soc), sizeof(oc_block));
  int nb = 0, off = 0;
  if (zero_dim > 0) {
    C->b[nb].type = CT_ZERO, C->b[nb].off = off, C->b[nb].dim = zero_dim;
    off += zero_dim, ++nb;
  }
  if (nonneg_dim > 0) {
    oc_block* B = &C->b[nb];
    B->type = CT_NONNEG, B->off = off, B->dim = nonneg_dim;
    B->w = (oc_real*)xcalloc((size_t)nonneg_dim, sizeof(oc_real));
    B->Hs = (oc_real*)xcalloc((size_t)nonneg_dim, sizeof(oc_real));
    for (int i = 0; i < nonneg_dim; ++i) B->w[i] = 1.0, B->Hs[i] = 1.0;
    C->degree += nonneg_dim;
    off += nonneg_dim, ++nb;
  }
  for (int k = 0; k < n_soc; ++k) {
    oc_block* B = &C->b[nb];
    int d = soc_dims[k];
    B->type = CT_SOC, B->off = off, B->dim = d;
    B->W = (oc_real*)xcalloc((size_t)d * d, sizeof(oc_real));
    B->Hs = (oc_real*)xcalloc((size_t)d * d, sizeof(oc_real));
    B->LU = (oc_real*)xcalloc((size_t)d * d, sizeof(oc_real));
    B->piv = (int*)xcalloc((size_t)d, sizeof(int));
    for (int i = 0; i < d; ++i) B->W[i + i * d] = 1.0, B->Hs[i + i * d] = 1.0, B->LU[i + i * d] = 1.0, B->piv[i] = i;
    C->degree += 1;
    off += d, ++nb;
  }
  C->nblocks = nb;
  C->total = off;
  C->identity = (oc_real*)xcalloc((size_t)off, sizeof(oc_real));
  for (int i = 0; i < nb; ++i) {
    oc_block* B = &C->b[i];
    if (B->type == CT_NONNEG)
      for (int j = 0; j < B->dim; ++j) C->identity[B->off + j] = 1.0;
    if (B->type == CT_SOC) C->identity[B->off] = 1.0;
  }
}

static void cones_free(oc_cones* C) {
  for (int i = 0; i < C->nblocks; ++i) {
    free(C->b[i].W), free(C->b[i].Hs), free(C->b[i].LU), free(C->b[i].piv), free(C->b[i].w);
  }
  free(C->b);
  free(C->identity);
  memset(C, 0, sizeof(*C));
}

/* SecondOrderCone::product: second_order_cone.cpp:43-48 */
void oc_soc_product(int d, const oc_real* u, const oc_real* v, oc_real* out) {
  oc_real o0 = vdot(d, u, v);
  for (int i = 1; i < d; ++i) out[i] = u[0] * v[i] + v[0] * u[i];
  out[0] = o0;
}

/* SecondOrderCone::inverseProduct: second_order_cone.cpp:50-68 (kTiny = 1e-12 floors) */
void oc_soc_inverse_product(int d, const oc_real* u, const oc_real* w, oc_real* out) {
  const oc_real kTiny = 1e-12;
  oc_real u0 = u[0];
  oc_real rho = u0 * u0 - vdot(d - 1, u + 1, u + 1);
  if (fabs(rho) < kTiny) rho = (rho >= 0 ? kTiny : -kTiny);
  if (fabs(u0) < kTiny) u0 = kTiny;
  oc_real nu = vdot(d - 1, u + 1, w + 1);
  oc_real w0 = w[0];
  out[0] = (u0 * w0 - nu) / rho;
  for (int i = 1; i < d; ++i) out[i] = ((nu / u0 - w0) * u[i] + (rho / u0) * w[i]) / rho;
}

/* SecondOrderCone::updateScaling: second_order_cone.cpp:70-159 */
static void soc_scaling_into(int d, const oc_real* s, const oc_real* z, oc_real* W, oc_real* Hs) {
  const oc_real kTiny = 1e-12, kTinyRel = 1e-12, kAbsFloor = DBL_MIN * 1e4;
  const oc_real kBetaMax = 1e8, kWMax = 1e8;
  oc_real s0 = s[0], z0 = z[0];
  oc_real s1n = vdot(d - 1, s + 1, s + 1), z1n = vdot(d - 1, z + 1, z + 1);
  oc_real a = sqrt(dmax(dmax(s0 * s0 - s1n, kTinyRel * s0 * s0), kAbsFloor));
  oc_real b = sqrt(dmax(dmax(z0 * z0 - z1n, kTinyRel * z0 * z0), kAbsFloor));
  oc_real beta = dmin(dmax(sqrt(a / b), 1.0 / kBetaMax), kBetaMax);
  oc_real c = sqrt(dmax((vdot(d, s, z) / (a * b) + 1.0) / 2.0, kTiny));
  oc_real* v = (oc_real*)xcalloc((size_t)d, sizeof(oc_real));
  v[0] = z0 / b;
  for (int i = 1; i < d; ++i) v[i] = -z[i] / b;
  v[0] += s0 / a;
  for (int i = 1; i < d; ++i) v[i] += s[i] / a;
  for (int i = 0; i < d; ++i) v[i] /= (2.0 * c);
  v[0] += 1.0;
  v[0] = dmax(v[0], kTiny);
  oc_real sc = sqrt(2.0 * v[0]);
  for (int i = 0; i < d; ++i) v[i] /= sc;
  for (int j = 0; j < d; ++j)
    for (int i = 0; i < d; ++i) W[i + j * d] = 2.0 * (v[i] * v[j]);
  W[0] -= 1.0;
  for (int i = 1; i < d; ++i) W[i + i * d] += 1.0;
  for (int k = 0; k < d * d; ++k) W[k] *= beta;
  for (int j = 0; j < d; ++j)
    for (int i = 0; i < d; ++i) {
      oc_real acc = 0.0;
      for (int k = 0; k < d; ++k) acc += W[i + k * d] * W[k + j * d];
      Hs[i + j * d] = acc;
    }
  int bad = !vfinite(d * d, W) || !vfinite(d * d, Hs);
  oc_real wmax = 0.0;
  for (int k = 0; k < d * d; ++k) wmax = dmax(wmax, fabs(W[k]));
  if (bad || wmax > kWMax) {
    for (int k = 0; k < d * d; ++k) W[k] = 0.0, Hs[k] = 0.0;
    for (int i = 0; i < d; ++i) W[i + i * d] = 1.0, Hs[i + i * d] = 1.0;
  }
  free(v);
}

void oc_soc_scaling(int d, const oc_real* s, const oc_real* z, oc_real* W, oc_real* Hs) {
  soc_scaling_into(d, s, z, W, Hs);
}

/* SecondOrderCone::maxStep: second_order_cone.cpp:174-215 */
oc_real oc_soc_max_step(int d, const oc_real* x, const oc_real* dx, oc_real alpha_max) {
  const oc_real kEps = DBL_MIN * 1e4;
  oc_real p0 = x[0], d0 = dx[0];
  oc_real a = d0 * d0 - vdot(d - 1, dx + 1, dx + 1);
  oc_real b = 2.0 * (p0 * d0 - vdot(d - 1, x + 1, dx + 1));
  oc_real c = p0 * p0 - vdot(d - 1, x + 1, x + 1);
  oc_real alpha = alpha_max;
#define FEAS(r) ((r) > 0 && (p0 + (r) * d0) >= 0)
  if (fabs(a) < kEps) {
    if (fabs(b) > kEps) {
      oc_real r = -c / b;
      if (FEAS(r)) alpha = dmin(alpha, r);
    }
    return alpha;
  }
  oc_real disc = b * b - 4.0 * a * c;
  if (disc < 0) return alpha;
  oc_real sq = sqrt(disc);
  oc_real qq = (b >= 0) ? -0.5 * (b + sq) : -0.5 * (b - sq);
  oc_real r1, r2;
  if (fabs(qq) > kEps) {
    r1 = qq / a;
    r2 = c / qq;
  } else {
    r1 = r2 = -b / (2.0 * a);
  }
  if (FEAS(r1)) alpha = dmin(alpha, r1);
  if (FEAS(r2)) alpha = dmin(alpha, r2);
#undef FEAS
  return alpha;
}

/* ConeSet::updateScaling: cone_set.cpp:48-54; Nonneg: nonnegative_cone.cpp:22-26 */
static void cones_update_scaling(oc_cones* C, const oc_real* s, const oc_real* z) {
  for (int i = 0; i < C->nblocks; ++i) {
    oc_block* B = &C->b[i];
    const oc_real* sb = s + B->off;
    const oc_real* zb = z + B->off;
    if (B->type == CT_NONNEG) {
      for (int j = 0; j < B->dim; ++j) {
        B->w[j] = sqrt(sb[j] / zb[j]);
        B->Hs[j] = B->w[j] * B->w[j];
      }
    } else if (B->type == CT_SOC) {
      int d = B->dim;
      soc_scaling_into(d, sb, zb, B->W, B->Hs);
      memcpy(B->LU, B->W, sizeof(oc_real) * (size_t)d * d);
      lu_factor(d, B->LU, B->piv);
    }
  }
}

/* ConeSet::product / inverseProduct: cone_set.cpp:34-46 (Zero -> 0, Nonneg elementwise) */
static void cones_product(const oc_cones* C, const oc_real* u, const oc_real* v, oc_real* out) {
  for (int i = 0; i < C->nblocks; ++i) {
    const oc_block* B = &C->b[i];
    int o = B->off, d = B->dim;
    if (B->type == CT_ZERO)
      for (int j = 0; j < d; ++j) out[o + j] = 0.0;
    else if (B->type == CT_NONNEG)
      for (int j = 0; j < d; ++j) out[o + j] = u[o + j] * v[o + j];
    else
      oc_soc_product(d, u + o, v + o, out + o);
  }
}
static void cones_inverse_product(const oc_cones* C, const oc_real* u, const oc_real* w, oc_real* out) {
  for (int i = 0; i < C->nblocks; ++i) {
    const oc_block* B = &C->b[i];
    int o = B->off, d = B->dim;
    if (B->type == CT_ZERO)
      for (int j = 0; j < d; ++j) out[o + j] = 0.0;
    else if (B->type == CT_NONNEG)
      for (int j = 0; j < d; ++j) out[o + j] = w[o + j] / u[o + j];
    else
      oc_soc_inverse_product(d, u + o, w + o, out + o);
  }
}

/* ConeSet::applyW / applyWInv / mulHs: cone_set.cpp:56-75 */
static void cones_apply_w(const oc_cones* C, const oc_real* x, oc_real* out) {
  for (int i = 0; i < C->nblocks; ++i) {
    const oc_block* B = &C->b[i];
    int o = B->off, d = B->dim;
    if (B->type == CT_ZERO)
      for (int j = 0; j < d; ++j) out[o + j] = 0.0;
    else if (B->type == CT_NONNEG)
      for (int j = 0; j < d; ++j) out[o + j] = B->w[j] * x[o + j];
    else
      for (int r = 0; r < d; ++r) {
        oc_real acc = 0.0;
        for (int k = 0; k < d; ++k) acc += B->W[r + k * d] * x[o + k];
        out[o + r] = acc;
      }
  }
}
static void cones_apply_winv(const oc_cones* C, const oc_real* x, oc_real* out) {
  for (int i = 0; i < C->nblocks; ++i) {
    const oc_block* B = &C->b[i];
    int o = B->off, d = B->dim;
    if (B->type == CT_ZERO)
      for (int j = 0; j < d; ++j) out[o + j] = 0.0;
    else if (B->type == CT_NONNEG)
      for (int j = 0; j < d; ++j) out[o + j] = x[o + j] / B->w[j];
    else
      lu_solve(d, B->LU, B->piv, x + o, out + o);
  }
}
static void cones_mul_hs(const oc_cones* C, const oc_real* x, oc_real* out) {
  for (int i = 0; i < C->nblocks; ++i) {
    const oc_block* B = &C->b[i];
    int o = B->off, d = B->dim;
    if (B->type == CT_ZERO)
      for (int j = 0; j < d; ++j) out[o + j] = 0.0;
    else if (B->type == CT_NONNEG)
      for (int j = 0; j < d; ++j) out[o + j] = B->Hs[j] * x[o + j];
    else
      for (int r = 0; r < d; ++r) {
        oc_real acc = 0.0;
        for (int k = 0; k < d; ++k) acc += B->Hs[r + k * d] * x[o + k];
        out[o + r] = acc;
      }
  }
}

/* ConeSet::margins: cone_set.cpp:77-90; per-block margins: nonnegative_cone.h:28,
 * second_order_cone.cpp:169-172, zero_cone.h (+inf). */
static void cones_margins(const oc_cones* C, const oc_real* x, oc_real* min_margin, oc_real* pos_sum) {
  oc_real mn = OC_INF, ps = 0.0;
  for (int i = 0; i < C->nblocks; ++i) {
    const oc_block* B = &C->b[i];
    const oc_real* xb = x + B->off;
    oc_real mg;
    if (B->type == CT_ZERO) continue;
    if (B->type == CT_NONNEG) {
      mg = xb[0];
      for (int j = 1; j < B->dim; ++j) mg = dmin(mg, xb[j]);
    } else {
      mg = xb[0] - vnorm(B->dim - 1, xb + 1);
    }
    if (isfinite(mg)) {
      mn = dmin(mn, mg);
      ps += dmax(mg, 0.0);
    }
  }
  if (!isfinite(mn)) mn = 0.0;
  *min_margin = mn;
  *pos_sum = ps;
}

/* ConeSet::scaledUnitShift: cone_set.cpp:92-98 */
static void cones_unit_shift(const oc_cones* C, oc_real* x, oc_real alpha) {
  for (int i = 0; i < C->nblocks; ++i) {
    const oc_block* B = &C->b[i];
    if (B->type == CT_NONNEG)
      for (int j = 0; j < B->dim; ++j) x[B->off + j] += alpha;
    else if (B->type == CT_SOC)
      x[B->off] += alpha;
  }
}

/* ConeSet::zeroPrimalZeroConeBlocks: cone_set.cpp:100-106 */
static void cones_zero_zero_blocks(const oc_cones* C, oc_real* s) {
  for (int i = 0; i < C->nblocks; ++i)
    if (C->b[i].type == CT_ZERO)
      for (int j = 0; j < C->b[i].dim; ++j) s[C->b[i].off + j] = 0.0;
}

/* ConeSet::maxStep: cone_set.cpp:108-115; Nonneg: nonnegative_cone.cpp:40-49 */
static oc_real cones_max_step(const oc_cones* C, const oc_real* x, const oc_real* dx, oc_real amax) {
  oc_real alpha = amax;
  for (int i = 0; i < C->nblocks; ++i) {
    const oc_block* B = &C->b[i];
    int o = B->off;
    if (B->type == CT_NONNEG) {
      for (int j = 0; j < B->dim; ++j)
        if (dx[o + j] < 0) alpha = dmin(alpha, -x[o + j] / dx[o + j]);
    } else if (B->type == CT_SOC) {
      alpha = oc_soc_max_step(B->dim, x + o, dx + o, alpha);
    }
  }
  return alpha;
}

/* ------------------------------------------------------------------ */
/* CSC helpers                                                         */
/* ------------------------------------------------------------------ */
typedef struct {
  int rows, cols, nnz;
  int* p;
  int* i;
  oc_real* x;
} oc_csc;

static void csc_copy(oc_csc* M, int rows, int cols, const int* p, const int* i, const oc_real* x) {
  M->rows = rows, M->cols = cols;
  M->nnz = p ? p[cols] : 0;
  M->p = (int*)xcalloc((size_t)cols + 1, sizeof(int));
  M->i = (int*)xcalloc((size_t)M->nnz, sizeof(int));
  M->x = (oc_real*)xcalloc((size_t)M->nnz, sizeof(oc_real));
  if (p) {
    memcpy(M->p, p, sizeof(int) * ((size_t)cols + 1));
    memcpy(M->i, i, sizeof(int) * (size_t)M->nnz);
    memcpy(M->x, x, sizeof(oc_real) * (size_t)M->nnz);
  }
}
static void csc_free(oc_csc* M) {
  free(M->p), free(M->i), free(M->x);
  memset(M, 0, sizeof(*M));
}
/* Eigen SparseMatrix::coeff(r, c) on a sorted column (binary search; 0 when absent). */
static oc_real csc_coeff(const oc_csc* M, int r, int c) {
  int lo = M->p[c], hi = M->p[c + 1];
  while (lo < hi) {
    int mid = (lo + hi) / 2;
    if (M->i[mid] < r)
      lo = mid + 1;
    else
      hi = mid;
  }
  if (lo < M->p[c + 1] && M->i[lo] == r) return M->x[lo];
  return 0.0;
}

/* P (symmetric, upper stored) * v: solver_impl.cpp:102 (selfadjointView<Upper>) */
static void pmul(const oc_csc* P, const oc_real* v, oc_real* out) {
  for (int i = 0; i < P->rows; ++i) out[i] = 0.0;
  for (int c = 0; c < P->cols; ++c)
    for (int k = P->p[c]; k < P->p[c + 1]; ++k) {
      int r = P->i[k];
      if (r > c) continue;
      out[r] += P->x[k] * v[c];
      if (r < c) out[c] += P->x[k] * v[r];
    }
}

/* ------------------------------------------------------------------ */
/* Equilibration: src/conicxx/kkt/equilibration.cpp:8-139              */
/* ------------------------------------------------------------------ */
typedef struct {
  int n, m;
  oc_real* d;
  oc_real* e;
  oc_real c;
} oc_equil;

static void equil_compute(oc_equil* E, const oc_csc* P, const oc_csc* A, const oc_cones* C,
                          const oc_settings* st) {
  int n = P->rows, m = A->rows;
  E->n = n, E->m = m;
  free(E->d), free(E->e);
  E->d = (oc_real*)xcalloc((size_t)n, sizeof(oc_real));
  E->e = (oc_real*)xcalloc((size_t)m, sizeof(oc_real));
  for (int i = 0; i < n; ++i) E->d[i] = 1.0;
  for (int i = 0; i < m; ++i) E->e[i] = 1.0;
  E->c = 1.0;
  if (!st->equilibrate || n == 0) return;
  oc_real* Pw = (oc_real*)xcalloc((size_t)P->nnz, sizeof(oc_real));
  oc_real* Aw = (oc_real*)xcalloc((size_t)A->nnz, sizeof(oc_real));
  memcpy(Pw, P->x, sizeof(oc_real) * (size_t)P->nnz);
  memcpy(Aw, A->x, sizeof(oc_real) * (size_t)A->nnz);
  oc_real* nx = (oc_real*)xcalloc((size_t)n, sizeof(oc_real));
  oc_real* nz = (oc_real*)xcalloc((size_t)m, sizeof(oc_real));
  oc_real* dx = (oc_real*)xcalloc((size_t)n, sizeof(oc_real));
  oc_real* dz = (oc_real*)xcalloc((size_t)m, sizeof(oc_real));
  const oc_real minS = st->equilibrate_min_scale, maxS = st->equilibrate_max_scale;
  for (int it = 0; it < st->equilibrate_max_iter; ++it) {
    for (int i = 0; i < n; ++i) nx[i] = 0.0;
    for (int i = 0; i < m; ++i) nz[i] = 0.0;
    for (int c = 0; c < P->cols; ++c)
      for (int k = P->p[c]; k < P->p[c + 1]; ++k) {
        int r = P->i[k];
        oc_real av = fabs(Pw[k]);
        nx[r] = dmax(nx[r], av);
        nx[c] = dmax(nx[c], av);
      }
    for (int c = 0; c < A->cols; ++c)
      for (int k = A->p[c]; k < A->p[c + 1]; ++k) {
        int r = A->i[k];
        oc_real av = fabs(Aw[k]);
        nx[c] = dmax(nx[c], av);
        nz[r] = dmax(nz[r], av);
      }
    for (int bi = 0; bi < C->nblocks; ++bi) {
      const oc_block* B = &C->b[bi];
      if (B->type != CT_SOC) continue;
      oc_real mx = nz[B->off];
      for (int j = 1; j < B->dim; ++j) mx = dmax(mx, nz[B->off + j]);
      for (int j = 0; j < B->dim; ++j) nz[B->off + j] = mx;
    }
    for (int i = 0; i < n; ++i) dx[i] = 1.0 / sqrt(nx[i] > 0 ? nx[i] : 1.0);
    for (int i = 0; i < m; ++i) dz[i] = 1.0 / sqrt(nz[i] > 0 ? nz[i] : 1.0);
    for (int c = 0; c < P->cols; ++c)
      for (int k = P->p[c]; k < P->p[c + 1]; ++k) Pw[k] *= dx[P->i[k]] * dx[c];
    for (int c = 0; c < A->cols; ++c)
      for (int k = A->p[c]; k < A->p[c + 1]; ++k) Aw[k] *= dz[A->i[k]] * dx[c];
    for (int i = 0; i < n; ++i) E->d[i] *= dx[i];
    for (int i = 0; i < m; ++i) E->e[i] *= dz[i];
  }
  for (int i = 0; i < n; ++i) E->d[i] = dmin(dmax(E->d[i], minS), maxS);
  for (int i = 0; i < m; ++i) E->e[i] = dmin(dmax(E->e[i], minS), maxS);
  oc_real sumAbs = 0.0;
  int count = P->nnz;
  for (int k = 0; k < P->nnz; ++k) sumAbs += fabs(Pw[k]);
  if (count > 0) {
    oc_real meanAbs = sumAbs / (oc_real)count;
    if (meanAbs > 0) E->c = 1.0 / dmax(meanAbs, minS);
  }
  E->c = dmin(dmax(E->c, minS), maxS);
  free(Pw), free(Aw), free(nx), free(nz), free(dx), free(dz);
}
static void equil_scale_p(const oc_equil* E, oc_csc* P) {
  for (int c = 0; c < P->cols; ++c)
    for (int k = P->p[c]; k < P->p[c + 1]; ++k) P->x[k] *= E->c * E->d[P->i[k]] * E->d[c];
}
static void equil_scale_a(const oc_equil* E, oc_csc* A) {
  for (int c = 0; c < A->cols; ++c)
    for (int k = A->p[c]; k < A->p[c + 1]; ++k) A->x[k] *= E->e[A->i[k]] * E->d[c];
}
static void equil_scale_q(const oc_equil* E, oc_real* q) {
  for (int i = 0; i < E->n; ++i) q[i] *= E->c * E->d[i];
}
static void equil_scale_b(const oc_equil* E, oc_real* b) {
  for (int i = 0; i < E->m; ++i) b[i] *= E->e[i];
}
static void equil_unscale_solution(const oc_equil* E, oc_real* x, oc_real* s, oc_real* z) {
  for (int i = 0; i < E->n; ++i) x[i] *= E->d[i];
  for (int i = 0; i < E->m; ++i) z[i] = z[i] * E->e[i] / E->c;
  for (int i = 0; i < E->m; ++i) s[i] /= E->e[i];
}
static void equil_scale_solution(const oc_equil* E, oc_real* x, oc_real* s, oc_real* z) {
  for (int i = 0; i < E->n; ++i) x[i] /= E->d[i];
  for (int i = 0; i < E->m; ++i) z[i] = z[i] * E->c / E->e[i];
  for (int i = 0; i < E->m; ++i) s[i] *= E->e[i];
}

/* ------------------------------------------------------------------ */
/* Ordering: exact minimum degree on the explicit elimination graph.   */
/* Replaces Eigen::AMDOrdering (qdldl_ldlt.cpp:31-34); ordering only    */
/* changes rounding.  Ties broken by lowest index.                      */
/* ------------------------------------------------------------------ */
static void min_degree_order(int N, const int* Kp, const int* Ki, int* perm) {
  size_t words = ((size_t)N + 63) / 64;
  uint64_t* adj = (uint64_t*)xcalloc((size_t)N * words, sizeof(uint64_t));
  int* deg = (int*)xcalloc((size_t)N, sizeof(int));
  char* done = (char*)xcalloc((size_t)N, 1);
  int* nbr = (int*)xcalloc((size_t)N, sizeof(int));
  for (int c = 0; c < N; ++c)
    for (int k = Kp[c]; k < Kp[c + 1]; ++k) {
      int r = Ki[k];
      if (r == c) continue;
      adj[(size_t)r * words + (size_t)c / 64] |= (uint64_t)1 << (c % 64);
      adj[(size_t)c * words + (size_t)r / 64] |= (uint64_t)1 << (r % 64);
    }
  for (int v = 0; v < N; ++v) {
    int dg = 0;
    for (size_t w = 0; w < words; ++w) dg += __builtin_popcountll(adj[(size_t)v * words + w]);
    deg[v] = dg;
  }
  for (int step = 0; step < N; ++step) {
    int p = -1, best = N + 1;
    for (int v = 0; v < N; ++v)
      if (!done[v] && deg[v] < best) best = deg[v], p = v;
    perm[step] = p;
    done[p] = 1;
    uint64_t* ap = adj + (size_t)p * words;
    int nn = 0;
    for (size_t w = 0; w < words; ++w) {
      uint64_t bits = ap[w];
      while (bits) {
        int t = __builtin_ctzll(bits);
        bits &= bits - 1;
        nbr[nn++] = (int)(w * 64 + (size_t)t);
      }
    }
    for (int a = 0; a < nn; ++a) {
      int i = nbr[a];
      uint64_t* ai = adj + (size_t)i * words;
      for (size_t w = 0; w < words; ++w) ai[w] |= ap[w];
      ai[i / 64] &= ~((uint64_t)1 << (i % 64));
      ai[p / 64] &= ~((uint64_t)1 << (p % 64));
      int dg = 0;
      for (size_t w = 0; w < words; ++w) dg += __builtin_popcountll(ai[w]);
      deg[i] = dg;
    }
    for (size_t w = 0; w < words; ++w) ap[w] = 0;
  }
  free(adj), free(deg), free(done), free(nbr);
}

/* ------------------------------------------------------------------ */
/* Sparse LDL^T (up-looking, elimination-tree driven): the published     */
/* QDLDL algorithm used at qdldl_ldlt.cpp:43-75 and the per-pivot        */
/* regularized variant of regularized_ldlt.cpp:107-139,142-254.          */
/* ------------------------------------------------------------------ */
typedef struct {
  int N;
  int* perm;  /* new -> old */
  int* iperm; /* old -> new */
  int* Up;    /* permuted upper triangle pattern */
  int* Ui;
  int* umap;  /* K lower value index -> U value index */
  oc_real* Ux;
  int* etree;
  int* Lnz;
  int* Lp;
  int* Li;
  oc_real* Lx;
  oc_real* D;
  oc_real* Dinv;
  oc_real* sign; /* expected pivot sign per permuted position */
  int* yIdx;
  int* stack;
  int* next;
  char* mark;
  oc_real* yv;
  oc_real flops;
  int info_ok;
  int num_reg;
} oc_ldl;

static void ldl_free(oc_ldl* F) {
  free(F->perm), free(F->iperm), free(F->Up), free(F->Ui), free(F->umap), free(F->Ux);
  free(F->etree), free(F->Lnz), free(F->Lp), free(F->Li), free(F->Lx), free(F->D), free(F->Dinv);
  free(F->sign), free(F->yIdx), free(F->stack), free(F->next), free(F->mark), free(F->yv);
  memset(F, 0, sizeof(*F));
}

/* analyzePattern: ordering, permuted upper triangle, etree + column counts. */
static int ldl_analyze(oc_ldl* F, int N, const int* Kp, const int* Ki, int nx) {
  ldl_free(F);
  F->N = N;
  F->perm = (int*)xcalloc((size_t)N, sizeof(int));
  F->iperm = (int*)xcalloc((size_t)N, sizeof(int));
  min_degree_order(N, Kp, Ki, F->perm);
  for (int k = 0; k < N; ++k) F->iperm[F->perm[k]] = k;
  int nnz = Kp[N];
  /* permuted upper: K(i,j), i>=j  ->  (min(pi,pj), max(pi,pj)) */
  F->Up = (int*)xcalloc((size_t)N + 1, sizeof(int));
  F->Ui = (int*)xcalloc((size_t)nnz, sizeof(int));
  F->umap = (int*)xcalloc((size_t)nnz, sizeof(int));
  F->Ux = (oc_real*)xcalloc((size_t)nnz, sizeof(oc_real));
  int* cnt = (int*)xcalloc((size_t)N + 1, sizeof(int));
  for (int c = 0; c < N; ++c)
    for (int k = Kp[c]; k < Kp[c + 1]; ++k) {
      int pi = F->iperm[Ki[k]], pj = F->iperm[c];
      int col = pi > pj ? pi : pj;
      cnt[col + 1]++;
    }
  for (int c = 0; c < N; ++c) cnt[c + 1] += cnt[c];
  memcpy(F->Up, cnt, sizeof(int) * ((size_t)N + 1));
  for (int c = 0; c < N; ++c)
    for (int k = Kp[c]; k < Kp[c + 1]; ++k) {
      int pi = F->iperm[Ki[k]], pj = F->iperm[c];
      int col = pi > pj ? pi : pj, row = pi > pj ? pj : pi;
      int dst = cnt[col]++;
      F->Ui[dst] = row;
      F->umap[k] = dst;
    }
  free(cnt);
  /* elimination tree + column counts (row-subtree walk) */
  F->etree = (int*)xcalloc((size_t)N, sizeof(int));
  F->Lnz = (int*)xcalloc((size_t)N, sizeof(int));
  int* flag = (int*)xcalloc((size_t)N, sizeof(int));
  for (int i = 0; i < N; ++i) F->etree[i] = -1, flag[i] = -1;
  for (int j = 0; j < N; ++j) {
    flag[j] = j;
    for (int k = F->Up[j]; k < F->Up[j + 1]; ++k) {
      int i = F->Ui[k];
      if (i > j) return 0;
      while (flag[i] != j) {
        if (F->etree[i] == -1) F->etree[i] = j;
        F->Lnz[i]++;
        flag[i] = j;
        i = F->etree[i];
      }
    }
  }
  free(flag);
  F->Lp = (int*)xcalloc((size_t)N + 1, sizeof(int));
  oc_real flops = 0.0;
  for (int i = 0; i < N; ++i) {
    F->Lp[i + 1] = F->Lp[i] + F->Lnz[i];
    oc_real cj = (oc_real)F->Lnz[i] + 1.0;
    flops += cj * (cj - 1.0);
  }
  F->flops = flops;
  F->Li = (int*)xcalloc((size_t)F->Lp[N], sizeof(int));
  F->Lx = (oc_real*)xcalloc((size_t)F->Lp[N], sizeof(oc_real));
  F->D = (oc_real*)xcalloc((size_t)N, sizeof(oc_real));
  F->Dinv = (oc_real*)xcalloc((size_t)N, sizeof(oc_real));
  F->sign = (oc_real*)xcalloc((size_t)N, sizeof(oc_real));
  for (int k = 0; k < N; ++k) F->sign[k] = (F->perm[k] < nx) ? 1.0 : -1.0;
  F->yIdx = (int*)xcalloc((size_t)N, sizeof(int));
  F->stack = (int*)xcalloc((size_t)N, sizeof(int));
  F->next = (int*)xcalloc((size_t)N, sizeof(int));
  F->mark = (char*)xcalloc((size_t)N, 1);
  F->yv = (oc_real*)xcalloc((size_t)N, sizeof(oc_real));
  F->info_ok = 1;
  return 1;
}

/* regularizePivot: regularized_ldlt.cpp:132-139 */
static int reg_pivot(oc_real* Dk, oc_real sgn, oc_real eps, oc_real delta) {
  if (*Dk * sgn >= eps) return 0;
  *Dk += sgn * delta;
  if (*Dk * sgn < eps) *Dk = sgn * delta;
  return 1;
}

/* Numeric factorization of the K lower values (already bumped by the caller). */
static void ldl_factor(oc_ldl* F, const oc_real* Kx, int regularize, oc_real eps, oc_real delta) {
  int N = F->N;
  int nnz = F->Up[N];
  (void)nnz;
  for (int k = 0; k < F->Up[N]; ++k) F->Ux[k] = 0.0;
  /* scatter the K lower values (umap is injective) */
  /* (the caller passes Kx of length nnz(K lower) in CSC order) */
  for (int k = 0; k < F->Up[N]; ++k) F->Ux[F->umap[k]] = Kx[k];
  for (int i = 0; i < N; ++i) F->next[i] = F->Lp[i], F->mark[i] = 0, F->yv[i] = 0.0;
  int nreg = 0;
  F->info_ok = 1;
  for (int k = 0; k < N; ++k) {
    oc_real Dk = 0.0;
    int ny = 0;
    /* nonzero pattern of row k of L: reach of column k's entries in the etree;
       each path is pushed so that descendants are processed before ancestors */
    for (int t = F->Up[k]; t < F->Up[k + 1]; ++t) {
      int i = F->Ui[t];
      if (i == k) {
        Dk += F->Ux[t];
        continue;
      }
      F->yv[i] += F->Ux[t];
      int len = 0;
      while (i != -1 && i < k && !F->mark[i]) {
        F->mark[i] = 1;
        F->stack[len++] = i;
        i = F->etree[i];
      }
      while (len > 0) F->yIdx[ny++] = F->stack[--len];
    }
    for (int t = ny - 1; t >= 0; --t) {
      int c = F->yIdx[t];
      oc_real yc = F->yv[c];
      int end = F->next[c];
      for (int j = F->Lp[c]; j < end; ++j) F->yv[F->Li[j]] -= F->Lx[j] * yc;
      oc_real lkc = yc * F->Dinv[c];
      F->Li[end] = k;
      F->Lx[end] = lkc;
      Dk -= yc * lkc;
      F->next[c] = end + 1;
      F->yv[c] = 0.0;
      F->mark[c] = 0;
    }
    if (regularize) {
      if (reg_pivot(&Dk, F->sign[k], eps, delta)) ++nreg;
    } else if (Dk == 0.0) {
      F->info_ok = 0;
    }
    F->D[k] = Dk;
    F->Dinv[k] = 1.0 / Dk;
  }
  F->num_reg = nreg;
  if (regularize && (oc_real)nreg > 0.05 * (oc_real)N) F->info_ok = 0;
}

/* QDLDL_solve with the permutation (qdldl_ldlt.cpp:71-75). */
static void ldl_solve(const oc_ldl* F, const oc_real* rhs, oc_real* x, oc_real* work) {
  int N = F->N;
  for (int k = 0; k < N; ++k) work[k] = rhs[F->perm[k]];
  for (int j = 0; j < N; ++j) {
    oc_real xj = work[j];
    for (int p = F->Lp[j]; p < F->Lp[j + 1]; ++p) work[F->Li[p]] -= F->Lx[p] * xj;
  }
  for (int j = 0; j < N; ++j) work[j] *= F->Dinv[j];
  for (int j = N - 1; j >= 0; --j) {
    oc_real acc = work[j];
    for (int p = F->Lp[j]; p < F->Lp[j + 1]; ++p) acc -= F->Lx[p] * work[F->Li[p]];
    work[j] = acc;
  }
  for (int k = 0; k < N; ++k) x[F->perm[k]] = work[k];
}

/* ------------------------------------------------------------------ */
/* KKT system: src/conicxx/kkt/kkt_system.cpp:22-311                   */
/* ------------------------------------------------------------------ */
typedef struct {
  int n, m, N;
  oc_real sreg_p, sreg_a, dyn_eps, dyn_delta, refine_tol;
  int refine_max, backend;
  int* Kp; /* K lower, CSC */
  int* Ki;
  oc_real* Kx;
  oc_real* Ktry;
  int* slot_pdiag; /* n */
  int npoff;
  int* poff_slot; int* poff_r; int* poff_c;
  int na;
  int* a_slot; int* a_r; int* a_c;
  int nhs;
  int* hs_slot; int* hs_blk; int* hs_a; int* hs_b;
  int* pP; int* iP; int pnnz; /* pattern fingerprints */
  int* pA; int* iA; int annz;
  oc_ldl F;
  int factorized;
  oc_real last_p, last_a;
  oc_real* w1; oc_real* w2; oc_real* w3;
} oc_kkt;

static void kkt_free(oc_kkt* K) {
  free(K->Kp), free(K->Ki), free(K->Kx), free(K->Ktry), free(K->slot_pdiag);
  free(K->poff_slot), free(K->poff_r), free(K->poff_c);
  free(K->a_slot), free(K->a_r), free(K->a_c);
  free(K->hs_slot), free(K->hs_blk), free(K->hs_a), free(K->hs_b);
  free(K->pP), free(K->iP), free(K->pA), free(K->iA);
  free(K->w1), free(K->w2), free(K->w3);
  ldl_free(&K->F);
  memset(K, 0, sizeof(*K));
}

typedef struct { int r, c, id; } oc_key;
static int key_cmp(const void* a, const void* b) {
  const oc_key* x = (const oc_key*)a;
  const oc_key* y = (const oc_key*)b;
  if (x->c != y->c) return x->c < y->c ? -1 : 1;
  if (x->r != y->r) return x->r < y->r ? -1 : 1;
  return x->id < y->id ? -1 : (x->id > y->id);
}

static int same_pattern(const int* p, const int* i, int cols, const int* pref, const int* iref) {
  for (int c = 0; c <= cols; ++c)
    if (p[c] != pref[c]) return 0;
  for (int k = 0; k < p[cols]; ++k)
    if (i[k] != iref[k]) return 0;
  return 1;
}

static void kkt_write_bump(const oc_kkt* K, oc_real* Kv, oc_real ep, oc_real ea) {
  /* writeRegularizedDiagonal: kkt_system.cpp:152-168 */
  if (ep != 0.0)
    for (int i = 0; i < K->n; ++i) Kv[K->slot_pdiag[i]] += ep;
  if (ea != 0.0)
    for (int t = 0; t < K->nhs; ++t)
      if (K->hs_a[t] == K->hs_b[t]) Kv[K->hs_slot[t]] -= ea;
}

/* tryFactorizeWithBump: kkt_system.cpp:222-240 */
static int kkt_try_factor(oc_kkt* K, oc_real ep, oc_real ea) {
  const oc_real* Kv = K->Kx;
  if (ep != 0.0 || ea != 0.0) {
    memcpy(K->Ktry, K->Kx, sizeof(oc_real) * (size_t)K->Kp[K->N]);
    kkt_write_bump(K, K->Ktry, ep, ea);
    Kv = K->Ktry;
  }
  ldl_factor(&K->F, Kv, K->backend == LS_REG, K->dyn_eps, K->dyn_delta);
  if (!K->F.info_ok) return 0;
  oc_real mn = 1.0;
  if (K->N > 0) {
    mn = fabs(K->F.D[0]);
    for (int i = 1; i < K->N; ++i) mn = dmin(mn, fabs(K->F.D[i]));
    for (int i = 0; i < K->N; ++i)
      if (isnan(K->F.D[i])) mn = NAN;
  }
  return isfinite(mn) && mn > K->dyn_eps;
}

/* factorizeWithRetry: kkt_system.cpp:242-264 */
static int kkt_factor_retry(oc_kkt* K) {
  oc_real ep = 0.0, ea = 0.0;
  for (int att = 0; att < 18; ++att) {
    if (kkt_try_factor(K, ep, ea)) {
      K->factorized = 1;
      K->last_p = ep, K->last_a = ea;
      return 1;
    }
    ep = dmax(ep * 10.0, K->dyn_delta);
    ea = dmax(ea * 10.0, K->dyn_delta);
  }
  K->factorized = 0;
  return 0;
}

/* y = K x, K symmetric from lower storage (selfadjointView<Lower>) */
static void kkt_spmv(const oc_kkt* K, const oc_real* x, oc_real* y) {
  for (int i = 0; i < K->N; ++i) y[i] = 0.0;
  for (int c = 0; c < K->N; ++c)
    for (int k = K->Kp[c]; k < K->Kp[c + 1]; ++k) {
      int r = K->Ki[k];
      y[r] += K->Kx[k] * x[c];
      if (r != c) y[c] += K->Kx[k] * x[r];
    }
}

/* escalateAndResolve: kkt_system.cpp:266-294 */
static int kkt_escalate(oc_kkt* K, const oc_real* rhs, oc_real* x) {
  oc_real ep = K->last_p, ea = K->last_a;
  for (int att = 0; att < 10; ++att) {
    ep = dmax(ep * 1000.0, K->dyn_delta);
    ea = dmax(ea * 1000.0, K->dyn_delta);
    if (!kkt_try_factor(K, ep, ea)) continue;
    ldl_solve(&K->F, rhs, x, K->w3);
    if (vfinite(K->N, x)) {
      K->last_p = ep, K->last_a = ea;
      return 1;
    }
  }
  return 0;
}

/* KktSystem::solve: kkt_system.cpp:296-311 */
static oc_real kkt_solve(oc_kkt* K, const oc_real* rhs, oc_real* x) {
  int N = K->N;
  ldl_solve(&K->F, rhs, x, K->w3);
  if (!vfinite(N, x)) kkt_escalate(K, rhs, x);
  oc_real rn = dmax(vnorm(N, rhs), 1e-30);
  oc_real* r = K->w1;
  oc_real* dx = K->w2;
  kkt_spmv(K, x, r);
  for (int i = 0; i < N; ++i) r[i] = rhs[i] - r[i];
  oc_real rel = vnorm(N, r) / rn;
  for (int it = 0; it < K->refine_max && rel > K->refine_tol; ++it) {
    ldl_solve(&K->F, r, dx, K->w3);
    for (int i = 0; i < N; ++i) x[i] += dx[i];
    kkt_spmv(K, x, r);
    for (int i = 0; i < N; ++i) r[i] = rhs[i] - r[i];
    rel = vnorm(N, r) / rn;
  }
  return rel;
}

/* KktSystem::updateData: kkt_system.cpp:119-137 */
static int kkt_update_data(oc_kkt* K, const oc_csc* P, const oc_csc* A) {
  if (P) {
    if (P->nnz != K->pnnz || !same_pattern(P->p, P->i, K->n, K->pP, K->iP)) return 0;
    for (int i = 0; i < K->n; ++i) K->Kx[K->slot_pdiag[i]] = csc_coeff(P, i, i) + K->sreg_p;
    for (int t = 0; t < K->npoff; ++t) K->Kx[K->poff_slot[t]] = csc_coeff(P, K->poff_r[t], K->poff_c[t]);
  }
  if (A) {
    if (A->nnz != K->annz || !same_pattern(A->p, A->i, K->n, K->pA, K->iA)) return 0;
    for (int t = 0; t < K->na; ++t) K->Kx[K->a_slot[t]] = csc_coeff(A, K->a_r[t], K->a_c[t]);
  }
  return kkt_factor_retry(K);
}

/* KktSystem::updateScalingAndFactorize: kkt_system.cpp:139-150 */
static int kkt_update_scaling(oc_kkt* K, const oc_cones* C) {
  for (int t = 0; t < K->nhs; ++t) {
    const oc_block* B = &C->b[K->hs_blk[t]];
    int a = K->hs_a[t], b = K->hs_b[t];
    oc_real h;
    if (B->type == CT_ZERO)
      h = 0.0;
    else if (B->type == CT_NONNEG)
      h = (a == b) ? B->Hs[a] : 0.0;
    else
      h = B->Hs[a + b * B->dim];
    oc_real v = -h;
    if (a == b) v -= K->sreg_a;
    K->Kx[K->hs_slot[t]] = v;
  }
  return kkt_factor_retry(K);
}

/* KktSystem::setup: kkt_system.cpp:22-117 (+ SparsityMap::finalize, sparsity_map.cpp:14-56) */
static int kkt_setup(oc_kkt* K, const oc_csc* P, const oc_csc* A, const oc_cones* C,
                     const oc_settings* st) {
  kkt_free(K);
  int n = P->rows, m = C->total, N = n + m;
  K->n = n, K->m = m, K->N = N;
  K->sreg_p = st->static_reg_P, K->sreg_a = st->static_reg_A;
  K->dyn_eps = st->dynamic_reg_eps, K->dyn_delta = st->dynamic_reg_delta;
  K->refine_max = st->refine_max_iter, K->refine_tol = st->refine_tol;
  K->backend = st->linear_solver;
  int maxslots = n + P->nnz + A->nnz;
  for (int bi = 0; bi < C->nblocks; ++bi) {
    int d = C->b[bi].dim;
    maxslots += (C->b[bi].type == CT_SOC) ? d * (d + 1) / 2 : d;
  }
  oc_key* keys = (oc_key*)xcalloc((size_t)maxslots, sizeof(oc_key));
  int ns = 0;
  K->slot_pdiag = (int*)xcalloc((size_t)n, sizeof(int));
  for (int i = 0; i < n; ++i) K->slot_pdiag[i] = ns, keys[ns].r = i, keys[ns].c = i, keys[ns].id = ns, ++ns;
  K->poff_slot = (int*)xcalloc((size_t)P->nnz, sizeof(int));
  K->poff_r = (int*)xcalloc((size_t)P->nnz, sizeof(int));
  K->poff_c = (int*)xcalloc((size_t)P->nnz, sizeof(int));
  for (int c = 0; c < P->cols; ++c)
    for (int k = P->p[c]; k < P->p[c + 1]; ++k) {
      int r = P->i[k];
      if (r >= c) continue;
      K->poff_slot[K->npoff] = ns, K->poff_r[K->npoff] = r, K->poff_c[K->npoff] = c, K->npoff++;
      keys[ns].r = c, keys[ns].c = r, keys[ns].id = ns, ++ns;
    }
  K->a_slot = (int*)xcalloc((size_t)A->nnz, sizeof(int));
  K->a_r = (int*)xcalloc((size_t)A->nnz, sizeof(int));
  K->a_c = (int*)xcalloc((size_t)A->nnz, sizeof(int));
  for (int c = 0; c < A->cols; ++c)
    for (int k = A->p[c]; k < A->p[c + 1]; ++k) {
      int r = A->i[k];
      K->a_slot[K->na] = ns, K->a_r[K->na] = r, K->a_c[K->na] = c, K->na++;
      keys[ns].r = n + r, keys[ns].c = c, keys[ns].id = ns, ++ns;
    }
  int nhsmax = maxslots - ns;
  K->hs_slot = (int*)xcalloc((size_t)nhsmax, sizeof(int));
  K->hs_blk = (int*)xcalloc((size_t)nhsmax, sizeof(int));
  K->hs_a = (int*)xcalloc((size_t)nhsmax, sizeof(int));
  K->hs_b = (int*)xcalloc((size_t)nhsmax, sizeof(int));
  for (int bi = 0; bi < C->nblocks; ++bi) {
    const oc_block* B = &C->b[bi];
    int off = B->off, d = B->dim;
    if (B->type == CT_SOC) {
      for (int a = 0; a < d; ++a)
        for (int b = 0; b <= a; ++b) {
          int t = K->nhs++;
          K->hs_slot[t] = ns, K->hs_blk[t] = bi, K->hs_a[t] = a, K->hs_b[t] = b;
          keys[ns].r = n + off + a, keys[ns].c = n + off + b, keys[ns].id = ns, ++ns;
        }
    } else {
      for (int a = 0; a < d; ++a) {
        int t = K->nhs++;
        K->hs_slot[t] = ns, K->hs_blk[t] = bi, K->hs_a[t] = a, K->hs_b[t] = a;
        keys[ns].r = n + off + a, keys[ns].c = n + off + a, keys[ns].id = ns, ++ns;
      }
    }
  }
  /* finalize: unique keys sorted by (col,row) -> CSC; slot -> offset */
  qsort(keys, (size_t)ns, sizeof(oc_key), key_cmp);
  int* slot_off = (int*)xcalloc((size_t)ns, sizeof(int));
  K->Kp = (int*)xcalloc((size_t)N + 1, sizeof(int));
  K->Ki = (int*)xcalloc((size_t)ns, sizeof(int));
  int nnz = 0;
  for (int t = 0; t < ns; ++t) {
    if (t == 0 || keys[t].r != keys[t - 1].r || keys[t].c != keys[t - 1].c) {
      K->Ki[nnz] = keys[t].r;
      K->Kp[keys[t].c + 1]++;
      ++nnz;
    }
    slot_off[keys[t].id] = nnz - 1;
  }
  for (int c = 0; c < N; ++c) K->Kp[c + 1] += K->Kp[c];
  free(keys);
  K->Kx = (oc_real*)xcalloc((size_t)nnz, sizeof(oc_real));
  K->Ktry = (oc_real*)xcalloc((size_t)nnz, sizeof(oc_real));
  for (int i = 0; i < n; ++i) K->slot_pdiag[i] = slot_off[K->slot_pdiag[i]];
  for (int t = 0; t < K->npoff; ++t) K->poff_slot[t] = slot_off[K->poff_slot[t]];
  for (int t = 0; t < K->na; ++t) K->a_slot[t] = slot_off[K->a_slot[t]];
  for (int t = 0; t < K->nhs; ++t) K->hs_slot[t] = slot_off[K->hs_slot[t]];
  free(slot_off);
  K->w1 = (oc_real*)xcalloc((size_t)N, sizeof(oc_real));
  K->w2 = (oc_real*)xcalloc((size_t)N, sizeof(oc_real));
  K->w3 = (oc_real*)xcalloc((size_t)N, sizeof(oc_real));
  if (!ldl_analyze(&K->F, N, K->Kp, K->Ki, n)) return 0;
  K->pnnz = P->nnz;
  K->pP = (int*)xcalloc((size_t)n + 1, sizeof(int));
  K->iP = (int*)xcalloc((size_t)P->nnz, sizeof(int));
  memcpy(K->pP, P->p, sizeof(int) * ((size_t)n + 1));
  memcpy(K->iP, P->i, sizeof(int) * (size_t)P->nnz);
  K->annz = A->nnz;
  K->pA = (int*)xcalloc((size_t)n + 1, sizeof(int));
  K->iA = (int*)xcalloc((size_t)A->nnz, sizeof(int));
  memcpy(K->pA, A->p, sizeof(int) * ((size_t)n + 1));
  memcpy(K->iA, A->i, sizeof(int) * (size_t)A->nnz);
  if (!kkt_update_data(K, P, A)) return 0;
  return kkt_update_scaling(K, C);
}

/* ------------------------------------------------------------------ */
/* IPM driver: src/conicxx/solver_impl.cpp                              */
/* ------------------------------------------------------------------ */
struct oc_solver {
  oc_settings st;
  int n, m, setup_done;
  oc_csc P, A;
  oc_real* q;
  oc_real* b;
  oc_cones C;
  oc_equil E;
  oc_kkt K;
  int zero_dim, nonneg_dim, n_soc;
  int* soc_dims;
  oc_real *x, *s, *z, tau, kappa;
  oc_real *Px, *rx, *rz;
  oc_real rtau, dqx, dbz, dsz, dxPx;
  oc_real *x1, *z1, *Px1;
  oc_real *dxa, *dza, *dsa, dtaua, dkappaa;
  oc_real *dx, *dz, *ds, dtau, dkappa;
  int have_warm;
  oc_real *wx, *ws, *wz;
  /* solution */
  int status, iterations;
  oc_real *xo, *so, *zo, objective, pres, dres, gap, mu_out;
  oc_real* trace;
  int ntrace;
  /* scratch */
  oc_real *rhs, *sol, *t1, *t2, *t3, *t4, *t5;
};

oc_solver* oc_create(void) {
  oc_solver* h = (oc_solver*)xcalloc(1, sizeof(oc_solver));
  h->tau = 1.0, h->kappa = 1.0; /* member defaults, solver_impl.h:72 */
  oc_default_settings(&h->st);
  return h;
}

static void solver_free_data(oc_solver* h) {
  csc_free(&h->P), csc_free(&h->A);
  cones_free(&h->C);
  free(h->E.d), free(h->E.e);
  memset(&h->E, 0, sizeof(h->E));
  kkt_free(&h->K);
  oc_real** vs[] = {&h->q, &h->b, &h->x, &h->s, &h->z, &h->Px, &h->rx, &h->rz, &h->x1, &h->z1,
                   &h->Px1, &h->dxa, &h->dza, &h->dsa, &h->dx, &h->dz, &h->ds, &h->wx, &h->ws,
                   &h->wz, &h->xo, &h->so, &h->zo, &h->trace, &h->rhs, &h->sol, &h->t1, &h->t2,
                   &h->t3, &h->t4, &h->t5};
  for (size_t i = 0; i < sizeof(vs) / sizeof(vs[0]); ++i) free(*vs[i]), *vs[i] = NULL;
  free(h->soc_dims);
  h->soc_dims = NULL;
}

void oc_destroy(oc_solver* h) {
  if (!h) return;
  solver_free_data(h);
  free(h);
}

#define ALLOC_V(ptr, len) (ptr) = (oc_real*)xcalloc((size_t)(len), sizeof(oc_real))

/* SolverImpl::setup: solver_impl.cpp:8-55 */
int oc_setup(oc_solver* h, int n, int m, const int* Pp, const int* Pi, const oc_real* Px,
             const oc_real* q, const int* Ap, const int* Ai, const oc_real* Ax, const oc_real* b,
             int zero_dim, int nonneg_dim, int n_soc, const int* soc_dims, const oc_settings* st) {
  h->setup_done = 0;
  if (zero_dim < 0 || nonneg_dim < 0 || n < 0) return 0;
  int total = zero_dim + nonneg_dim;
  for (int k = 0; k < n_soc; ++k) {
    if (soc_dims[k] < 1) return 0;
    total += soc_dims[k];
  }
  if (m != total) return 0;
  solver_free_data(h);
  h->st = *st;
  h->n = n, h->m = m;
  h->zero_dim = zero_dim, h->nonneg_dim = nonneg_dim, h->n_soc = n_soc;
  h->soc_dims = (int*)xcalloc((size_t)n_soc, sizeof(int));
  if (n_soc) memcpy(h->soc_dims, soc_dims, sizeof(int) * (size_t)n_soc);
  cones_init(&h->C, zero_dim, nonneg_dim, n_soc, soc_dims);
  csc_copy(&h->P, n, n, Pp, Pi, Px);
  csc_copy(&h->A, m, n, Ap, Ai, Ax);
  ALLOC_V(h->q, n), ALLOC_V(h->b, m);
  memcpy(h->q, q, sizeof(oc_real) * (size_t)n);
  memcpy(h->b, b, sizeof(oc_real) * (size_t)m);
  equil_compute(&h->E, &h->P, &h->A, &h->C, &h->st);
  if (h->st.equilibrate) {
    equil_scale_p(&h->E, &h->P);
    equil_scale_a(&h->E, &h->A);
    equil_scale_q(&h->E, h->q);
    equil_scale_b(&h->E, h->b);
  }
  if (!kkt_setup(&h->K, &h->P, &h->A, &h->C, &h->st)) return 0;
  int N = n + m;
  ALLOC_V(h->x, n), ALLOC_V(h->s, m), ALLOC_V(h->z, m);
  ALLOC_V(h->Px, n), ALLOC_V(h->rx, n), ALLOC_V(h->rz, m);
  ALLOC_V(h->x1, n), ALLOC_V(h->z1, m), ALLOC_V(h->Px1, n);
  ALLOC_V(h->dxa, n), ALLOC_V(h->dza, m), ALLOC_V(h->dsa, m);
  ALLOC_V(h->dx, n), ALLOC_V(h->dz, m), ALLOC_V(h->ds, m);
  ALLOC_V(h->wx, n), ALLOC_V(h->ws, m), ALLOC_V(h->wz, m);
  ALLOC_V(h->xo, n), ALLOC_V(h->so, m), ALLOC_V(h->zo, m);
  ALLOC_V(h->rhs, N), ALLOC_V(h->sol, N);
  ALLOC_V(h->t1, N), ALLOC_V(h->t2, N), ALLOC_V(h->t3, N), ALLOC_V(h->t4, N), ALLOC_V(h->t5, N);
  h->have_warm = 0;
  h->status = ST_UNSOLVED;
  h->iterations = 0;
  h->setup_done = 1;
  return 1;
}

/* SolverImpl::updateData: solver_impl.cpp:57-91 */
int oc_update_data(oc_solver* h, const int* Pp, const int* Pi, const oc_real* Px, const oc_real* q,
                   const int* Ap, const int* Ai, const oc_real* Ax, const oc_real* b) {
  if (!h->setup_done) return 0;
  oc_csc Ps, As;
  memset(&Ps, 0, sizeof(Ps));
  memset(&As, 0, sizeof(As));
  oc_real* qs = NULL;
  oc_real* bs = NULL;
  int n = h->n, m = h->m, ok = 1;
  if (Px) {
    csc_copy(&Ps, n, n, Pp, Pi, Px);
    if (h->st.equilibrate) equil_scale_p(&h->E, &Ps);
  }
  if (Ax) {
    csc_copy(&As, m, n, Ap, Ai, Ax);
    if (h->st.equilibrate) equil_scale_a(&h->E, &As);
  }
  if (q) {
    ALLOC_V(qs, n);
    memcpy(qs, q, sizeof(oc_real) * (size_t)n);
    if (h->st.equilibrate) equil_scale_q(&h->E, qs);
  }
  if (b) {
    ALLOC_V(bs, m);
    memcpy(bs, b, sizeof(oc_real) * (size_t)m);
    if (h->st.equilibrate) equil_scale_b(&h->E, bs);
  }
  if (!kkt_update_data(&h->K, Px ? &Ps : NULL, Ax ? &As : NULL)) ok = 0;
  if (ok) {
    if (Px) csc_free(&h->P), h->P = Ps, memset(&Ps, 0, sizeof(Ps));
    if (Ax) csc_free(&h->A), h->A = As, memset(&As, 0, sizeof(As));
    if (q) memcpy(h->q, qs, sizeof(oc_real) * (size_t)n);
    if (b) memcpy(h->b, bs, sizeof(oc_real) * (size_t)m);
  }
  csc_free(&Ps), csc_free(&As);
  free(qs), free(bs);
  return ok;
}

/* SolverImpl::setWarmStart: solver_impl.cpp:93-100 */
void oc_set_warm_start(oc_solver* h, const oc_real* x, const oc_real* s, const oc_real* z) {
  if (!h->setup_done) return;
  memcpy(h->wx, x, sizeof(oc_real) * (size_t)h->n);
  memcpy(h->ws, s, sizeof(oc_real) * (size_t)h->m);
  memcpy(h->wz, z, sizeof(oc_real) * (size_t)h->m);
  if (h->st.equilibrate) equil_scale_solution(&h->E, h->wx, h->ws, h->wz);
  h->have_warm = 1;
}

void oc_set_settings(oc_solver* h, const oc_settings* st) { h->st = *st; }

/* shiftToInteriorCold / ensureStrictlyInteriorWarm: solver_impl.cpp:104-124 */
static void shift_cold(oc_solver* h, oc_real* v) {
  oc_real mn, ps;
  cones_margins(&h->C, v, &mn, &ps);
  oc_real deg = (oc_real)(h->C.degree > 1 ? h->C.degree : 1);
  oc_real target = dmax(1.0, 0.1 * ps / deg);
  if (mn <= 0.0) {
    cones_unit_shift(&h->C, v, -mn);
    cones_unit_shift(&h->C, v, target);
  } else if (mn < target) {
    cones_unit_shift(&h->C, v, target - mn);
  } else {
    cones_unit_shift(&h->C, v, 0.0);
  }
}
static void ensure_interior_warm(oc_solver* h, oc_real* v) {
  oc_real mn, ps;
  cones_margins(&h->C, v, &mn, &ps);
  if (mn <= 1e-8) cones_unit_shift(&h->C, v, -mn + 1e-6);
}

/* computeInitialPoint: solver_impl.cpp:126-158 */
static int initial_point(oc_solver* h) {
  int n = h->n, m = h->m;
  cones_update_scaling(&h->C, h->C.identity, h->C.identity);
  if (!kkt_update_scaling(&h->K, &h->C)) return 0;
  if (h->P.nnz == 0) {
    for (int i = 0; i < n; ++i) h->rhs[i] = 0.0;
    for (int i = 0; i < m; ++i) h->rhs[n + i] = h->b[i];
    kkt_solve(&h->K, h->rhs, h->sol);
    for (int i = 0; i < n; ++i) h->x[i] = h->sol[i];
    for (int i = 0; i < m; ++i) h->s[i] = -h->sol[n + i];
    for (int i = 0; i < n; ++i) h->rhs[i] = -h->q[i];
    for (int i = 0; i < m; ++i) h->rhs[n + i] = 0.0;
    kkt_solve(&h->K, h->rhs, h->sol);
    for (int i = 0; i < m; ++i) h->z[i] = h->sol[n + i];
  } else {
    for (int i = 0; i < n; ++i) h->rhs[i] = -h->q[i];
    for (int i = 0; i < m; ++i) h->rhs[n + i] = h->b[i];
    kkt_solve(&h->K, h->rhs, h->sol);
    for (int i = 0; i < n; ++i) h->x[i] = h->sol[i];
    for (int i = 0; i < m; ++i) h->z[i] = h->sol[n + i], h->s[i] = -h->sol[n + i];
  }
  cones_zero_zero_blocks(&h->C, h->s);
  shift_cold(h, h->s);
  shift_cold(h, h->z);
  h->tau = 1.0, h->kappa = 1.0;
  return vfinite(n, h->x) && vfinite(m, h->s) && vfinite(m, h->z);
}

/* computeConstantSolve: solver_impl.cpp:165-175 */
static int constant_solve(oc_solver* h) {
  int n = h->n, m = h->m;
  for (int i = 0; i < n; ++i) h->rhs[i] = -h->q[i];
  for (int i = 0; i < m; ++i) h->rhs[n + i] = h->b[i];
  kkt_solve(&h->K, h->rhs, h->sol);
  memcpy(h->x1, h->sol, sizeof(oc_real) * (size_t)n);
  memcpy(h->z1, h->sol + n, sizeof(oc_real) * (size_t)m);
  pmul(&h->P, h->x1, h->Px1);
  return vfinite(n, h->x1) && vfinite(m, h->z1);
}

/* computeResiduals: solver_impl.cpp:177-186 */
static void residuals(oc_solver* h) {
  int n = h->n, m = h->m;
  pmul(&h->P, h->x, h->Px);
  oc_real* Atz = h->t1;
  for (int i = 0; i < n; ++i) Atz[i] = 0.0;
  for (int i = 0; i < m; ++i) h->rz[i] = 0.0;
  for (int c = 0; c < n; ++c)
    for (int k = h->A.p[c]; k < h->A.p[c + 1]; ++k) {
      Atz[c] += h->A.x[k] * h->z[h->A.i[k]];
      h->rz[h->A.i[k]] += h->A.x[k] * h->x[c];
    }
  for (int i = 0; i < n; ++i) h->rx[i] = -Atz[i] - h->Px[i] - h->tau * h->q[i];
  for (int i = 0; i < m; ++i) h->rz[i] = h->rz[i] + h->s[i] - h->tau * h->b[i];
  h->dqx = vdot(n, h->q, h->x);
  h->dbz = vdot(m, h->b, h->z);
  h->dsz = vdot(m, h->s, h->z);
  h->dxPx = vdot(n, h->x, h->Px);
  h->rtau = h->dqx + h->dbz + h->kappa + h->dxPx / h->tau;
}

static oc_real compute_mu(const oc_solver* h) {
  return (h->dsz + h->tau * h->kappa) / (oc_real)(h->C.degree + 1);
}

/* shared tau/kappa elimination of computeAffineStep / computeCombinedStep
 * (solver_impl.cpp:200-231, 255-286) */
static int tau_elim(oc_solver* h, const oc_real* xv, const oc_real* zv, oc_real dtau_rhs,
                    oc_real dkappa_rhs, oc_real* dtau_out) {
  int n = h->n, m = h->m;
  oc_real* xi = h->t2;
  oc_real* Pv = h->t3;
  for (int i = 0; i < n; ++i) xi[i] = h->x[i] / h->tau;
  pmul(&h->P, xv, Pv);
  oc_real num = dtau_rhs - dkappa_rhs / h->tau + vdot(n, h->q, xv) + vdot(m, h->b, zv) + 2.0 * vdot(n, xi, Pv);
  oc_real* xm = h->t4;
  for (int i = 0; i < n; ++i) xm[i] = xi[i] - h->x1[i];
  pmul(&h->P, xm, Pv);
  oc_real den = h->kappa / h->tau - vdot(n, h->q, h->x1) - vdot(m, h->b, h->z1) + vdot(n, xm, Pv) -
               vdot(n, h->x1, h->Px1);
  if (den == 0.0 || !isfinite(num) || !isfinite(den)) return 0;
  *dtau_out = num / den;
  return 1;
}

/* computeAffineStep: solver_impl.cpp:192-231 */
static int affine_step(oc_solver* h) {
  int n = h->n, m = h->m;
  for (int i = 0; i < n; ++i) h->rhs[i] = h->rx[i];
  for (int i = 0; i < m; ++i) h->rhs[n + i] = h->s[i] - h->rz[i];
  kkt_solve(&h->K, h->rhs, h->sol);
  const oc_real* xv = h->sol;
  const oc_real* zv = h->sol + n;
  if (!vfinite(n + m, h->sol)) return 0;
  oc_real dtau_rhs = h->rtau, dkappa_rhs = h->tau * h->kappa;
  if (!tau_elim(h, xv, zv, dtau_rhs, dkappa_rhs, &h->dtaua)) return 0;
  for (int i = 0; i < n; ++i) h->dxa[i] = xv[i] + h->dtaua * h->x1[i];
  for (int i = 0; i < m; ++i) h->dza[i] = zv[i] + h->dtaua * h->z1[i];
  oc_real* Hz = h->t5;
  cones_mul_hs(&h->C, h->dza, Hz);
  for (int i = 0; i < m; ++i) h->dsa[i] = -(Hz[i] + h->s[i]);
  h->dkappaa = -(dkappa_rhs + h->kappa * h->dtaua) / h->tau;
  return isfinite(h->dtaua) && isfinite(h->dkappaa) && vfinite(n, h->dxa) && vfinite(m, h->dza) &&
         vfinite(m, h->dsa);
}

/* computeCombinedStep: solver_impl.cpp:233-286 */
static int combined_step(oc_solver* h, oc_real sigma, oc_real mu) {
  int n = h->n, m = h->m;
  oc_real* lam = (oc_real*)xcalloc((size_t)m, sizeof(oc_real));
  oc_real* lamp = (oc_real*)xcalloc((size_t)m, sizeof(oc_real));
  oc_real* wds = (oc_real*)xcalloc((size_t)m, sizeof(oc_real));
  oc_real* wdz = (oc_real*)xcalloc((size_t)m, sizeof(oc_real));
  oc_real* corr = (oc_real*)xcalloc((size_t)m, sizeof(oc_real));
  oc_real* dsc = (oc_real*)xcalloc((size_t)m, sizeof(oc_real));
  oc_real* lid = (oc_real*)xcalloc((size_t)m, sizeof(oc_real));
  oc_real* dsk = (oc_real*)xcalloc((size_t)m, sizeof(oc_real));
  int ok = 0;
  cones_apply_w(&h->C, h->z, lam);
  cones_product(&h->C, lam, lam, lamp);
  cones_apply_winv(&h->C, h->dsa, wds);
  cones_apply_w(&h->C, h->dza, wdz);
  cones_product(&h->C, wds, wdz, corr);
  for (int i = 0; i < m; ++i) dsc[i] = lamp[i] + corr[i] - sigma * mu * h->C.identity[i];
  cones_inverse_product(&h->C, lam, dsc, lid);
  cones_apply_w(&h->C, lid, dsk);
  for (int i = 0; i < n; ++i) h->rhs[i] = (1.0 - sigma) * h->rx[i];
  for (int i = 0; i < m; ++i) h->rhs[n + i] = dsk[i] - (1.0 - sigma) * h->rz[i];
  kkt_solve(&h->K, h->rhs, h->sol);
  const oc_real* xv = h->sol;
  const oc_real* zv = h->sol + n;
  if (vfinite(n + m, h->sol)) {
    oc_real dtau_rhs = (1.0 - sigma) * h->rtau;
    oc_real dkappa_rhs = h->tau * h->kappa + h->dtaua * h->dkappaa - sigma * mu;
    if (tau_elim(h, xv, zv, dtau_rhs, dkappa_rhs, &h->dtau)) {
      for (int i = 0; i < n; ++i) h->dx[i] = xv[i] + h->dtau * h->x1[i];
      for (int i = 0; i < m; ++i) h->dz[i] = zv[i] + h->dtau * h->z1[i];
      oc_real* Hz = h->t5;
      cones_mul_hs(&h->C, h->dz, Hz);
      for (int i = 0; i < m; ++i) h->ds[i] = -(Hz[i] + dsk[i]);
      h->dkappa = -(dkappa_rhs + h->kappa * h->dtau) / h->tau;
      ok = isfinite(h->dtau) && isfinite(h->dkappa) && vfinite(n, h->dx) && vfinite(m, h->dz) &&
           vfinite(m, h->ds);
    }
  }
  free(lam), free(lamp), free(wds), free(wdz), free(corr), free(dsc), free(lid), free(dsk);
  return ok;
}

/* computeStepLength: solver_impl.cpp:288-296 */
static oc_real step_length(const oc_solver* h, const oc_real* ds, const oc_real* dz, oc_real dtau,
                          oc_real dkappa) {
  oc_real alpha = 1.0;
  if (dtau < 0) alpha = dmin(alpha, -h->tau / dtau);
  if (dkappa < 0) alpha = dmin(alpha, -h->kappa / dkappa);
  alpha = cones_max_step(&h->C, h->z, dz, alpha);
  alpha = cones_max_step(&h->C, h->s, ds, alpha);
  return alpha;
}

/* finalizeSolution: solver_impl.cpp:342-365 */
static void finalize(oc_solver* h) {
  int n = h->n, m = h->m;
  int inf = (h->status == ST_PRIMAL_INF || h->status == ST_DUAL_INF);
  oc_real scale = inf ? 1.0 / dmax(h->kappa, 1e-30) : 1.0 / dmax(h->tau, 1e-30);
  for (int i = 0; i < n; ++i) h->xo[i] = h->x[i] * scale;
  for (int i = 0; i < m; ++i) h->so[i] = h->s[i] * scale, h->zo[i] = h->z[i] * scale;
  if (h->st.equilibrate) equil_unscale_solution(&h->E, h->xo, h->so, h->zo);
  oc_real pobj = (0.5 * h->dxPx / h->tau + h->dqx) / h->tau;
  if (h->st.equilibrate) pobj /= h->E.c;
  h->objective = pobj;
  oc_real gap = h->dsz / (h->tau * h->tau);
  if (h->st.equilibrate) gap /= h->E.c;
  h->gap = gap;
  h->pres = vnorm(m, h->rz) / dmax(h->tau, 1e-30);
  h->dres = vnorm(n, h->rx) / dmax(h->tau, 1e-30);
  h->mu_out = compute_mu(h);
}

/* SolverImpl::solve: solver_impl.cpp:367-466 */
int oc_solve(oc_solver* h) {
  h->status = ST_UNSOLVED;
  h->iterations = 0;
  h->objective = h->pres = h->dres = h->gap = h->mu_out = 0.0;
  h->ntrace = 0;
  if (!h->setup_done) {
    h->status = ST_NUMERR;
    return h->status;
  }
  int n = h->n, m = h->m;
  free(h->trace);
  int tcap = h->st.max_iter > 0 ? h->st.max_iter : 1;
  h->trace = (oc_real*)xcalloc((size_t)tcap * 8, sizeof(oc_real));
  int use_warm = h->st.warm_start && h->have_warm;
  if (use_warm) {
    memcpy(h->x, h->wx, sizeof(oc_real) * (size_t)n);
    memcpy(h->s, h->ws, sizeof(oc_real) * (size_t)m);
    memcpy(h->z, h->wz, sizeof(oc_real) * (size_t)m);
    h->tau = 1.0, h->kappa = 1.0;
    cones_zero_zero_blocks(&h->C, h->s);
    ensure_interior_warm(h, h->s);
    ensure_interior_warm(h, h->z);
  } else if (!initial_point(h)) {
    h->status = ST_NUMERR;
    finalize(h);
    return h->status;
  }
  int final_status = ST_MAXIT, iterations = 0;
  oc_real* saff = (oc_real*)xcalloc((size_t)m, sizeof(oc_real));
  oc_real* zaff = (oc_real*)xcalloc((size_t)m, sizeof(oc_real));
  for (int iter = 0; iter < h->st.max_iter; ++iter) {
    iterations = iter + 1;
    oc_real* tr = h->trace + 8 * iter;
    cones_update_scaling(&h->C, h->s, h->z);
    if (!kkt_update_scaling(&h->K, &h->C) || !constant_solve(h)) {
      final_status = ST_NUMERR;
      break;
    }
    residuals(h);
    oc_real mu = compute_mu(h);
    oc_real nrx = vnorm(n, h->rx) / h->tau;
    oc_real nrz = vnorm(m, h->rz) / h->tau;
    oc_real pobj = (0.5 * h->dxPx / h->tau + h->dqx) / h->tau;
    oc_real dobj = (-0.5 * h->dxPx / h->tau - h->dbz) / h->tau;
    tr[0] = mu, tr[1] = nrx, tr[2] = nrz, tr[3] = pobj, tr[4] = dobj;
    h->ntrace = iter + 1;
    oc_real gap_abs = fabs(pobj - dobj);
    oc_real gap_tol = h->st.tol_gap_abs + h->st.tol_gap_rel * dmax(fabs(pobj), fabs(dobj));
    if (nrx < h->st.tol_feas && nrz < h->st.tol_feas && gap_abs < gap_tol && h->tau > 1e-6) {
      final_status = ST_SOLVED;
      break;
    }
    {
      oc_real ratio = h->tau / dmax(h->kappa, 1e-30);
      if (ratio <= h->st.tol_infeas) {
        if (h->dbz < -h->st.tol_infeas * dmax(1.0, vnorm(m, h->z))) {
          final_status = ST_PRIMAL_INF;
          break;
        }
        if (h->dqx < -h->st.tol_infeas * dmax(1.0, vnorm(n, h->x))) {
          final_status = ST_DUAL_INF;
          break;
        }
      }
    }
    if (!affine_step(h)) {
      final_status = ST_NUMERR;
      break;
    }
    oc_real aa = step_length(h, h->dsa, h->dza, h->dtaua, h->dkappaa);
    for (int i = 0; i < m; ++i) saff[i] = h->s[i] + aa * h->dsa[i], zaff[i] = h->z[i] + aa * h->dza[i];
    oc_real tau_a = h->tau + aa * h->dtaua;
    oc_real kap_a = h->kappa + aa * h->dkappaa;
    oc_real mu_aff = (vdot(m, saff, zaff) + tau_a * kap_a) / (oc_real)(h->C.degree + 1);
    oc_real sigma;
    if (mu > 0) {
      oc_real r = dmax(mu_aff / mu, 0.0);
      sigma = dmin(dmax(r * r * r, 0.0), 1.0);
    } else {
      sigma = 1.0;
    }
    tr[5] = aa, tr[6] = sigma;
    if (!combined_step(h, sigma, mu)) {
      final_status = ST_NUMERR;
      break;
    }
    oc_real alpha = step_length(h, h->ds, h->dz, h->dtau, h->dkappa) * h->st.max_step_fraction;
    tr[7] = alpha;
    if (!isfinite(alpha) || alpha <= 0.0) {
      final_status = ST_NUMERR;
      break;
    }
    for (int i = 0; i < n; ++i) h->x[i] += alpha * h->dx[i];
    for (int i = 0; i < m; ++i) h->s[i] += alpha * h->ds[i], h->z[i] += alpha * h->dz[i];
    h->tau += alpha * h->dtau;
    h->kappa += alpha * h->dkappa;
    oc_real sc = dmax(h->tau, h->kappa);
    if (sc > 1e8 || sc < 1e-8) {
      oc_real inv = 1.0 / sc;
      for (int i = 0; i < n; ++i) h->x[i] *= inv;
      for (int i = 0; i < m; ++i) h->s[i] *= inv, h->z[i] *= inv;
      h->tau *= inv, h->kappa *= inv;
    }
  }
  free(saff), free(zaff);
  h->status = final_status;
  h->iterations = iterations;
  finalize(h);
  if (final_status == ST_SOLVED && h->st.warm_start) {
    memcpy(h->wx, h->x, sizeof(oc_real) * (size_t)n);
    memcpy(h->ws, h->s, sizeof(oc_real) * (size_t)m);
    memcpy(h->wz, h->z, sizeof(oc_real) * (size_t)m);
    h->have_warm = 1;
  }
  return h->status;
}

int oc_get_solution(const oc_solver* h, oc_real* x, oc_real* s, oc_real* z, oc_real* info,
                    int* iterations) {
  if (h->xo && x) memcpy(x, h->xo, sizeof(oc_real) * (size_t)h->n);
  if (h->so && s) memcpy(s, h->so, sizeof(oc_real) * (size_t)h->m);
  if (h->zo && z) memcpy(z, h->zo, sizeof(oc_real) * (size_t)h->m);
  if (info) {
    info[0] = h->pres, info[1] = h->dres, info[2] = h->gap, info[3] = h->mu_out;
    info[4] = h->objective;
  }
  if (iterations) *iterations = h->iterations;
  return h->status;
}

int oc_get_trace(const oc_solver* h, oc_real* out, int max_rows) {
  int r = h->ntrace < max_rows ? h->ntrace : max_rows;
  if (r > 0) memcpy(out, h->trace, sizeof(oc_real) * 8 * (size_t)r);
  return r;
}

void oc_get_kkt_stats(const oc_solver* h, oc_real* out4) {
  out4[0] = h->K.N;
  out4[1] = h->K.Kp ? h->K.Kp[h->K.N] : 0;
  out4[2] = h->K.F.Lp ? (oc_real)h->K.F.Lp[h->K.N] + h->K.N : 0;
  out4[3] = h->K.F.flops;
}

/* ---- per-kernel entry points ---- */
void oc_cone_margins(int zero_dim, int nonneg_dim, int n_soc, const int* soc_dims,
                     const oc_real* x, oc_real* min_margin, oc_real* pos_sum) {
  oc_cones C;
  cones_init(&C, zero_dim, nonneg_dim, n_soc, soc_dims);
  cones_margins(&C, x, min_margin, pos_sum);
  cones_free(&C);
}

oc_real oc_cone_max_step(int zero_dim, int nonneg_dim, int n_soc, const int* soc_dims,
                        const oc_real* x, const oc_real* dx, oc_real alpha_max) {
  oc_cones C;
  cones_init(&C, zero_dim, nonneg_dim, n_soc, soc_dims);
  oc_real a = cones_max_step(&C, x, dx, alpha_max);
  cones_free(&C);
  return a;
}

void oc_cone_scaling_apply(int zero_dim, int nonneg_dim, int n_soc, const int* soc_dims,
                           const oc_real* s, const oc_real* z, const oc_real* v, oc_real* lambda,
                           oc_real* Hv, oc_real* Wv, oc_real* Winv_v) {
  oc_cones C;
  cones_init(&C, zero_dim, nonneg_dim, n_soc, soc_dims);
  cones_update_scaling(&C, s, z);
  cones_apply_w(&C, z, lambda);
  cones_mul_hs(&C, v, Hv);
  cones_apply_w(&C, v, Wv);
  cones_apply_winv(&C, v, Winv_v);
  cones_free(&C);
}

void oc_equilibrate(int n, int m, const int* Pp, const int* Pi, const oc_real* Px, const int* Ap,
                    const int* Ai, const oc_real* Ax, int zero_dim, int nonneg_dim, int n_soc,
                    const int* soc_dims, const oc_settings* st, oc_real* d, oc_real* e, oc_real* c) {
  oc_cones C;
  oc_csc P, A;
  oc_equil E;
  memset(&E, 0, sizeof(E));
  cones_init(&C, zero_dim, nonneg_dim, n_soc, soc_dims);
  csc_copy(&P, n, n, Pp, Pi, Px);
  csc_copy(&A, m, n, Ap, Ai, Ax);
  equil_compute(&E, &P, &A, &C, st);
  memcpy(d, E.d, sizeof(oc_real) * (size_t)n);
  memcpy(e, E.e, sizeof(oc_real) * (size_t)m);
  *c = E.c;
  free(E.d), free(E.e);
  csc_free(&P), csc_free(&A);
  cones_free(&C);
}

oc_real oc_kkt_solve(int n, int m, const int* Pp, const int* Pi, const oc_real* Px, const int* Ap,
                    const int* Ai, const oc_real* Ax, int zero_dim, int nonneg_dim, int n_soc,
                    const int* soc_dims, const oc_real* s, const oc_real* z, const oc_settings* st,
                    const oc_real* rhs, oc_real* x_out, int* ok) {
  oc_cones C;
  oc_csc P, A;
  oc_kkt K;
  memset(&K, 0, sizeof(K));
  cones_init(&C, zero_dim, nonneg_dim, n_soc, soc_dims);
  csc_copy(&P, n, n, Pp, Pi, Px);
  csc_copy(&A, m, n, Ap, Ai, Ax);
  oc_real rel = NAN;
  *ok = kkt_setup(&K, &P, &A, &C, st);
  if (*ok && s && z) {
    cones_update_scaling(&C, s, z);
    *ok = kkt_update_scaling(&K, &C);
  }
  if (*ok) rel = kkt_solve(&K, rhs, x_out);
  kkt_free(&K);
  csc_free(&P), csc_free(&A);
  cones_free(&C);
  return rel;
}
