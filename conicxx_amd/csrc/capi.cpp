// C-ABI of the solver (declared in include/conicxx_c.h).  No exception
// crosses this boundary: HIP or allocation failures become return codes or
// a NumericalError status, as the reference's Solver never throws
// (include/conicxx/solver.h:43-47, 86-89).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <exception>
#include <new>
#include <vector>

#include "../../include/conicxx_c.h"
#include "batch.hpp"
#include "equilibrate.hpp"
#include "solver.hpp"
#include "symbolic.hpp"

extern "C" {

struct cx_solver {
  cx::Solver* impl;
};

void cx_default_settings(cx_settings* s) {
  if (!s) return;
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
  s->linear_solver = CX_LS_QDLDL;
  s->factor_precision = CX_FACTOR_FP64;
  s->fp32_reg = 1e-4;
}

cx_solver* cx_create(void) {
  try {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return nullptr;
    cx_solver* h = new cx_solver;
    h->impl = new cx::Solver();
    return h;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "[conicxx_amd] cx_create failed: %s\n", e.what());
    return nullptr;
  }
}

void cx_destroy(cx_solver* h) {
  if (!h) return;
  try {
    delete h->impl;
  } catch (...) {
  }
  delete h;
}

int cx_setup(cx_solver* h, int n, int m, const int* P_colptr, const int* P_rowind,
             const double* P_val, const double* q, const int* A_colptr, const int* A_rowind,
             const double* A_val, const double* b, int zero_dim, int nonneg_dim, int n_soc,
             const int* soc_dims, const cx_settings* settings) {
  if (!h) return 0;
  try {
    cx_settings st;
    if (settings)
      st = *settings;
    else
      cx_default_settings(&st);
    return h->impl->setup(n, m, P_colptr, P_rowind, P_val, q, A_colptr, A_rowind, A_val, b, zero_dim,
                          nonneg_dim, n_soc, soc_dims, st)
               ? 1
               : 0;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "[conicxx_amd] setup failed: %s\n", e.what());
    return 0;
  }
}

int cx_update_data(cx_solver* h, const int* P_colptr, const int* P_rowind, const double* P_val,
                   const double* q, const int* A_colptr, const int* A_rowind, const double* A_val,
                   const double* b) {
  if (!h) return 0;
  try {
    return h->impl->update_data(P_colptr, P_rowind, P_val, q, A_colptr, A_rowind, A_val, b) ? 1 : 0;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "[conicxx_amd] updateData failed: %s\n", e.what());
    return 0;
  }
}

int cx_update_data_noalloc(cx_solver* h, const int* P_colptr, const int* P_rowind, const double* P_val,
                           const double* q, const int* A_colptr, const int* A_rowind, const double* A_val,
                           const double* b) {
  if (!h) return 0;
  try {
    return h->impl->update_data_noalloc(P_colptr, P_rowind, P_val, q, A_colptr, A_rowind, A_val, b) ? 1 : 0;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "[conicxx_amd] updateData (no allocation) failed: %s\n", e.what());
    return 0;
  }
}

void cx_set_warm_start(cx_solver* h, const double* x, const double* s, const double* z) {
  if (!h) return;
  try {
    h->impl->set_warm_start(x, s, z);
  } catch (...) {
  }
}

void cx_set_settings(cx_solver* h, const cx_settings* settings) {
  if (h && settings) h->impl->set_settings(*settings);
}

void cx_get_settings(const cx_solver* h, cx_settings* out) {
  if (h && out) *out = h->impl->settings();
}

int cx_solve(cx_solver* h) {
  if (!h) return CX_NUMERICAL_ERROR;
  try {
    return h->impl->solve();
  } catch (...) {
    return CX_NUMERICAL_ERROR;
  }
}

int cx_get_solution(const cx_solver* h, double* x, double* s, double* z, double* objective,
                    cx_info* info) {
  if (!h) return CX_UNSOLVED;
  return h->impl->get_solution(x, s, z, objective, info);
}

const char* cx_status_string(int status) {
  switch (status) {
    case CX_UNSOLVED: return "Unsolved";
    case CX_SOLVED: return "Solved";
    case CX_PRIMAL_INFEASIBLE: return "PrimalInfeasible";
    case CX_DUAL_INFEASIBLE: return "DualInfeasible";
    case CX_MAX_ITERATIONS: return "MaxIterations";
    case CX_NUMERICAL_ERROR: return "NumericalError";
  }
  return "Unknown";
}

int cx_get_dims(const cx_solver* h, int* n, int* m) {
  if (!h || !n || !m) return 0;
  h->impl->dims(n, m);
  return 1;
}

void cx_get_kkt_stats(const cx_solver* h, double* stats8) {
  if (h && stats8) h->impl->kkt_stats(stats8);
}

void cx_get_profile(const cx_solver* h, double* out13) {
  if (h && out13) h->impl->profile(out13);
}

int cx_get_kernel_stats(const cx_solver* h, double* out, int max_kinds) {
  if (!h || !out) return 0;
  return h->impl->kernel_stats(out, max_kinds);
}

int cx_backward(cx_solver* h, const double* dl_dx, const double* dl_ds, const double* dl_dz,
                double* dP, double* dq, double* dA, double* db) {
  if (!h) return 0;
  try {
    return h->impl->backward(dl_dx, dl_ds, dl_dz, dP, dq, dA, db);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "[conicxx_amd] backward failed: %s\n", e.what());
    return 0;
  }
}

double cx_kkt_factor_solve(cx_solver* h, const double* s, const double* z, const double* rhs, double* x,
                           int refine_max_iter, int* factorizations) {
  if (factorizations) *factorizations = 0;
  if (!h || !rhs || !x) return NAN;
  try {
    return h->impl->kkt_factor_solve(s, z, rhs, x, refine_max_iter, factorizations);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "[conicxx_amd] kkt_factor_solve failed: %s\n", e.what());
    return NAN;
  }
}

struct cx_batch {
  cx::BatchSolver* impl;
};

cx_batch* cx_batch_create(void) {
  try {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return nullptr;
    cx_batch* h = new cx_batch;
    h->impl = new cx::BatchSolver();
    return h;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "[conicxx_amd] cx_batch_create failed: %s\n", e.what());
    return nullptr;
  }
}

void cx_batch_destroy(cx_batch* h) {
  if (!h) return;
  try {
    delete h->impl;
  } catch (...) {
  }
  delete h;
}

int cx_batch_setup(cx_batch* h, int nprob, const int* n, const int* m, const int* const* P_colptr,
                   const int* const* P_rowind, const double* const* P_val, const double* const* q,
                   const int* const* A_colptr, const int* const* A_rowind, const double* const* A_val,
                   const double* const* b, const int* zero_dim, const int* nonneg_dim, const int* n_soc,
                   const int* const* soc_dims, const cx_settings* settings) {
  if (!h) return 0;
  try {
    cx_settings st;
    if (settings)
      st = *settings;
    else
      cx_default_settings(&st);
    return h->impl->setup(nprob, n, m, P_colptr, P_rowind, P_val, q, A_colptr, A_rowind, A_val, b, zero_dim,
                          nonneg_dim, n_soc, soc_dims, st)
               ? 1
               : 0;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "[conicxx_amd] batch setup failed: %s\n", e.what());
    return 0;
  }
}

int cx_batch_update_data(cx_batch* h, const int* const* P_colptr, const int* const* P_rowind,
                         const double* const* P_val, const double* const* q, const int* const* A_colptr,
                         const int* const* A_rowind, const double* const* A_val, const double* const* b) {
  if (!h) return 0;
  try {
    return h->impl->update_data(P_colptr, P_rowind, P_val, q, A_colptr, A_rowind, A_val, b) ? 1 : 0;
  } catch (...) {
    return 0;
  }
}

int cx_batch_solve(cx_batch* h, int* status) {
  if (!h) return -1;
  try {
    return h->impl->solve(status);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "[conicxx_amd] batch solve failed: %s\n", e.what());
    return -1;
  }
}

int cx_batch_get_solution(cx_batch* h, int k, double* x, double* s, double* z, double* objective,
                          cx_info* info) {
  if (!h) return CX_UNSOLVED;
  try {
    return h->impl->get_solution(k, x, s, z, objective, info);
  } catch (...) {
    return CX_NUMERICAL_ERROR;
  }
}

void cx_batch_set_settings(cx_batch* h, const cx_settings* settings) {
  if (h && settings) h->impl->set_settings(*settings);
}

int cx_batch_autotune(cx_batch* h, int reps) {
  if (!h) return 0;
  try {
    return h->impl->autotune(reps);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "[conicxx_amd] batch autotune failed: %s\n", e.what());
    return 0;
  }
}

int cx_batch_set_tuning(cx_batch* h, int lds_doubles, int grid) {
  if (!h) return 0;
  try {
    return h->impl->set_tuning(lds_doubles, grid) ? 1 : 0;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "[conicxx_amd] batch set_tuning failed: %s\n", e.what());
    return 0;
  }
}

void cx_batch_get_tuning(const cx_batch* h, double* out4) {
  if (h && out4) h->impl->tuning(out4);
}

void cx_batch_get_stats(cx_batch* h, int k, double* out13) {
  if (!h || !out13) return;
  try {
    h->impl->stats(k, out13);
  } catch (...) {
  }
}

void cx_batch_set_differentiable(cx_batch* h, int on) {
  if (!h) return;
  try {
    h->impl->set_differentiable(on != 0);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "[conicxx_amd] set_differentiable failed: %s\n", e.what());
  }
}

int cx_batch_backward(cx_batch* h, const double* const* dl_dx, const double* const* dl_ds,
                      const double* const* dl_dz, double* const* dP, double* const* dq, double* const* dA,
                      double* const* db) {
  if (!h) return -1;
  try {
    return h->impl->backward(dl_dx, dl_ds, dl_dz, dP, dq, dA, db);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "[conicxx_amd] batch backward failed: %s\n", e.what());
    return -1;
  }
}

double cx_batch_last_backward_ms(const cx_batch* h) { return h ? h->impl->last_backward_ms() : 0.0; }

double cx_batch_last_solve_ms(const cx_batch* h) { return h ? h->impl->last_solve_ms() : 0.0; }

int cx_equilibrate(int n, int m, const int* P_colptr, const int* P_rowind, const double* P_val,
                   const int* A_colptr, const int* A_rowind, const double* A_val, int zero_dim,
                   int nonneg_dim, int n_soc, const int* soc_dims, const cx_settings* settings, double* d,
                   double* e, double* c) {
  if (n < 0 || m < 0 || n_soc < 0 || !P_colptr || !A_colptr) return 0;
  try {
    cx_settings st;
    cx_default_settings(&st);
    if (settings) st = *settings;
    std::vector<int> soc_off(n_soc), sd(soc_dims, soc_dims + n_soc);
    int off = zero_dim + nonneg_dim;
    for (int k = 0; k < n_soc; ++k) soc_off[k] = off, off += sd[k];
    if (off != m) return 0;
    std::vector<double> Pv(P_val, P_val + P_colptr[n]), Av(A_val, A_val + A_colptr[n]), ed, ee;
    double ec = 1.0;
    cx::ruiz_equilibrate(n, m, P_colptr, P_rowind, Pv, A_colptr, A_rowind, Av, soc_off, sd, st, ed, ee, ec);
    if (d) std::copy(ed.begin(), ed.end(), d);
    if (e) std::copy(ee.begin(), ee.end(), e);
    if (c) *c = ec;
    return 1;
  } catch (const std::exception&) {
    return 0;
  }
}

int cx_analyze_pattern(int N, const int* colptr, const int* rowind, int* perm, int* colcount,
                       double* stats8) {
  try {
    cx::LowerPattern K;
    K.N = N;
    K.colptr.assign(colptr, colptr + N + 1);
    K.rowind.assign(rowind, rowind + colptr[N]);
    cx::SupernodalPlan P = cx::plan_supernodes(K);
    if (perm)
      for (int k = 0; k < N; ++k) perm[k] = P.perm[k];
    if (colcount) {
      // exact counts of the final order (a topological order of its
      // elimination tree, not necessarily a postorder: count in one)
      std::vector<int> par = cx::etree_of(K, P.perm);
      std::vector<int> cc = cx::column_counts(K, P.perm, par, cx::postorder(par));
      for (int k = 0; k < N; ++k) colcount[k] = cc[k];
    }
    if (stats8) {
      stats8[0] = P.N;
      stats8[1] = K.colptr[N];
      stats8[2] = P.ns;
      stats8[3] = P.nlevels;
      stats8[4] = P.nnzL;
      stats8[5] = P.flops;
      stats8[6] = P.max_front;
      stats8[7] = P.nnzL_exact;
    }
    return 1;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "[conicxx_amd] analyze failed: %s\n", e.what());
    return 0;
  }
}

}  // extern "C"
