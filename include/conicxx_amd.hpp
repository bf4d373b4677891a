// conicxx_amd.hpp -- header-only C++17 facade over the C-ABI (conicxx_c.h).
//
// Mirrors the reference's public classes (include/conicxx/solver.h:33-95,
// settings.h:21-62, cone_spec.h:15-33, solution.h:9-39) member for member, so
// a caller of conicxx::Solver switches by changing the include and the
// namespace.  Matrices are taken through any type with the compressed-column
// interface of Eigen::SparseMatrix<double, ColMajor, int> (rows(), cols(),
// nonZeros(), outerIndexPtr(), innerIndexPtr(), valuePtr()) or through the
// plain CscView below; vectors through any type with data() and size().
// Eigen itself is not required.  Like the reference, nothing here throws from
// setup/updateData/solve; a missing GPU makes setup() return false.
#pragma once

#include <string>
#include <utility>
#include <vector>

#include "conicxx_c.h"

namespace conicxx_amd {

using Scalar = double;
using Index = int;

// Non-owning compressed-column matrix (int32 indices, double values).
struct CscView {
  Index nrows = 0, ncols = 0;
  const Index* colptr = nullptr;  // ncols + 1
  const Index* rowind = nullptr;  // colptr[ncols]
  const Scalar* values = nullptr;
  Index rows() const { return nrows; }
  Index cols() const { return ncols; }
  const Index* outerIndexPtr() const { return colptr; }
  const Index* innerIndexPtr() const { return rowind; }
  const Scalar* valuePtr() const { return values; }
};

// Owning compressed-column matrix: what the conversions below return.
struct CscMatrix {
  Index nrows = 0, ncols = 0;
  std::vector<Index> colptr, rowind;
  std::vector<Scalar> values;
  Index rows() const { return nrows; }
  Index cols() const { return ncols; }
  Index nonZeros() const { return ncols > 0 ? colptr[ncols] : 0; }
  const Index* outerIndexPtr() const { return colptr.data(); }
  const Index* innerIndexPtr() const { return rowind.data(); }
  const Scalar* valuePtr() const { return values.data(); }
  // entry (i, j), 0 where the pattern has none
  Scalar coeff(Index i, Index j) const {
    for (Index e = colptr[j]; e < colptr[j + 1]; ++e)
      if (rowind[e] == i) return values[e];
    return 0.0;
  }
};

// Non-owning dense matrix, column-major unless col_major is false.
struct DenseView {
  Index nrows = 0, ncols = 0;
  const Scalar* data = nullptr;
  bool col_major = true;
  Index rows() const { return nrows; }
  Index cols() const { return ncols; }
};

// Non-owning row-major (CSR) matrix.
struct CsrView {
  Index nrows = 0, ncols = 0;
  const Index* rowptr = nullptr;  // nrows + 1
  const Index* colind = nullptr;  // rowptr[nrows]
  const Scalar* values = nullptr;
};

// conicxx::toSparse / toSparseUpperTriangular (include/conicxx/convert.h:17-41):
// every entry of a dense matrix explicit (zeros included, so the pattern does
// not depend on the values); the upper triangle for P (warning when it is not
// symmetric); a CSR matrix re-laid out by columns.  Each warns once on stderr
// unless warn is false.  An invalid input gives an empty 0 x 0 matrix.
inline CscMatrix toSparse(const DenseView& d, const char* name = nullptr, bool warn = true) {
  CscMatrix out;
  out.colptr.assign(static_cast<size_t>(d.ncols) + 1, 0);
  out.rowind.resize(static_cast<size_t>(d.nrows) * d.ncols);
  out.values.resize(out.rowind.size());
  if (cx_dense_to_csc(d.nrows, d.ncols, d.data, d.col_major ? 1 : 0, name, warn ? 1 : 0, out.colptr.data(),
                      out.rowind.data(), out.values.data()) < 0)
    return CscMatrix{};
  out.nrows = d.nrows, out.ncols = d.ncols;
  return out;
}

inline CscMatrix toSparseUpperTriangular(const DenseView& d, const char* name = nullptr, bool warn = true) {
  if (d.nrows != d.ncols) return CscMatrix{};
  const size_t n = static_cast<size_t>(d.nrows);
  CscMatrix out;
  out.colptr.assign(n + 1, 0);
  out.rowind.resize(n * (n + 1) / 2);
  out.values.resize(out.rowind.size());
  if (cx_dense_to_csc_upper(d.nrows, d.data, d.col_major ? 1 : 0, name, warn ? 1 : 0, out.colptr.data(),
                            out.rowind.data(), out.values.data()) < 0)
    return CscMatrix{};
  out.nrows = out.ncols = d.nrows;
  return out;
}

inline CscMatrix toSparse(const CsrView& r, const char* name = nullptr, bool warn = true) {
  if (r.nrows < 0 || r.ncols < 0 || !r.rowptr) return CscMatrix{};
  const size_t nnz = static_cast<size_t>(r.rowptr[r.nrows]);
  CscMatrix out;
  out.colptr.assign(static_cast<size_t>(r.ncols) + 1, 0);
  out.rowind.resize(nnz);
  out.values.resize(nnz);
  if (cx_csr_to_csc(r.nrows, r.ncols, r.rowptr, r.colind, r.values, name, warn ? 1 : 0, out.colptr.data(),
                    out.rowind.data(), out.values.data()) < 0)
    return CscMatrix{};
  out.nrows = r.nrows, out.ncols = r.ncols;
  return out;
}

enum class Status {
  Unsolved = CX_UNSOLVED,
  Solved = CX_SOLVED,
  PrimalInfeasible = CX_PRIMAL_INFEASIBLE,
  DualInfeasible = CX_DUAL_INFEASIBLE,
  MaxIterations = CX_MAX_ITERATIONS,
  NumericalError = CX_NUMERICAL_ERROR,
};

inline const char* toString(Status s) { return cx_status_string(static_cast<int>(s)); }

enum class LinearSolverBackend { Eigen = CX_LS_EIGEN, Qdldl = CX_LS_QDLDL, RegularizedLdlt = CX_LS_REGULARIZED };
enum class FactorPrecision { FP64 = CX_FACTOR_FP64, FP32Refined = CX_FACTOR_FP32_REFINED };

struct Settings {
  Scalar tol_feas = 1e-8;
  Scalar tol_gap_abs = 1e-8;
  Scalar tol_gap_rel = 1e-8;
  Scalar tol_infeas = 1e-7;
  int max_iter = 200;
  Scalar static_reg_P = 1e-8;
  Scalar static_reg_A = 1e-8;
  Scalar dynamic_reg_eps = 1e-14;
  Scalar dynamic_reg_delta = 1e-7;
  int refine_max_iter = 3;
  Scalar refine_tol = 1e-12;
  Scalar max_step_fraction = 0.99;
  bool equilibrate = true;
  int equilibrate_max_iter = 10;
  Scalar equilibrate_min_scale = 1e-4;
  Scalar equilibrate_max_scale = 1e4;
  bool warm_start = true;
  bool verbose = false;
  bool record_timings = false;
  bool validate_inputs = true;
  LinearSolverBackend linear_solver = LinearSolverBackend::Qdldl;
  FactorPrecision factor_precision = FactorPrecision::FP64;
  Scalar fp32_reg = 1e-4;

  cx_settings to_c() const {
    cx_settings c;
    c.tol_feas = tol_feas;
    c.tol_gap_abs = tol_gap_abs;
    c.tol_gap_rel = tol_gap_rel;
    c.tol_infeas = tol_infeas;
    c.max_iter = max_iter;
    c.static_reg_P = static_reg_P;
    c.static_reg_A = static_reg_A;
    c.dynamic_reg_eps = dynamic_reg_eps;
    c.dynamic_reg_delta = dynamic_reg_delta;
    c.refine_max_iter = refine_max_iter;
    c.refine_tol = refine_tol;
    c.max_step_fraction = max_step_fraction;
    c.equilibrate = equilibrate;
    c.equilibrate_max_iter = equilibrate_max_iter;
    c.equilibrate_min_scale = equilibrate_min_scale;
    c.equilibrate_max_scale = equilibrate_max_scale;
    c.warm_start = warm_start;
    c.verbose = verbose;
    c.record_timings = record_timings;
    c.validate_inputs = validate_inputs;
    c.linear_solver = static_cast<int>(linear_solver);
    c.factor_precision = static_cast<int>(factor_precision);
    c.fp32_reg = fp32_reg;
    return c;
  }
};

struct ConeSpec {
  Index zero_dim = 0;
  Index nonneg_dim = 0;
  std::vector<Index> soc_dims;
  Index totalDim() const {
    Index t = zero_dim + nonneg_dim;
    for (Index d : soc_dims) t += d;
    return t;
  }
  bool isValid() const {
    if (zero_dim < 0 || nonneg_dim < 0) return false;
    for (Index d : soc_dims)
      if (d < 1) return false;
    return true;
  }
};

struct Info {
  int iterations = 0;
  Scalar primal_residual = 0;
  Scalar dual_residual = 0;
  Scalar duality_gap = 0;
  Scalar mu = 0;
  Scalar setup_time_s = 0;
  Scalar solve_time_s = 0;
};

struct Solution {
  Status status = Status::Unsolved;
  std::vector<Scalar> x, s, z;
  Scalar objective = 0;
  Info info;
  bool ok() const { return status == Status::Solved; }
};

class Solver {
 public:
  Solver() : h_(cx_create()) {}
  ~Solver() { cx_destroy(h_); }
  Solver(Solver&& o) noexcept { *this = std::move(o); }
  Solver& operator=(Solver&& o) noexcept {
    if (this != &o) {
      cx_destroy(h_);
      h_ = o.h_;
      o.h_ = nullptr;
      st_ = o.st_;
      sol_ = std::move(o.sol_);
      n_ = o.n_;
      m_ = o.m_;
    }
    return *this;
  }
  Solver(const Solver&) = delete;
  Solver& operator=(const Solver&) = delete;

  // false without a GPU or on invalid input (solver_impl.cpp:11-16)
  template <class SpP, class VecQ, class SpA, class VecB>
  bool setup(const SpP& P, const VecQ& q, const SpA& A, const VecB& b, const ConeSpec& cone,
             const Settings& settings = Settings{}) {
    if (!h_) return false;
    st_ = settings;
    const Index n = static_cast<Index>(P.cols()), m = static_cast<Index>(A.rows());
    if (P.rows() != n || A.cols() != n || static_cast<Index>(q.size()) != n ||
        static_cast<Index>(b.size()) != m || !cone.isValid() || cone.totalDim() != m)
      return false;
    const cx_settings c = settings.to_c();
    const int ok = cx_setup(h_, n, m, P.outerIndexPtr(), P.innerIndexPtr(), P.valuePtr(), q.data(),
                            A.outerIndexPtr(), A.innerIndexPtr(), A.valuePtr(), b.data(), cone.zero_dim,
                            cone.nonneg_dim, static_cast<int>(cone.soc_dims.size()),
                            cone.soc_dims.empty() ? nullptr : cone.soc_dims.data(), &c);
    if (ok) n_ = n, m_ = m;
    return ok != 0;
  }

  // dense P and/or A (src/conicxx/solver.cpp:17-30): converted with
  // toSparseUpperTriangular / toSparse, which warn on stderr.  There is no
  // dense updateData (solver.h:55-62): convert once and update the values.
  template <class VecQ, class SpA, class VecB>
  bool setup(const DenseView& P, const VecQ& q, const SpA& A, const VecB& b, const ConeSpec& cone,
             const Settings& settings = Settings{}) {
    const CscMatrix Pc = toSparseUpperTriangular(P, "P");
    return Pc.ncols == P.ncols && setup(Pc, q, A, b, cone, settings);
  }
  template <class SpP, class VecQ, class VecB>
  bool setup(const SpP& P, const VecQ& q, const DenseView& A, const VecB& b, const ConeSpec& cone,
             const Settings& settings = Settings{}) {
    const CscMatrix Ac = toSparse(A, "A");
    return Ac.ncols == A.ncols && setup(P, q, Ac, b, cone, settings);
  }
  template <class VecQ, class VecB>
  bool setup(const DenseView& P, const VecQ& q, const DenseView& A, const VecB& b, const ConeSpec& cone,
             const Settings& settings = Settings{}) {
    const CscMatrix Pc = toSparseUpperTriangular(P, "P"), Ac = toSparse(A, "A");
    return Pc.ncols == P.ncols && Ac.ncols == A.ncols && setup(Pc, q, Ac, b, cone, settings);
  }

  // nullptr = unchanged (solver.h:77); false on pattern mismatch or failure
  template <class SpP, class VecQ, class SpA, class VecB>
  bool updateData(const SpP* P, const VecQ* q, const SpA* A, const VecB* b) {
    if (!h_) return false;
    if ((q && static_cast<Index>(q->size()) != n_) || (b && static_cast<Index>(b->size()) != m_)) return false;
    return cx_update_data(h_, P ? P->outerIndexPtr() : nullptr, P ? P->innerIndexPtr() : nullptr,
                          P ? P->valuePtr() : nullptr, q ? q->data() : nullptr,
                          A ? A->outerIndexPtr() : nullptr, A ? A->innerIndexPtr() : nullptr,
                          A ? A->valuePtr() : nullptr, b ? b->data() : nullptr) != 0;
  }

  template <class V>
  void setWarmStart(const V& x, const V& s, const V& z) {
    if (h_) cx_set_warm_start(h_, x.data(), s.data(), z.data());
  }

  void setSettings(const Settings& settings) {
    st_ = settings;
    if (h_) {
      const cx_settings c = settings.to_c();
      cx_set_settings(h_, &c);
    }
  }
  const Settings& settings() const { return st_; }

  const Solution& solve() {
    if (!h_) {
      sol_ = Solution{};
      sol_.status = Status::NumericalError;
      return sol_;
    }
    cx_solve(h_);
    return fetch();
  }
  const Solution& solution() const { return sol_; }

 private:
  const Solution& fetch() {
    sol_.x.assign(n_, 0.0);
    sol_.s.assign(m_, 0.0);
    sol_.z.assign(m_, 0.0);
    cx_info info;
    sol_.status = static_cast<Status>(
        cx_get_solution(h_, sol_.x.data(), sol_.s.data(), sol_.z.data(), &sol_.objective, &info));
    sol_.info = Info{info.iterations, info.primal_residual, info.dual_residual, info.duality_gap,
                     info.mu,         info.setup_time_s,    info.solve_time_s};
    return sol_;
  }

  cx_solver* h_ = nullptr;
  Settings st_{};
  Solution sol_{};
  Index n_ = 0, m_ = 0;
};

}  // namespace conicxx_amd
