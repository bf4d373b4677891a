// Drives the header-only C++ facade (include/conicxx_amd.hpp) the way a caller
// of the reference's conicxx::Solver would: the closed-form QPs of the
// reference's solver tests (test/solver/test_solve_qp.cpp:9-66) and a
// per-timestep updateData()+solve() loop on a box QP whose optimum is
// max(c, 0) (test/solver/test_update_reuse.cpp:32-73).
// Prints one line per check and exits 0 if all pass; exits 3 when no GPU is
// present (setup() returns false, as documented).
#include <cmath>
#include <cstdio>
#include <vector>

#include "conicxx_amd.hpp"

namespace cx = conicxx_amd;

struct Csc {
  std::vector<int> p, i;
  std::vector<double> x;
  int r, c;
  cx::CscView view() const { return cx::CscView{r, c, p.data(), i.data(), x.data()}; }
};

static int fails = 0;
static void check(const char* what, bool ok) {
  std::printf("%s %s\n", ok ? "PASS" : "FAIL", what);
  if (!ok) ++fails;
}

int main() {
  cx::Settings def;
  cx_settings c;
  cx_default_settings(&c);
  check("Settings defaults equal cx_default_settings",
        c.tol_feas == def.tol_feas && c.max_iter == def.max_iter && c.refine_tol == def.refine_tol &&
            c.linear_solver == (int)def.linear_solver && c.warm_start == (int)def.warm_start);
  check("toString(Solved)", std::string(cx::toString(cx::Status::Solved)) == "Solved");

  // conversions (test/convert/test_convert.cpp:9-104), host only
  {
    const double a[6] = {1.0, 0.0, 0.0, 5.0, 3.0, 0.0};  // 2 x 3, column-major
    const cx::CscMatrix S = cx::toSparse(cx::DenseView{2, 3, a}, "A", false);
    bool same = S.rows() == 2 && S.cols() == 3 && S.nonZeros() == 6;
    for (int j = 0; j < 3; ++j)
      for (int i = 0; i < 2; ++i) same = same && S.coeff(i, j) == a[j * 2 + i];
    check("toSparse keeps every dense entry, zeros included", same);
    const double p[9] = {4.0, 1.0, 2.0, 1.0, 5.0, 3.0, 2.0, 3.0, 6.0};
    const cx::CscMatrix U = cx::toSparseUpperTriangular(cx::DenseView{3, 3, p}, "P", false);
    check("toSparseUpperTriangular keeps the 6 entries i <= j",
          U.nonZeros() == 6 && U.coeff(0, 2) == 2.0 && U.coeff(2, 0) == 0.0 && U.coeff(1, 1) == 5.0);
    const int rp[3] = {0, 2, 3}, ci[3] = {0, 1, 1};
    const double cv[3] = {1.0, 2.0, 3.0};
    const cx::CscMatrix C = cx::toSparse(cx::CsrView{2, 2, rp, ci, cv}, "A", false);
    check("toSparse(CSR) re-lays out by columns",
          C.nonZeros() == 3 && C.coeff(0, 0) == 1.0 && C.coeff(0, 1) == 2.0 && C.coeff(1, 1) == 3.0 &&
              C.coeff(1, 0) == 0.0 && C.colptr[1] == 1);
  }

  // min 0.5 (x0^2 + x1^2)  s.t. x0 + x1 = 1  ->  x = (0.5, 0.5), objective 0.25
  Csc P{{0, 1, 2}, {0, 1}, {1.0, 1.0}, 2, 2};
  Csc A{{0, 1, 2}, {0, 0}, {1.0, 1.0}, 1, 2};
  std::vector<double> q{0.0, 0.0}, b{1.0};
  cx::ConeSpec spec;
  spec.zero_dim = 1;
  cx::Solver solver;
  if (!solver.setup(P.view(), q, A.view(), b, spec)) {
    std::printf("setup returned false (no GPU device)\n");
    return fails ? 1 : 3;
  }
  const cx::Solution& s1 = solver.solve();
  check("equality QP solved", s1.ok());
  check("equality QP x", std::fabs(s1.x[0] - 0.5) < 1e-6 && std::fabs(s1.x[1] - 0.5) < 1e-6);
  check("equality QP objective", std::fabs(s1.objective - 0.25) < 1e-6);

  // min 0.5 |x - c|^2 s.t. x >= 0 (as -x + s = 0, s >= 0): x = max(c, 0)
  Csc Ab{{0, 1, 2}, {0, 1}, {-1.0, -1.0}, 2, 2};
  std::vector<double> bb{0.0, 0.0};
  cx::ConeSpec box;
  box.nonneg_dim = 2;
  const double cs[3][2] = {{1.0, 2.0}, {-1.0, 3.0}, {0.5, -0.5}};
  cx::Solver sb;
  std::vector<double> qb{-cs[0][0], -cs[0][1]};
  check("box QP setup", sb.setup(P.view(), qb, Ab.view(), bb, box));
  for (int t = 0; t < 3; ++t) {
    qb = {-cs[t][0], -cs[t][1]};
    const cx::CscView* none = nullptr;
    if (t > 0) check("box QP updateData(q)", sb.updateData(none, &qb, none, (const std::vector<double>*)nullptr));
    const cx::Solution& s = sb.solve();
    const double e0 = std::fmax(cs[t][0], 0.0), e1 = std::fmax(cs[t][1], 0.0);
    char msg[96];
    std::snprintf(msg, sizeof msg, "box QP timestep %d: x = max(c, 0)", t);
    check(msg, s.ok() && std::fabs(s.x[0] - e0) < 1e-6 && std::fabs(s.x[1] - e1) < 1e-6);
  }
  // a pattern change is rejected
  Csc Pbad{{0, 1, 3}, {0, 0, 1}, {1.0, 0.1, 1.0}, 2, 2};
  const std::vector<double>* nv = nullptr;
  const cx::CscView pv = Pbad.view();
  check("updateData rejects a new pattern", !sb.updateData(&pv, nv, (const cx::CscView*)nullptr, nv));

  // dense P and A (test/solver/test_solve_dense_input.cpp:58-78): optimum (0, 2)
  const double Pd[4] = {1.0, 0.0, 0.0, 1.0}, Ad[4] = {-1.0, 0.0, 0.0, -1.0};
  std::vector<double> qd{1.0, -2.0}, bd{0.0, 0.0};
  cx::Solver sd;
  check("dense setup", sd.setup(cx::DenseView{2, 2, Pd}, qd, cx::DenseView{2, 2, Ad}, bd, box));
  const cx::Solution& s3 = sd.solve();
  check("dense-input QP x = (0, 2)", s3.ok() && std::fabs(s3.x[0]) < 1e-6 && std::fabs(s3.x[1] - 2.0) < 1e-6);
  return fails ? 1 : 0;
}
