// Host Ruiz equilibration of one problem (Equilibration::compute and
// scaleProblem, src/conicxx/equilibration.cpp:8-127), shared by the
// single-problem and batched solvers.
#pragma once

#include <algorithm>
#include <cmath>
#include <vector>

#include "../../include/conicxx_c.h"

namespace cx {

// Computes d (n), e (m) and the cost scale c, and scales the P (upper CSC) and
// A (CSC) values in place: P <- c D P D, A <- E A D.  With equilibration off,
// d = e = 1 and c = 1 and nothing is scaled.  SOC rows share one scale per
// block (equilibration.cpp:47-57); d, e are clamped after the passes
// (equilibration.cpp:84-85); c = 1 / max(mean |P|, min_scale), clamped, and 1
// when P has no entries.
inline void ruiz_equilibrate(int n, int m, const int* Pp, const int* Pi, std::vector<double>& Pv,
                             const int* Ap, const int* Ai, std::vector<double>& Av,
                             const std::vector<int>& soc_off, const std::vector<int>& soc_dims,
                             const cx_settings& st, std::vector<double>& ed, std::vector<double>& ee,
                             double& ec) {
  ed.assign(n, 1.0);
  ee.assign(m, 1.0);
  ec = 1.0;
  if (!st.equilibrate || n <= 0) return;
  const int nP = Pp[n], nA = Ap[n];
  std::vector<double> Pw(Pv), Aw(Av), nx(n), nz(m), dx(n), dz(m);
  const double minS = st.equilibrate_min_scale, maxS = st.equilibrate_max_scale;
  for (int it = 0; it < st.equilibrate_max_iter; ++it) {
    std::fill(nx.begin(), nx.end(), 0.0);
    std::fill(nz.begin(), nz.end(), 0.0);
    for (int c = 0; c < n; ++c)
      for (int k = Pp[c]; k < Pp[c + 1]; ++k) {
        double av = std::fabs(Pw[k]);
        nx[Pi[k]] = std::max(nx[Pi[k]], av);
        nx[c] = std::max(nx[c], av);
      }
    for (int c = 0; c < n; ++c)
      for (int k = Ap[c]; k < Ap[c + 1]; ++k) {
        double av = std::fabs(Aw[k]);
        nx[c] = std::max(nx[c], av);
        nz[Ai[k]] = std::max(nz[Ai[k]], av);
      }
    for (size_t k = 0; k < soc_dims.size(); ++k) {
      int o = soc_off[k], d = soc_dims[k];
      double mx = nz[o];
      for (int j = 1; j < d; ++j) mx = std::max(mx, nz[o + j]);
      for (int j = 0; j < d; ++j) nz[o + j] = mx;
    }
    for (int i = 0; i < n; ++i) dx[i] = 1.0 / std::sqrt(nx[i] > 0 ? nx[i] : 1.0);
    for (int i = 0; i < m; ++i) dz[i] = 1.0 / std::sqrt(nz[i] > 0 ? nz[i] : 1.0);
    for (int c = 0; c < n; ++c)
      for (int k = Pp[c]; k < Pp[c + 1]; ++k) Pw[k] *= dx[Pi[k]] * dx[c];
    for (int c = 0; c < n; ++c)
      for (int k = Ap[c]; k < Ap[c + 1]; ++k) Aw[k] *= dz[Ai[k]] * dx[c];
    for (int i = 0; i < n; ++i) ed[i] *= dx[i];
    for (int i = 0; i < m; ++i) ee[i] *= dz[i];
  }
  for (int i = 0; i < n; ++i) ed[i] = std::min(std::max(ed[i], minS), maxS);
  for (int i = 0; i < m; ++i) ee[i] = std::min(std::max(ee[i], minS), maxS);
  double sumAbs = 0.0;
  for (int k = 0; k < nP; ++k) sumAbs += std::fabs(Pw[k]);
  if (nP > 0) {
    double meanAbs = sumAbs / (double)nP;
    if (meanAbs > 0) ec = 1.0 / std::max(meanAbs, minS);
  }
  ec = std::min(std::max(ec, minS), maxS);
  // scaleProblem (equilibration.cpp:102-127)
  for (int c = 0; c < n; ++c)
    for (int k = Pp[c]; k < Pp[c + 1]; ++k) Pv[k] *= ec * ed[Pi[k]] * ed[c];
  for (int c = 0; c < n; ++c)
    for (int k = Ap[c]; k < Ap[c + 1]; ++k) Av[k] *= ee[Ai[k]] * ed[c];
  (void)nA;
}

}  // namespace cx
