// convert.cpp -- dense and row-major (CSR) inputs to the solver's canonical
// compressed-column layout (host code; no device work).
//
// The conversions of include/conicxx/convert.h / src/conicxx/convert.cpp:22-79:
//   * a dense matrix keeps every entry explicit, exact zeros included, so the
//     pattern of a converted matrix does not depend on its values and stays
//     valid for updateData() across conversions;
//   * the upper-triangle conversion for P keeps the n(n+1)/2 entries i <= j and
//     warns when the input is not numerically symmetric (the lower triangle is
//     ignored);
//   * a CSR matrix is re-laid out column by column (rows ascending).
// Each conversion prints a one-line warning on stderr unless `warn` is 0:
// dense input carries no sparsity for the factorization to exploit.
#include <algorithm>
#include <cmath>
#include <cstdio>

#include "conicxx_c.h"

namespace {

void warn_line(const char* what, const char* name, int warn) {
  if (!warn) return;
  if (name && *name)
    std::fprintf(stderr, "[conicxx_amd] warning: %s (matrix: %s)\n", what, name);
  else
    std::fprintf(stderr, "[conicxx_amd] warning: %s\n", what);
}

inline double at(const double* d, int rows, int cols, int col_major, int i, int j) {
  return col_major ? d[(long)j * rows + i] : d[(long)i * cols + j];
}

}  // namespace

extern "C" {

long cx_dense_to_csc(int rows, int cols, const double* dense, int col_major, const char* name, int warn,
                     int* colptr, int* rowind, double* values) {
  if (rows < 0 || cols < 0 || (!dense && rows && cols) || !colptr) return -1;
  if ((long)rows * cols > 0x7fffffffL) return -1;  // int32 indices
  warn_line(
      "dense matrix converted to compressed columns with every entry explicit (zeros included, so "
      "the pattern is the same for any values); no sparsity is left to exploit: pass a sparse "
      "matrix for large or sparse problems",
      name, warn);
  long e = 0;
  for (int j = 0; j < cols; ++j) {
    colptr[j] = (int)e;
    for (int i = 0; i < rows; ++i, ++e) {
      rowind[e] = i;
      values[e] = at(dense, rows, cols, col_major, i, j);
    }
  }
  colptr[cols] = (int)e;
  return e;
}

long cx_dense_to_csc_upper(int n, const double* dense, int col_major, const char* name, int warn, int* colptr,
                           int* rowind, double* values) {
  if (n < 0 || (!dense && n) || !colptr) return -1;
  if ((long)n * (n + 1) / 2 > 0x7fffffffL) return -1;
  if (warn && n > 0) {
    double asym = 0.0, big = 1.0;
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) {
        const double a = at(dense, n, n, col_major, i, j);
        asym = std::max(asym, std::fabs(a - at(dense, n, n, col_major, j, i)));
        big = std::max(big, std::fabs(a));
      }
    if (asym > 1e-8 * big)
      warn_line(
          "the matrix given as P is not numerically symmetric: only its upper triangle is read and "
          "the lower triangle is ignored; check that P equals its transpose",
          name, warn);
  }
  warn_line(
      "dense matrix converted to compressed columns (upper triangle, every entry explicit) for P; "
      "no sparsity is left to exploit: pass a sparse matrix for large or sparse problems",
      name, warn);
  long e = 0;
  for (int j = 0; j < n; ++j) {
    colptr[j] = (int)e;
    for (int i = 0; i <= j; ++i, ++e) {
      rowind[e] = i;
      values[e] = at(dense, n, n, col_major, i, j);
    }
  }
  colptr[n] = (int)e;
  return e;
}

long cx_csr_to_csc(int rows, int cols, const int* rowptr, const int* colind, const double* values_in,
                   const char* name, int warn, int* colptr, int* rowind, double* values) {
  if (rows < 0 || cols < 0 || !rowptr || !colptr) return -1;
  const long nnz = rowptr[rows];
  for (int i = 0; i < rows; ++i)
    if (rowptr[i + 1] < rowptr[i]) return -1;
  for (long e = 0; e < nnz; ++e)
    if (colind[e] < 0 || colind[e] >= cols) return -1;
  warn_line(
      "row-major (CSR) matrix re-laid out as compressed columns; build the matrix in compressed "
      "columns to avoid this one-time conversion",
      name, warn);
  // counting sort by column; rows come out ascending within each column
  std::fill(colptr, colptr + cols + 1, 0);
  for (long e = 0; e < nnz; ++e) ++colptr[colind[e] + 1];
  for (int j = 0; j < cols; ++j) colptr[j + 1] += colptr[j];
  for (int i = 0; i < rows; ++i)
    for (int e = rowptr[i]; e < rowptr[i + 1]; ++e) {
      const int d = colptr[colind[e]]++;
      rowind[d] = i;
      values[d] = values_in[e];
    }
  for (int j = cols; j > 0; --j) colptr[j] = colptr[j - 1];
  colptr[0] = 0;
  return nnz;
}

}  // extern "C"
