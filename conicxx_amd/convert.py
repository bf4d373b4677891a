"""Input conversions to the solver's compressed-column layout
(conicxx::toSparse / toSparseUpperTriangular, include/conicxx/convert.h:17-41,
src/conicxx/convert.cpp:22-79), done by the library's host code
(cx_dense_to_csc, cx_dense_to_csc_upper, cx_csr_to_csc).

  to_sparse(dense)                 every entry explicit, exact zeros included,
                                   so the pattern does not depend on the values
  to_sparse_upper_triangular(P)    the entries i <= j (P is read as its upper
                                   triangle); warns when P is not symmetric
  to_sparse(csr)                   a row-major matrix re-laid out by columns

Each prints a one-line warning on stderr unless warn=False: dense input
carries no sparsity for the factorization to exploit.  Solver.setup() applies
them to dense or CSR arguments; updateData() takes compressed columns only (as
in the reference, solver.h:55-62), so convert once and update the values.
"""
import ctypes

import numpy as np
import scipy.sparse as sp

from . import _lib


def _name(name):
    return name.encode() if name else None


def to_sparse(M, name=None, warn=True):
    """Dense array or CSR matrix -> scipy CSC with the library's layout."""
    if sp.issparse(M):
        R = sp.csr_matrix(M)
        if not isinstance(M, (sp.csr_matrix, sp.csr_array)):
            return sp.csc_matrix(M)  # already column-major or coordinate input: no relayout
        R.sort_indices()
        rows, cols = R.shape
        colptr = np.zeros(cols + 1, dtype=np.int32)
        rowind = np.zeros(R.nnz, dtype=np.int32)
        vals = np.zeros(R.nnz)
        a1, p1 = _lib.i32(R.indptr); a2, p2 = _lib.i32(R.indices); a3, p3 = _lib.f64(R.data)
        nnz = _lib.lib().cx_csr_to_csc(rows, cols, p1, p2, p3, _name(name), int(warn),
                                       colptr.ctypes.data_as(_lib._i32p), rowind.ctypes.data_as(_lib._i32p),
                                       vals.ctypes.data_as(_lib._f64p))
        if nnz < 0:
            raise ValueError("invalid CSR matrix")
        return sp.csc_matrix((vals, rowind, colptr), shape=(rows, cols))
    D = np.asfortranarray(np.asarray(M, dtype=np.float64))
    if D.ndim != 2:
        raise ValueError("a dense matrix must be two-dimensional")
    rows, cols = D.shape
    colptr = np.zeros(cols + 1, dtype=np.int32)
    rowind = np.zeros(rows * cols, dtype=np.int32)
    vals = np.zeros(rows * cols)
    nnz = _lib.lib().cx_dense_to_csc(rows, cols, D.ctypes.data_as(_lib._f64p), 1, _name(name), int(warn),
                                     colptr.ctypes.data_as(_lib._i32p), rowind.ctypes.data_as(_lib._i32p),
                                     vals.ctypes.data_as(_lib._f64p))
    if nnz < 0:
        raise ValueError("matrix too large for 32-bit indices")
    return sp.csc_matrix((vals, rowind, colptr), shape=(rows, cols))


def to_sparse_upper_triangular(M, name=None, warn=True):
    """Square dense array -> scipy CSC holding its upper triangle (i <= j)."""
    D = np.asfortranarray(np.asarray(M, dtype=np.float64))
    if D.ndim != 2 or D.shape[0] != D.shape[1]:
        raise ValueError("P must be a square matrix")
    n = D.shape[0]
    colptr = np.zeros(n + 1, dtype=np.int32)
    rowind = np.zeros(n * (n + 1) // 2, dtype=np.int32)
    vals = np.zeros(n * (n + 1) // 2)
    nnz = _lib.lib().cx_dense_to_csc_upper(n, D.ctypes.data_as(_lib._f64p), 1, _name(name), int(warn),
                                           colptr.ctypes.data_as(_lib._i32p), rowind.ctypes.data_as(_lib._i32p),
                                           vals.ctypes.data_as(_lib._f64p))
    if nnz < 0:
        raise ValueError("matrix too large for 32-bit indices")
    return sp.csc_matrix((vals, rowind, colptr), shape=(n, n))
