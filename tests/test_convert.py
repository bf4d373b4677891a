"""Input conversions of the library (cx_dense_to_csc, cx_dense_to_csc_upper,
cx_csr_to_csc through conicxx_amd.convert), host code only: the cases of the
reference's test/convert/test_convert.cpp:9-104."""
import numpy as np
import pytest
import scipy.sparse as sp

from conicxx_amd.convert import to_sparse, to_sparse_upper_triangular


def test_dense_keeps_every_entry_including_zeros(capfd):
    A = np.array([[1.0, 0.0, 3.0], [0.0, 5.0, 0.0]])
    S = to_sparse(A)
    assert capfd.readouterr().err != ""
    assert S.shape == (2, 3) and S.nnz == 6
    assert np.array_equal(S.toarray(), A)
    assert np.array_equal(S.indptr, [0, 2, 4, 6]) and np.array_equal(S.indices, [0, 1, 0, 1, 0, 1])


def test_dense_pattern_stable_across_changing_zeros():
    S1 = to_sparse(np.array([[1.0, 2.0], [3.0, 4.0]]), warn=False)
    S2 = to_sparse(np.array([[1.0, 0.0], [3.0, 4.0]]), warn=False)
    assert S1.nnz == S2.nnz
    assert np.array_equal(S1.indptr, S2.indptr) and np.array_equal(S1.indices, S2.indices)


def test_warn_false_suppresses_warning(capfd):
    S = to_sparse(np.eye(2), "A", warn=False)
    assert capfd.readouterr().err == ""
    assert S.nnz == 4


def test_upper_triangular_extracts_only_upper_triangle(capfd):
    P = np.array([[4.0, 1.0, 2.0], [1.0, 5.0, 3.0], [2.0, 3.0, 6.0]])
    S = to_sparse_upper_triangular(P)
    capfd.readouterr()
    assert S.nnz == 6
    assert np.array_equal(S.toarray(), np.triu(P))


def test_upper_triangular_warns_on_asymmetric_input(capfd):
    to_sparse_upper_triangular(np.array([[1.0, 2.0], [0.0, 1.0]]), "P")
    err = capfd.readouterr().err
    assert "not numerically symmetric" in err and "(matrix: P)" in err


def test_upper_triangular_no_asymmetry_warning_on_symmetric_input(capfd):
    to_sparse_upper_triangular(np.array([[1.0, 2.0], [2.0, 1.0]]), "P")
    err = capfd.readouterr().err
    assert "not numerically symmetric" not in err and err != ""


def test_row_major_converts_to_column_major(capfd):
    csr = sp.csr_matrix((np.array([1.0, 2.0, 3.0]), np.array([0, 1, 1]), np.array([0, 2, 3])), shape=(2, 2))
    csc = to_sparse(csr)
    assert capfd.readouterr().err != ""
    assert isinstance(csc, sp.csc_matrix) and csc.shape == (2, 2)
    assert np.array_equal(csc.toarray(), [[1.0, 2.0], [0.0, 3.0]])
    assert np.array_equal(csc.indptr, [0, 1, 3]) and np.array_equal(csc.indices, [0, 0, 1])


def test_row_major_random_matches_scipy():
    M = sp.random(37, 23, density=0.2, random_state=1, format="csr")
    C = to_sparse(M, warn=False)
    ref = sp.csc_matrix(M)
    ref.sort_indices()
    assert np.array_equal(C.indptr, ref.indptr) and np.array_equal(C.indices, ref.indices)
    assert np.array_equal(C.data, ref.data)


def test_invalid_inputs():
    with pytest.raises(ValueError):
        to_sparse_upper_triangular(np.ones((2, 3)))
    with pytest.raises(ValueError):
        to_sparse(np.ones(3))
