"""Pivots, canonical kernels and echelon forms (CPU): the whole-array versions
in linac_amd.linear_algebra_tools against the oracle's restatement of
linac/linear_algebra_tools.py, the defining identity R K = 0, and the
reference's large-fraction round trip (tests/test_null_space.py:65-67 on its
fixture test_coo_large_fractions.json, copied to tests/golden)."""
import os
from fractions import Fraction

import numpy
import pytest

from conftest import GOLDEN
from linac_amd import linear_algebra_tools as lat
from linac_amd.sparse_matrix_tools import matrix_from_json_coo
from oracle import crref, modp, pivots


def _random_rref(rows, cols, rank, p, seed):
    A = modp.low_rank(rows, cols, rank, p, seed)
    R, piv = crref.rref_modp(A, p)
    return lat.drop_bottom_zero_rows(R.astype(numpy.int64)), piv


@pytest.mark.parametrize("shape, rank, seed", [((12, 20), 9, 0), ((30, 30), 30, 1), ((5, 40), 5, 2), ((40, 25), 11, 3)])
def test_kernel_matches_oracle_and_annihilates(shape, rank, seed):
    p = 2 ** 31 - 1
    R, piv = _random_rref(*shape, rank, p, seed)
    assert lat.pivot_columns_from_row_reduced_echelon_form(R) == piv == pivots.pivot_columns(R)
    assert lat.non_pivot_columns_from_row_reduced_echelon_form(R) == pivots.non_pivot_columns(R)
    K = lat.canonical_kernel_from_row_reduced_echelon_form(R)
    Ko = pivots.canonical_kernel(R)
    assert K.dtype == Ko.dtype and numpy.array_equal(K, Ko)
    assert K.shape == (shape[1], shape[1] - len(piv))
    # R K = 0 over GF(p) (the -1 entries are p - 1 there)
    Kp = numpy.mod(K, p)
    assert not numpy.any(modp.matmul_modp(R % p, Kp, p))
    # pivots handed over from the device give the same kernel with no scan
    assert numpy.array_equal(lat.canonical_kernel_from_row_reduced_echelon_form(R, pivots=piv), K)
    assert lat.non_pivot_columns_from_canonical_kernel(K) == pivots.kernel_non_pivot_rows(K)
    assert lat.pivot_columns_from_canonical_kernel(K) == piv
    assert numpy.array_equal(lat.row_reduced_echelon_form_from_canonical_kernel(K), R)
    assert numpy.array_equal(pivots.rref_from_canonical_kernel(K), R)


def test_large_fraction_round_trip():
    """reference tests/test_null_space.py:65-67."""
    rref = matrix_from_json_coo(os.path.join(GOLDEN, "test_coo_large_fractions.json"))
    K = lat.canonical_kernel_from_row_reduced_echelon_form(rref)
    back = lat.row_reduced_echelon_form_from_canonical_kernel(K)
    assert numpy.all(rref == back)
    assert numpy.all(K == pivots.canonical_kernel(rref))
    # R K = 0 exactly over the rationals, on a few kernel columns
    for j in range(0, K.shape[1], max(1, K.shape[1] // 5)):
        col = K[:, j]
        assert all(sum((a * b for a, b in zip(row, col)), Fraction(0)) == 0 for row in rref)


def test_float_kernel_and_permutations():
    rs = numpy.random.RandomState(4)
    R = numpy.zeros((3, 7))
    for i, c in enumerate([0, 2, 5]):
        R[i, c] = 1.0
        R[i, c + 1:] = rs.standard_normal(7 - c - 1) * (numpy.arange(c + 1, 7) != 2) * (numpy.arange(c + 1, 7) != 5)
    K = lat.canonical_kernel_from_row_reduced_echelon_form(R)
    assert numpy.allclose(R @ K, 0) and numpy.array_equal(K, pivots.canonical_kernel(R))
    assert numpy.allclose(lat.row_reduced_echelon_form_from_canonical_kernel(K), R)
    perm = [3, 0, 2, 1]
    assert lat.invert_permutation(perm) == (1, 3, 2, 0)
    P = lat.permutation_matrix_from_permutation(perm)
    assert numpy.array_equal(P @ numpy.arange(4), numpy.array(perm))
    assert lat.permutation_matrix_from_permutation([]).shape == (0, 0)


def test_full_rank_and_zero_rank_edges():
    R = numpy.identity(4, dtype=numpy.int64)
    K = lat.canonical_kernel_from_row_reduced_echelon_form(R)
    assert K.shape == (4, 0)
    assert lat.non_pivot_columns_from_canonical_kernel(K) == []
    Z = numpy.zeros((0, 3), dtype=numpy.int64)
    assert numpy.array_equal(lat.canonical_kernel_from_row_reduced_echelon_form(Z), -numpy.identity(3, dtype=int))
