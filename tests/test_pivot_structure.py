"""Synthetic matrices with a prescribed RREF pivot structure; the oracle must
recover it (CPU)."""
import numpy
import pytest

from oracle import crref, modp, pivots

KNOWN_MEDIUM_PIVOTS = [0, 1, 2, 3, 4, 5, 6, 7, 8, 10, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22, 23, 24, 25, 26, 27,
                       38, 39, 40, 41, 59]  # reference tests/test_null_space.py:27


def planted_pivot_matrix(field, seed, rows=32, cols=86, piv=KNOWN_MEDIUM_PIVOTS):
    """E @ R with R in RREF with pivot columns `piv`, E a random rows x rank matrix."""
    rs = numpy.random.RandomState(seed)
    r = len(piv)
    if field:
        R = numpy.zeros((r, cols), dtype=numpy.int64)
        for k, c in enumerate(piv):
            R[k, c] = 1
            free = [j for j in range(c + 1, cols) if j not in piv]
            R[k, free] = rs.randint(0, field, size=len(free))
        E = rs.randint(0, field, size=(rows, r))
        return modp.matmul_modp(E, R, field)
    R = numpy.zeros((r, cols), dtype=complex)
    for k, c in enumerate(piv):
        R[k, c] = 1
        free = [j for j in range(c + 1, cols) if j not in piv]
        R[k, free] = rs.standard_normal(len(free)) + 1j * rs.standard_normal(len(free))
    E = rs.standard_normal((rows, r)) + 1j * rs.standard_normal((rows, r))
    return E @ R


@pytest.mark.parametrize("field", [0, 2 ** 31 - 1])
def test_oracle_recovers_planted_pivots(field):
    A = planted_pivot_matrix(field, seed=3)
    if field:
        R, piv = crref.rref_modp(A, field)
    else:
        R, piv = crref.rref_c128(A)
    assert piv == KNOWN_MEDIUM_PIVOTS
    R = pivots.drop_bottom_zero_rows(R)
    assert R.shape == (31, 86)
    assert pivots.pivot_columns(R) == KNOWN_MEDIUM_PIVOTS
