"""Pins the oracle's rook (pivoting=2) and total (pivoting=3) pivoting
branches (oracle/rref.py, restating linac/row_reduce.py:49-58), which the
GPU path does not use: the column permutation they track in variable_order
must give a matrix whose exact RREF (rational arithmetic) the result equals,
and the first pivot must be the one each rule names (total: the largest |x|
of the matrix; rook: the first entry above 0.1 in row-major order)."""
import fractions

import numpy
import pytest

from oracle import rref


def _exact_rref(A):
    M = [[fractions.Fraction(x) for x in row] for row in A]
    rows, cols = len(M), len(M[0])
    r = 0
    for c in range(cols):
        piv = next((i for i in range(r, rows) if M[i][c] != 0), None)
        if piv is None:
            continue
        M[r], M[piv] = M[piv], M[r]
        h = M[r][c]
        M[r] = [x / h for x in M[r]]
        for i in range(rows):
            if i != r and M[i][c] != 0:
                f = M[i][c]
                M[i] = [a - f * b for a, b in zip(M[i], M[r])]
        r += 1
        if r == rows:
            break
    return numpy.array([[float(x) for x in row] for row in M])


@pytest.mark.parametrize("pivoting", [2, 3])
@pytest.mark.parametrize("seed, shape", [(0, (5, 7)), (1, (6, 6)), (2, (7, 4)), (3, (4, 9))])
def test_column_permutation_and_rref(pivoting, seed, shape):
    rs = numpy.random.RandomState(seed)
    A = rs.randint(-9, 10, size=shape).astype(float)
    if seed == 3:
        A[:, 2] = A[:, 0] + A[:, 1]  # a dependent column
    R, order = rref.row_reduce(A, pivoting=pivoting, scaling=False)
    assert sorted(order) == list(range(shape[1]))
    want = _exact_rref(A[:, order])
    assert numpy.allclose(R, want, atol=1e-9)


def test_first_pivot_follows_each_rule():
    A = numpy.array([[0.05, 0.0, 3.0, -7.0],
                     [0.0, 0.5, -9.5, 1.0],
                     [2.0, 0.0, 0.0, 4.0]])
    _, order3 = rref.row_reduce(A, pivoting=3, scaling=False)
    assert order3[0] == 2  # |-9.5| is the largest entry
    _, order2 = rref.row_reduce(A, pivoting=2, scaling=False)
    assert order2[0] == 2  # 3.0 is the first entry above 0.1 in row-major order
    B = numpy.array([[0.0, 0.2, 5.0], [1.0, 0.0, 0.0]])
    _, order2b = rref.row_reduce(B, pivoting=2, scaling=False)
    assert order2b[0] == 1  # 0.2 precedes 5.0 in row-major order
    _, order3b = rref.row_reduce(B, pivoting=3, scaling=False)
    assert order3b[0] == 2
