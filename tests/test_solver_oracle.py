"""The oracle's restatement of ``iterative_gaussian_solver``
(``linac/linear_system_solver.py:24-79``) on systems whose answer is fixed by
construction: a planted solution with zero entries (dropped as zero, the rest
returned exactly), and a duplicated column (its later copy is redundant: the
unknown and its row are dropped and the system restricted to the others is
consistent with the planted values)."""
import numpy
import pytest

from oracle import modp, solver

P = 2 ** 31 - 1


def _planted(n, seed, zeros, dup=None):
    rs = numpy.random.RandomState(seed)
    A = rs.randint(0, P, size=(n, n)).astype(numpy.int64)
    x = rs.randint(1, P, size=n).astype(numpy.int64)
    x[list(zeros)] = 0
    if dup is not None:
        c, d = dup
        A[:, d] = A[:, c]
        x[d] = 0
    b = modp.matmul_modp(A, x.reshape(-1, 1), P)[:, 0]
    return numpy.concatenate([A, b.reshape(-1, 1)], axis=1).astype(numpy.uint32), x


@pytest.mark.parametrize("iters", [1, 3])
def test_planted_zeros(iters):
    M, x = _planted(40, 1, zeros={3, 17, 39})
    sol, hist = solver.iterative_gaussian_solver(M, max_iterations=iters, field_characteristic=P)
    assert sol == x.tolist()
    assert hist[0] == ([], [3, 17, 39])
    assert len(hist) == min(iters, 2)  # the second pass drops nothing new and stops


def test_duplicated_column_is_redundant():
    M, x = _planted(30, 2, zeros={5}, dup=(4, 11))
    sol, hist = solver.iterative_gaussian_solver(M, max_iterations=4, field_characteristic=P)
    assert hist[0] == ([11], [5])
    assert sol == x.tolist() and sol[11] == 0


def test_float_planted():
    rs = numpy.random.RandomState(3)
    A = rs.standard_normal((25, 25))
    x = rs.standard_normal(25)
    x[[2, 20]] = 0
    M = numpy.concatenate([A, (A @ x).reshape(-1, 1)], axis=1)
    sol, hist = solver.iterative_gaussian_solver(M, max_iterations=2, field_characteristic=0)
    assert hist[0] == ([], [2, 20])
    assert numpy.allclose(sol, x, atol=1e-9)


def test_everything_redundant_stops():
    M = numpy.zeros((6, 7), dtype=numpy.uint32)
    sol, hist = solver.iterative_gaussian_solver(M, max_iterations=5, field_characteristic=P)
    assert sol == [0] * 6 and hist == [(list(range(6)), [])]
