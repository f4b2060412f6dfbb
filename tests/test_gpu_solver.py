"""``iterative_gaussian_solver`` on the device against the oracle's
restatement of the reference loop (``linac/linear_system_solver.py:24-79``):
GF(p) solutions equal exactly, float solutions within 1e-9 with the same
dropped unknowns.  The systems plant zero unknowns, duplicated columns and a
cascade (a column that becomes redundant only once a zero unknown's row is
gone, which changes the next iteration's solution)."""
import numpy
import pytest

from oracle import modp, solver

pytestmark = pytest.mark.gpu


def _system(n, p, seed, zeros, dups, cascade=None):
    rs = numpy.random.RandomState(seed)
    if p:
        A = rs.randint(0, min(p, 2 ** 62), size=(n, n)).astype(numpy.int64) % p
        x = rs.randint(1, min(p, 2 ** 62), size=n).astype(numpy.int64) % p
    else:
        A = rs.standard_normal((n, n))
        x = rs.standard_normal(n)
    x[list(zeros)] = 0
    for c, d in dups:
        A[:, d] = A[:, c]
    if cascade is not None:  # column e equals column c except in row r, whose unknown is zero
        c, e, r = cascade
        A[:, e] = A[:, c]
        A[r, e] = (A[r, c] + 1) % p if p else A[r, c] + 1.0
        x[r] = 0
    if p and p < 2 ** 31:
        b = modp.matmul_modp(A, x.reshape(-1, 1), p)[:, 0]
    elif p:
        b = numpy.array((A.astype(object) @ x.astype(object)) % p, dtype=numpy.int64)
    else:
        b = A @ x
    M = numpy.concatenate([A, b.reshape(-1, 1)], axis=1)
    if p:
        M = M.astype(numpy.uint32 if p < 2 ** 32 else numpy.uint64)
    return M


CASES = [
    (2 ** 31 - 1, 300, 1, {7, 100, 299}, [(3, 50), (60, 61)], None),
    (2 ** 31 - 1, 257, 2, {10}, [], (20, 30, 10)),
    (4294967291, 200, 3, {0, 5}, [(1, 2)], (40, 41, 5)),
    (2 ** 61 - 1, 120, 4, {9}, [(11, 12)], (13, 14, 9)),
    (0, 400, 5, {1, 200}, [(2, 3)], (50, 51, 1)),
]


@pytest.mark.parametrize("iters", [1, 5])
@pytest.mark.parametrize("p, n, seed, zeros, dups, cascade", CASES)
def test_solver_matches_oracle(gpu, p, n, seed, zeros, dups, cascade, iters):
    from linac_amd.linear_system_solver import iterative_gaussian_solver
    M = _system(n, p, seed, zeros, dups, cascade)
    want, hist = solver.iterative_gaussian_solver(M, max_iterations=iters, field_characteristic=p)
    got = iterative_gaussian_solver(M, max_iterations=iters, field_characteristic=p)
    assert len(got) == n
    if p:
        assert got == want
    else:
        zw = [i for i, v in enumerate(want) if v == 0]
        zg = [i for i, v in enumerate(got) if v == 0]
        assert zg == zw
        assert numpy.allclose(numpy.array(got, dtype=float), numpy.array(want, dtype=float), rtol=1e-9, atol=1e-9)
    if cascade is not None and iters > 1:
        assert len(hist) >= 3  # the cascade drops a column in the second pass


def test_solver_complex_and_small_edges(gpu):
    from linac_amd.linear_system_solver import iterative_gaussian_solver
    rs = numpy.random.RandomState(6)
    A = rs.standard_normal((60, 60)) + 1j * rs.standard_normal((60, 60))
    x = rs.standard_normal(60) + 1j * rs.standard_normal(60)
    x[[4, 44]] = 0
    M = numpy.concatenate([A, (A @ x).reshape(-1, 1)], axis=1)
    want, _ = solver.iterative_gaussian_solver(M, max_iterations=3)
    got = iterative_gaussian_solver(M, max_iterations=3)
    assert [i for i, v in enumerate(got) if v == 0] == [i for i, v in enumerate(want) if v == 0] == [4, 44]
    assert numpy.allclose(numpy.array(got), numpy.array(want), rtol=1e-9, atol=1e-9)
    # 1 x 2 and all-zero systems
    assert iterative_gaussian_solver(numpy.array([[3, 6]], dtype=numpy.uint32), field_characteristic=7) == [2]
    assert iterative_gaussian_solver(numpy.zeros((5, 6), dtype=numpy.uint32), max_iterations=3,
                                     field_characteristic=7) == [0] * 5
    with pytest.raises(ValueError):
        iterative_gaussian_solver(numpy.zeros((5, 5)))
