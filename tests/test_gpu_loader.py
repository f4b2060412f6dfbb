"""Matrix loader on the GPU against the loader oracle and the reference's
known solution (reference tests/test_matrix_loader.py:9-57)."""
import numpy
import pytest

from oracle import crref, loader
from test_loader_oracle import ANSATZE, KNOWN, PREFACTORS, Field, dict_points

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("p", [2 ** 31 - 1, 2 ** 31 - 19, 2 ** 31 + 11, 2 ** 61 - 1, 2 ** 61 + 15])
def test_load_matrices_known_solution(gpu, p):
    n = sum(map(len, ANSATZE))
    pts, f = dict_points(p, n, 0)
    As = gpu.load_matrices(PREFACTORS, ANSATZE, pts)
    Ref = loader.load_matrices_reference(PREFACTORS, ANSATZE, pts)
    assert [A.shape for A in As] == [A.shape for A in Ref]
    for A, B in zip(As, Ref):
        # the reference's GPU loader stores uint64 for p > 2**31-1
        # (pycuda_matrix_loader.py:63) while its CPU fallback, restated by the
        # oracle, switches at 2**32-1 (:223): compare values, check the GPU rule
        assert A.dtype == (numpy.uint64 if p > 2 ** 31 - 1 else numpy.uint32) or A.size == 0
        assert numpy.array_equal(A.astype(numpy.uint64), B.astype(numpy.uint64))
    Ab = numpy.block([A.astype(numpy.uint64) for A in As] + [f.astype(numpy.uint64).reshape(-1, 1)])
    R = gpu.cuda_row_reduce(Ab, p)
    assert numpy.array_equal(R[:, :-1], numpy.identity(n, dtype=R.dtype))
    assert R[:, -1].tolist() == KNOWN  # the check the reference test forgot to assert


@pytest.mark.parametrize("field", [Field("mpf", 0, 64), Field("mpc", 0, 64)])
def test_load_matrices_float_complex(gpu, field):
    n = sum(map(len, ANSATZE))
    rs = numpy.random.RandomState(2)
    x, y, z = (rs.random_sample(n) + 0.5 for _ in range(3))
    if field.name == "mpc":
        x = x + 1j * rs.random_sample(n)
    pts = {'field': field, '1': numpy.ones(n), 'x': x, 'y': y, 'z': z,
           PREFACTORS[0]: x / (x ** 2 + y + z), PREFACTORS[1]: y / (x + y + z ** 2), PREFACTORS[2]: z / (x + y ** 2 + z)}
    As = gpu.load_matrices(PREFACTORS, ANSATZE, pts)
    Ref = loader.load_matrices_reference(PREFACTORS, ANSATZE, pts)
    for A, B in zip(As, Ref):
        if B.size:
            assert numpy.allclose(A, B.astype(A.dtype), rtol=1e-14, atol=0)
    assert As[0].dtype == (numpy.float64 if field.name == "mpf" else numpy.complex128)
    f = pts[PREFACTORS[0]] * (x + 2 * y + 3 * z + 5 * x * y * z) + pts[PREFACTORS[1]] * (x * x + 7 * x * y + 11 * y * z)
    Ab = numpy.block([As[0], As[1], f.reshape(-1, 1)])
    sol = numpy.linalg.solve(Ab[:, :-1], Ab[:, -1])
    assert numpy.allclose(sol, KNOWN, atol=1e-6)


@pytest.mark.parametrize("p, dtype", [(2 ** 31 - 1, numpy.uint32), (65537, numpy.uint32), (2, numpy.uint32),
                                      (2 ** 61 - 1, numpy.uint64), (2 ** 31 + 11, numpy.uint64)])
def test_cuda_load_matrix_random_vs_oracle(gpu, p, dtype):
    rs = numpy.random.RandomState(p % 1000)
    bases = rs.randint(0, 2 ** 62, size=(777, 13), dtype=numpy.int64).astype(object)
    lind = rs.randint(0, 13, size=(1031, 7)).astype(numpy.uint32)
    A = gpu.cuda_load_matrix(bases, lind, (777, 1031), field_characteristic=p)
    assert A.dtype == dtype
    ref = crref.load_matrix_modp(numpy.array(bases % p, dtype=numpy.uint64), lind, p)
    assert numpy.array_equal(A.astype(numpy.uint64), ref)


def test_stacked_systems_one_launch(gpu):
    p = 2 ** 31 - 1
    rs = numpy.random.RandomState(4)
    S, rows, bl = 8, 300, 10
    stack = rs.randint(0, p, size=(S, rows, bl)).astype(numpy.uint32)
    lind = rs.randint(0, bl, size=(250, 7)).astype(numpy.uint32)
    out = gpu.load_matrices_stacked(stack, lind, p)
    assert out.shape == (S, rows, 250)
    for s in range(S):
        assert numpy.array_equal(out[s], gpu.cuda_load_matrix(stack[s], lind, field_characteristic=p))
        assert numpy.array_equal(out[s].astype(numpy.uint64),
                                 crref.load_matrix_modp(stack[s].astype(numpy.uint64), lind, p))


def test_long_basis_and_degree_edges(gpu):
    p = 2 ** 31 - 1
    rs = numpy.random.RandomState(5)
    bases = rs.randint(0, p, size=(40, 3000)).astype(numpy.uint32)
    lind = rs.randint(0, 3000, size=(70, 32)).astype(numpy.uint32)
    A = gpu.cuda_load_matrix(bases, lind, field_characteristic=p)
    assert numpy.array_equal(A.astype(numpy.uint64), crref.load_matrix_modp(bases.astype(numpy.uint64), lind, p))
    one = gpu.cuda_load_matrix(bases[:, :5], numpy.array([[4]], dtype=numpy.uint32), field_characteristic=p)
    assert numpy.array_equal(one[:, 0], bases[:, 4])
