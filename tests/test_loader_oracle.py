"""The loader oracle pinned by the reference's known solution
(reference tests/test_matrix_loader.py:9-22,54-57) with dict-form points."""
import collections

import numpy
import pytest

from oracle import crref, loader

Field = collections.namedtuple("Field", "name characteristic digits")

PREFACTORS = ['(x)/(x^2+y+z)', '(y)/(x+y+z^2)', '(z)/(x+y^2+z)']
ANSATZE = [
    [['x'], ['y'], ['z'], ['x', 'x'], ['x', 'y'], ['x', 'z'], ['y', 'y'], ['y', 'z'], ['x', 'y', 'z']],
    [['x'], ['y'], ['z'], ['x', 'x'], ['x', 'y'], ['x', 'z'], ['y', 'y'], ['y', 'z'], ['x', 'y', 'z']],
    [],
]
KNOWN = [1, 2, 3, 0, 0, 0, 0, 0, 5, 0, 0, 0, 1, 7, 0, 0, 11, 0]


def dict_points(p, n, seed):
    """Points as a dict of columns, prefactors evaluated exactly mod p."""
    rs = numpy.random.RandomState(seed)
    x, y, z = (rs.randint(1, p - 1, size=n).astype(object) for _ in range(3))
    inv = lambda v: numpy.array([pow(int(t), p - 2, p) for t in v], dtype=object)  # noqa: E731
    pts = {'field': Field("finite field", p, 1), '1': numpy.ones(n, dtype=object), 'x': x, 'y': y, 'z': z}
    pts[PREFACTORS[0]] = x * inv((x * x + y + z) % p) % p
    pts[PREFACTORS[1]] = y * inv((x + y + z * z) % p) % p
    pts[PREFACTORS[2]] = z * inv((x + y * y + z) % p) % p
    # f(x, y, z) of the reference test
    f = (pts[PREFACTORS[0]] * (x + 2 * y + 3 * z + 5 * x * y * z) + pts[PREFACTORS[1]] * (x * x + 7 * x * y + 11 * y * z)) % p
    return pts, f


@pytest.mark.parametrize("p", [2 ** 31 - 1, 2 ** 31 - 19, 2 ** 31 + 11, 2 ** 61 - 1])
def test_loader_oracle_gives_known_solution(p):
    n = sum(map(len, ANSATZE))
    pts, f = dict_points(p, n, 0)
    As = loader.load_matrices_reference(PREFACTORS, ANSATZE, pts)
    assert [A.shape for A in As] == [(18, 9), (18, 9), (18, 0)]
    assert As[0].dtype == (numpy.uint64 if p > 2 ** 32 - 1 else numpy.uint32)
    Ab = numpy.concatenate([A.astype(object) for A in As] + [f.reshape(-1, 1)], axis=1)
    R, piv = crref.rref_modp(Ab, p)
    assert piv == list(range(18))
    assert R[:, -1].tolist() == KNOWN


def test_c_and_python_loader_agree():
    p = 2 ** 31 - 1
    rs = numpy.random.RandomState(3)
    bases = rs.randint(0, p, size=(50, 11)).astype(numpy.uint64)
    lind = rs.randint(0, 11, size=(70, 7)).astype(numpy.uint32)
    assert numpy.array_equal(crref.load_matrix_modp(bases, lind, p), loader.load_reference(bases, lind, p).astype(numpy.uint64))
