"""The oracle pinned against the reference's golden vector, sympy, brute force
and closed-form certificates.  CPU only."""
import hashlib
import json
import os

import numpy
import pytest

from conftest import GOLDEN, load_small_cases
from oracle import crref, modp, rref


def test_c_oracle_reproduces_reference_golden_vector():
    """reference tests/test_row_reduce_random_cuda.py:52-63 on the C oracle."""
    p = 2 ** 31 - 19
    numpy.random.seed(0)
    A = numpy.random.randint(p, size=(4097, 4097), dtype=numpy.uint32)[:-1, :]
    R, piv = crref.rref_modp(A, p)
    known = numpy.load(os.path.join(GOLDEN, "known_result_4096_4097_2to32m19.npy"), allow_pickle=False)
    assert piv == list(range(4096))
    assert numpy.array_equal(R[:, -1], known)
    # and the certificate that does not depend on any elimination: A[:, :n] x == A[:, n]
    x = known.astype(numpy.int64).reshape(-1, 1)
    assert numpy.array_equal(modp.matmul_modp(A[:, :4096], x, p), A[:, 4096:].astype(numpy.int64))


@pytest.mark.parametrize("case", range(len(load_small_cases())))
def test_oracles_match_sympy_fixtures(case):
    p, A, R, piv = load_small_cases()[case]
    Rc, pc = crref.rref_modp(A, p)
    assert pc == piv
    assert numpy.array_equal(Rc.astype(numpy.uint64), R)
    if p < 2 ** 31:
        Rn = rref.rref_modp(A.astype(numpy.int64), p)
        assert numpy.array_equal(Rn.astype(numpy.uint64), R)
    assert modp.check_rref_shape(R, piv, p)
    assert modp.freivalds(A.astype(numpy.int64) if p < 2 ** 31 else A, R, piv, p) if p < 2 ** 31 else True


def test_c0_config_matches_sympy_checksum():
    cfg = json.load(open(os.path.join(GOLDEN, "c0_config.json")))
    p = cfg["p"]
    rs = numpy.random.RandomState(0)
    L = rs.randint(p, size=(500, 400), dtype=numpy.uint32)
    U = rs.randint(p, size=(400, 500), dtype=numpy.uint32)
    A = modp.matmul_modp(L, U, p).astype(numpy.uint32)
    assert hashlib.sha256(A.tobytes()).hexdigest() == cfg["input_sha256"]
    R, piv = crref.rref_modp(A, p)
    assert piv == cfg["pivots"] and len(piv) == cfg["rank"] == 400
    assert hashlib.sha256(numpy.ascontiguousarray(R).tobytes()).hexdigest() == cfg["rref_sha256"]


@pytest.mark.parametrize("seed", range(4))
def test_brute_force_and_numpy_restatement_agree(seed):
    p = 10007
    rs = numpy.random.RandomState(seed)
    A = rs.randint(p, size=(9, 13))
    A[:, 1] = 0
    A[4] = A[2] * 3 % p
    Rb, pb = modp.brute_rref(A, p)
    Rn = rref.rref_modp(A, p)
    Rc, pc = crref.rref_modp(A, p)
    assert numpy.array_equal(Rn.astype(object), Rb) and numpy.array_equal(Rc.astype(object), Rb)
    assert pb == pc


def test_planted_mistake_is_caught():
    """A wrong entry in an otherwise valid RREF fails the Freivalds certificate."""
    p = 2 ** 31 - 1
    A = modp.low_rank(40, 60, 25, p, 3)
    R, piv = crref.rref_modp(A, p)
    assert modp.freivalds(A, R, piv, p)
    bad = R.copy()
    bad[3, 45] = (int(bad[3, 45]) + 1) % p
    assert not modp.freivalds(A, bad, piv, p)
    bad = R.copy()
    bad[[0, 1]] = bad[[1, 0]]
    assert not modp.check_rref_shape(bad, piv, p)


def test_planted_solution_certificate():
    p = 2 ** 31 - 1
    Ab, X = modp.planted(60, p, 7, extra=2)
    R, piv = crref.rref_modp(Ab, p)
    assert piv == list(range(60))
    assert numpy.array_equal(R[:, 60:].astype(numpy.int64), X)


def test_u64_prime_oracle_against_planted():
    p = 2 ** 61 - 1
    rs = numpy.random.RandomState(5)
    A = rs.randint(0, 2 ** 61 - 1, size=(30, 30), dtype=numpy.int64)
    X = rs.randint(0, 2 ** 61 - 1, size=(30, 1), dtype=numpy.int64)
    b = numpy.array((A.astype(object) @ X.astype(object)) % p, dtype=numpy.uint64)
    R, piv = crref.rref_modp(numpy.concatenate([A.astype(numpy.uint64), b], axis=1), p)
    assert numpy.array_equal(R[:, -1], X[:, 0].astype(numpy.uint64))


def test_reference_error_behaviour():
    with pytest.raises(ValueError):
        rref.row_reduce(numpy.eye(3), pivoting=4)
    with pytest.raises(ZeroDivisionError):
        rref.rref_modp(numpy.array([[2, 1], [1, 1]]), 4)


def test_float_oracles_agree_and_solve():
    rs = numpy.random.RandomState(11)
    A = rs.random_sample((40, 40))
    b = rs.random_sample((40, 2))
    Ab = numpy.concatenate([A, b], axis=1)
    Rn = rref.rref_float(Ab)
    Rc, piv = crref.rref_f64(Ab)
    assert piv == list(range(40))
    assert numpy.allclose(Rn, Rc, rtol=0, atol=1e-12)
    assert numpy.allclose(Rc[:, 40:], numpy.linalg.solve(A, b), rtol=1e-10, atol=1e-12)
    C = rs.standard_normal((30, 31)) + 1j * rs.standard_normal((30, 31))
    Rz, pz = crref.rref_c128(C)
    Rzn = rref.rref_float(C)
    assert numpy.allclose(Rz, Rzn, atol=1e-12)
    assert numpy.allclose(Rz[:, 30], numpy.linalg.solve(C[:, :30], C[:, 30]), atol=1e-10)
