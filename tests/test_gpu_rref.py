"""GF(p) RREF on the GPU against the golden fixtures, the oracle and
size-independent certificates.  Bit-exact throughout."""
import hashlib
import json
import os

import numpy
import pytest

from conftest import GOLDEN, load_small_cases
from oracle import crref, modp

pytestmark = pytest.mark.gpu

METHODS = ["plain", "blocked"]
LIMB_P_MAX = 4278124287  # largest prime of the 32-bit int8-limb GEMM (blocked method)
P63 = 9223372036854775783  # largest prime below 2**63 (64-bit blocked method: 2**32 <= p < 2**63)
P32_UP = 4294967311  # smallest prime above 2**32


def blocked_applies(p, mod64=None):
    """Whether the blocked method serves p with the storage the dtype rules pick
    (uint64 for p > 2**32 - 1, or when forced)."""
    wide = p > 2 ** 32 - 1 if mod64 is None else mod64
    return (2 ** 32 <= p < 2 ** 63) if wide else (2 < p <= LIMB_P_MAX)


def _rref(gpu, A, p, method, **kw):
    try:
        return gpu.cuda_row_reduce(A, p, method=method, **kw)
    except ValueError as e:
        if "does not apply" in str(e):
            pytest.skip(f"{method} not applicable: {e}")
        raise


@pytest.mark.parametrize("method", METHODS)
def test_reference_golden_vector(gpu, method):
    """reference tests/test_row_reduce_random_cuda.py:52-63."""
    p = 2 ** 31 - 19
    numpy.random.seed(0)
    A = numpy.random.randint(p, size=(4097, 4097), dtype=numpy.uint32)[:-1, :]
    R = _rref(gpu, A, p, method)
    known = numpy.load(os.path.join(GOLDEN, "known_result_4096_4097_2to32m19.npy"), allow_pickle=False)
    assert R.dtype == numpy.uint32 and R.shape == (4096, 4097)
    assert numpy.array_equal(R[:, -1], known)
    assert numpy.array_equal(R[:, :-1], numpy.eye(4096, dtype=numpy.uint32))


@pytest.mark.parametrize("case", range(len(load_small_cases())))
def test_sympy_fixtures(gpu, case):
    p, A, R, piv = load_small_cases()[case]
    for method in ["plain", "auto"] + (["blocked"] if blocked_applies(p) else []):
        out, pv, rank = gpu.rref_with_pivots(A, p, method=method)
        assert out.dtype == (numpy.uint64 if p > 2 ** 32 - 1 else numpy.uint32)
        assert numpy.array_equal(out.astype(numpy.uint64), R), method
        assert pv == piv and rank == len(piv)


@pytest.mark.parametrize("p, n", [(2 ** 31 - 1, 1024), (2 ** 31 - 1, 2048), (2 ** 31 - 1, 3072), (2 ** 31 + 11, 3072),
                                  (2 ** 61 - 1, 2048), (P63, 1100)])
@pytest.mark.parametrize("method", METHODS)
def test_full_rank_gives_identity(gpu, p, n, method):
    """reference tests/test_row_reduce_random_cuda.py:23-49 (seeded here)."""
    A = numpy.random.RandomState(n).randint(0, p, size=(n, n), dtype=numpy.int64)
    R = _rref(gpu, A, p, method)
    assert numpy.array_equal(R, numpy.identity(n, dtype=R.dtype))


@pytest.mark.parametrize("shape, rank, p, seed", [
    ((300, 300), 250, 2 ** 31 - 1, 0), ((257, 513), 200, 2 ** 31 - 19, 1), ((600, 170), 170, 65537, 2),
    ((129, 1000), 129, 2 ** 31 + 11, 3), ((400, 400), 17, 2 ** 61 - 1, 4), ((333, 444), 300, 2, 5),
    ((200, 300), 150, 3, 6), ((1000, 1000), 700, 2 ** 31 - 1, 7), ((64, 4096), 64, 2 ** 32 - 5, 8),
    ((300, 700), 200, 4278124237, 9), ((520, 520), 520, 4278124237, 10),  # largest limb prime: sums wrap 32 bits
    ((300, 420), 60, P63, 11), ((150, 330), 40, P32_UP, 12),  # ends of the 64-bit limb range
])
@pytest.mark.parametrize("method", ["plain", "auto", "blocked"])
def test_matches_oracle_bit_exact(gpu, shape, rank, p, seed, method):
    rows, cols = shape
    if method == "blocked" and not blocked_applies(p):
        with pytest.raises(ValueError):
            gpu.rref_with_pivots(numpy.eye(4, dtype=numpy.uint64 if p > 2 ** 32 - 1 else numpy.uint32), p,
                                 method=method)
        return
    if p < 2 ** 31:
        A = modp.low_rank(rows, cols, rank, p, seed)
    else:
        rs = numpy.random.RandomState(seed)
        L = rs.randint(0, p, size=(rows, rank), dtype=numpy.int64).astype(object)
        U = rs.randint(0, p, size=(rank, cols), dtype=numpy.int64).astype(object)
        A = numpy.array((L @ U) % p, dtype=numpy.uint64)
    A[:, 3] = 0
    Rg, pg, rk = gpu.rref_with_pivots(A, p, method=method)
    Ro, po = crref.rref_modp(A, p)
    assert pg == po and rk == len(po) <= min(rank, rows)
    assert numpy.array_equal(Rg.astype(numpy.uint64), Ro.astype(numpy.uint64))


def test_c0_config_checksum(gpu):
    cfg = json.load(open(os.path.join(GOLDEN, "c0_config.json")))
    p = cfg["p"]
    rs = numpy.random.RandomState(0)
    L = rs.randint(p, size=(500, 400), dtype=numpy.uint32)
    U = rs.randint(p, size=(400, 500), dtype=numpy.uint32)
    A = modp.matmul_modp(L, U, p).astype(numpy.uint32)
    for method in ["plain", "auto", "blocked"]:
        R, piv, rank = gpu.rref_with_pivots(A, p, method=method)
        assert piv == cfg["pivots"] and rank == 400
        assert hashlib.sha256(numpy.ascontiguousarray(R).tobytes()).hexdigest() == cfg["rref_sha256"]


@pytest.mark.parametrize("method", METHODS)
def test_planted_solution_large(gpu, method):
    p = 2 ** 31 - 1
    Ab, X = modp.planted(5000, p, 12, extra=3)
    R, piv, rank = gpu.rref_with_pivots(Ab.astype(numpy.uint32), p, method=method) if method != "blocked" else \
        (_rref(gpu, Ab.astype(numpy.uint32), p, method), None, None)
    assert numpy.array_equal(R[:, 5000:].astype(numpy.int64), X)
    assert numpy.array_equal(R[:, :5000], numpy.eye(5000, dtype=numpy.uint32))


@pytest.mark.parametrize("rows, cols, rank, dup, seed", [(8400, 520, 300, False, 41), (9000, 600, 600, True, 42)])
def test_blocked_many_row_stripes(gpu, rows, cols, rank, dup, seed):
    """Over 128 stripes of 64 rows: each inner GEMM workgroup owns a whole
    stripe and captures its own left operand (fewer stripes share them)."""
    p = 2 ** 31 - 1
    rs = numpy.random.RandomState(seed)
    if dup:  # every row twice, shuffled: half of each panel's candidates reduce to zero
        half = modp.low_rank(rows // 2, cols, min(rank, rows // 2), p, seed)
        A = numpy.concatenate([half, half], axis=0)[rs.permutation(rows)]
    else:
        A = modp.low_rank(rows, cols, rank, p, seed)
    Rg, pg, rk = gpu.rref_with_pivots(A, p, method="blocked")
    Ro, po = crref.rref_modp(A, p)
    assert pg == po and rk == len(po)
    assert numpy.array_equal(Rg, Ro)


def test_blocked_planted_many_stripes(gpu):
    p = 2 ** 31 - 1
    Ab, X = modp.planted(9000, p, 7, extra=2)
    R = _rref(gpu, Ab.astype(numpy.uint32), p, "blocked")
    assert numpy.array_equal(R[:, 9000:].astype(numpy.int64), X)
    assert numpy.array_equal(R[:, :9000], numpy.eye(9000, dtype=numpy.uint32))


def test_u64_planted(gpu):
    p = 2 ** 61 - 1
    n = 700
    rs = numpy.random.RandomState(9)
    A = rs.randint(0, p, size=(n, n), dtype=numpy.int64).astype(object)
    X = rs.randint(0, p, size=(n, 1), dtype=numpy.int64).astype(object)
    Ab = numpy.concatenate([A, (A @ X) % p], axis=1)
    R = gpu.cuda_row_reduce(Ab, p)
    assert R.dtype == numpy.uint64
    assert R[:, -1].tolist() == X[:, 0].tolist()


def _u64_structured(rows, cols, p, seed):
    """Uniform rows over GF(p), then every fifth row replaced by the sum of two
    earlier rows, three zero columns and two duplicated columns: a
    rank-deficient input with irregular pivots, built without big-integer
    products."""
    rs = numpy.random.RandomState(seed)
    hi = numpy.uint64(p)
    A = (rs.randint(0, 2 ** 62, size=(rows, cols), dtype=numpy.int64).astype(numpy.uint64) % hi)
    for r in range(10, rows, 5):
        i, j = rs.randint(0, r, size=2)
        A[r] = (A[i] + A[j]) % hi  # both < p < 2**63: the sum fits 64 bits
    A[:, [3, cols // 2, cols - 2]] = 0
    A[:, 7] = A[:, 1]
    A[:, cols - 1] = A[:, cols // 3]
    return A


@pytest.mark.parametrize("p", [2 ** 61 - 1, P63, P32_UP])
@pytest.mark.parametrize("shape", [(700, 900), (1100, 530)])
def test_u64_blocked_matches_oracle(gpu, p, shape):
    """64-bit storage through the blocked method (eight int8 limbs, integer
    fold) against the C oracle, bit-exact, over several 256-column panels."""
    A = _u64_structured(shape[0], shape[1], p, seed=shape[0] + p % 1000)
    Rg, pg, rk = gpu.rref_with_pivots(A, p, method="blocked")
    Ro, po = crref.rref_modp(A, p)
    assert pg == po and rk == len(po) < min(shape)
    assert Rg.dtype == numpy.uint64 and numpy.array_equal(Rg, Ro.astype(numpy.uint64))


@pytest.mark.parametrize("p", [2 ** 61 - 1, P63])
def test_u64_blocked_planted(gpu, p):
    """[A | A x] over GF(p) with N = 2600 (eleven panels): the solution column
    comes back exactly, and the blocked and plain kernels agree bit for bit."""
    n = 2600
    rs = numpy.random.RandomState(31)
    A = rs.randint(0, 2 ** 62, size=(n, n), dtype=numpy.int64).astype(numpy.uint64) % numpy.uint64(p)
    x = [int(v) for v in rs.randint(0, 2 ** 62, size=n, dtype=numpy.int64) % p]
    b = numpy.array([sum(int(a) * xi for a, xi in zip(row.tolist(), x)) % p for row in A], dtype=numpy.uint64)
    Ab = numpy.concatenate([A, b[:, None]], axis=1)
    Rb, pb, rb = gpu.rref_with_pivots(Ab, p, method="blocked")
    assert rb == n and pb == list(range(n))
    assert [int(v) for v in Rb[:, -1]] == x
    assert numpy.array_equal(Rb[:, :n], numpy.eye(n, dtype=numpy.uint64))
    Rp, pp, rp = gpu.rref_with_pivots(Ab, p, method="plain")
    assert numpy.array_equal(Rp, Rb)


@pytest.mark.parametrize("n", [5000, 9000])
def test_u64_blocked_matches_plain_many_stripes(gpu, n):
    """64-bit storage with many 64-row stripes (6 and 3 inner-GEMM workgroups
    per stripe), rank-deficient rows and columns: blocked and plain agree bit
    for bit (the plain kernel is pinned against the oracle above)."""
    p = 2 ** 61 - 1
    A = _u64_structured(n, n // 2, p, seed=n)
    Rb, pb, rb = gpu.rref_with_pivots(A, p, method="blocked")
    Rp, pp, rp = gpu.rref_with_pivots(A, p, method="plain")
    assert pb == pp and rb == rp < n // 2
    assert numpy.array_equal(Rb, Rp)


def test_u64_blocked_thresholds(gpu):
    """64-bit storage of a prime below 2**32 (forced by the caller) is outside
    the 64-bit limb range: "blocked" refuses it, "auto" takes the plain kernel."""
    p = 2 ** 31 - 1
    A = numpy.random.RandomState(3).randint(0, p, size=(40, 50)).astype(numpy.uint64)
    with pytest.raises(ValueError):
        gpu.rref_with_pivots(A, p, _mod64=True, method="blocked")
    R, piv, rank = gpu.rref_with_pivots(A, p, _mod64=True, method="auto")
    Ro, po = crref.rref_modp(A, p)
    assert R.dtype == numpy.uint64 and piv == po and numpy.array_equal(R, Ro.astype(numpy.uint64))


def test_edge_cases(gpu):
    p = 2 ** 31 - 1
    assert gpu.cuda_row_reduce(numpy.zeros((0, 5), dtype=numpy.uint32), p).shape == (0, 5)
    assert gpu.cuda_row_reduce(numpy.zeros((4, 0), dtype=numpy.uint32), p).shape == (4, 0)
    assert gpu.cuda_row_reduce(numpy.array([[5]]), p).tolist() == [[1]]
    assert gpu.cuda_row_reduce(numpy.array([[0]]), p).tolist() == [[0]]
    Z = gpu.cuda_row_reduce(numpy.zeros((37, 53), dtype=numpy.uint32), p)
    assert not Z.any()
    # non-canonical and negative inputs are reduced mod p first
    A = numpy.array([[p + 2, -1], [3, 2 * p + 4]], dtype=numpy.int64)
    assert numpy.array_equal(gpu.cuda_row_reduce(A, p), crref.rref_modp(numpy.mod(A, p), p)[0])
    # caller's array is not modified; dtype follows the reference
    B = numpy.random.RandomState(1).randint(0, p, size=(50, 60))
    B0 = B.copy()
    R = gpu.cuda_row_reduce(B, p)
    assert numpy.array_equal(B, B0) and R.dtype == numpy.uint32
    assert gpu.cuda_row_reduce(B, p, _mod64=True).dtype == numpy.uint64
    assert numpy.array_equal(gpu.cuda_row_reduce(B, p, _mod64=True), R.astype(numpy.uint64))


def test_repeatable(gpu):
    p = 2 ** 31 - 1
    A = modp.low_rank(700, 900, 600, p, 21)
    R1 = gpu.cuda_row_reduce(A, p)
    R2 = gpu.cuda_row_reduce(A, p)
    assert numpy.array_equal(R1, R2)


@pytest.mark.parametrize("shape, structure, seed", [
    ((700, 900), "duplicate_rows", 0), ((1100, 600), "zero_blocks", 1), ((530, 530), "rank_one_panel", 2),
    ((1500, 1300), "low_rank", 3), ((257, 2049), "wide", 4), ((2049, 255), "tall", 5),
])
def test_blocked_structured_inputs(gpu, shape, structure, seed):
    """Rank-deficient panels that force repeated inner steps."""
    p = 2 ** 31 - 1
    rows, cols = shape
    rs = numpy.random.RandomState(seed)
    if structure == "duplicate_rows":
        half = modp.low_rank(rows // 2, cols, rows // 2, p, seed)
        A = numpy.concatenate([half, half], axis=0)[rs.permutation(rows - rows % 2)]
    elif structure == "zero_blocks":
        A = modp.low_rank(rows, cols, 500, p, seed)
        A[: rows // 2, : cols // 3] = 0
        A[:, 300:340] = 0
    elif structure == "rank_one_panel":
        A = modp.low_rank(rows, cols, 400, p, seed)
        v = rs.randint(0, p, size=(1, 64))
        A[:, 100:164] = modp.matmul_modp(rs.randint(0, p, size=(rows, 1)), v, p)
    elif structure == "low_rank":
        A = modp.low_rank(rows, cols, 37, p, seed)
    else:
        A = modp.low_rank(rows, cols, min(rows, cols), p, seed)
    Rg, pg, rk = gpu.rref_with_pivots(A, p, method="blocked")
    Ro, po = crref.rref_modp(A, p)
    assert pg == po and rk == len(po)
    assert numpy.array_equal(Rg, Ro)


@pytest.mark.parametrize("case", ["1x1", "1x1zero", "zeros", "row", "column", "late_pivots", "zero_panel",
                                  "one_tile", "few_rows", "identity_perm"])
def test_blocked_edge_shapes(gpu, case):
    """Shapes and structures at the edges of the panel / tile logic, blocked
    method against the oracle (bit-exact)."""
    p = 2 ** 31 - 1
    rs = numpy.random.RandomState(11)
    if case == "1x1":
        A = numpy.array([[5]])
    elif case == "1x1zero":
        A = numpy.array([[0]])
    elif case == "zeros":
        A = numpy.zeros((300, 530), dtype=numpy.int64)
    elif case == "row":
        A = rs.randint(0, p, size=(1, 700))
        A[0, :3] = 0
    elif case == "column":
        A = rs.randint(0, p, size=(700, 1))
    elif case == "late_pivots":  # the first 600 columns are zero: pivots only in the third panel
        A = numpy.zeros((200, 700), dtype=numpy.int64)
        A[:, 600:] = rs.randint(0, p, size=(200, 100))
    elif case == "zero_panel":  # a whole 256-column panel without pivots between two full ones
        A = rs.randint(0, p, size=(400, 800))
        A[:, 256:512] = 0
    elif case == "one_tile":
        A = rs.randint(0, p, size=(64, 64))
    elif case == "few_rows":
        A = rs.randint(0, p, size=(5, 900))
    else:  # a permuted identity: pivot rows far from their columns
        A = numpy.identity(333, dtype=numpy.int64)[rs.permutation(333)] * 7
    A = numpy.asarray(A, dtype=numpy.int64)
    Rg, pg, rk = gpu.rref_with_pivots(A, p, method="blocked")
    Ro, po = crref.rref_modp(A, p)
    assert pg == po and rk == len(po)
    assert numpy.array_equal(Rg.astype(numpy.uint64), Ro.astype(numpy.uint64))
