"""Exact modular GEMM with index lists (linac_gemm_sub_u32) and the resident
incremental RREF built on it, against the oracle (bit-exact): after every
addition the resident RREF equals the oracle's RREF of all rows so far, and
the growing-matrix loop returns what the reference's loop
(linac/vector_spaces/over_function_fields.py:74-101, re-reducing the stacked
matrix each time) returns."""
import numpy
import pytest

from oracle import crref, modp

pytestmark = pytest.mark.gpu


def _dev(a):
    import torch
    return torch.from_numpy(numpy.ascontiguousarray(numpy.asarray(a, dtype=numpy.uint32)).view(numpy.int32)).cuda()


def _host(t):
    return t.cpu().numpy().view(numpy.uint32)


def _matmul_exact(A, B, p):
    return numpy.array((A.astype(object) @ B.astype(object)) % p, dtype=numpy.uint64)


@pytest.mark.parametrize("m, n, k, p", [(70, 130, 40, 2 ** 31 - 1), (300, 257, 700, 2 ** 31 - 1),
                                        (65, 64, 1100, 4278124237), (33, 90, 600, 3), (1, 1, 1, 65537)])
def test_gemm_sub_matches_exact_product(gpu, m, n, k, p):
    """Depth beyond 512 runs in chunks; index lists select columns of A and rows of B."""
    from linac_amd.incremental import gemm_sub
    rs = numpy.random.RandomState(m + n + k)
    C = rs.randint(0, p, size=(m, n), dtype=numpy.int64)
    A = rs.randint(0, p, size=(m, k + 5), dtype=numpy.int64)
    B = rs.randint(0, p, size=(k + 3, n), dtype=numpy.int64)
    a_cols = rs.permutation(k + 5)[:k]
    b_rows = rs.permutation(k + 3)[:k]
    dC = _dev(C)
    gemm_sub(dC, _dev(A), _dev(B), p, a_cols=a_cols, b_rows=b_rows)
    want = (C.astype(object) - _matmul_exact(A[:, a_cols], B[b_rows], p).astype(object)) % p
    assert numpy.array_equal(_host(dC).astype(object), want)
    # identity index lists (NULL) and C aliasing A: C <- C - C[:, :k'] B[:k']
    kk = min(n, k)
    dC2 = _dev(C)
    gemm_sub(dC2, dC2, _dev(B), p, k=kk)
    want2 = (C.astype(object) - _matmul_exact(C[:, :kk], B[:kk], p).astype(object)) % p
    assert numpy.array_equal(_host(dC2).astype(object), want2)


def test_gemm_sub_rejects_bad_input(gpu):
    from linac_amd.incremental import gemm_sub
    C, A, B = _dev(numpy.zeros((4, 8))), _dev(numpy.zeros((4, 6))), _dev(numpy.zeros((6, 8)))
    with pytest.raises(ValueError):
        gemm_sub(C, A, B, 2 ** 31 - 1, a_cols=[0, 6])  # column 6 outside A's six
    with pytest.raises(ValueError):
        gemm_sub(C, A, B, 2 ** 31 - 1, b_rows=[7])
    with pytest.raises(ValueError):
        gemm_sub(C, A, B, 2 ** 31 - 2, k=2)  # not prime


@pytest.mark.parametrize("cols, dim, chunks, p", [(300, 180, [25, 40, 1, 0, 60, 90, 30], 2 ** 31 - 1),
                                                  (129, 129, [64, 64, 10], 65537),
                                                  (264, 140, [60, 70, 40], 4278124237)])
def test_resident_rref_equals_rref_of_all_rows(gpu, cols, dim, chunks, p):
    from linac_amd.incremental import ResidentRREF
    rs = numpy.random.RandomState(cols + dim)
    V = rs.randint(0, p, size=(dim, cols), dtype=numpy.int64)
    V[:, 5] = 0  # a column without pivots
    inc = ResidentRREF(cols, p)
    rows = numpy.zeros((0, cols), dtype=numpy.int64)
    for c in chunks:
        new = modp.matmul_modp(rs.randint(0, p, size=(c, dim)), V, p) if p < 2 ** 31 else \
            numpy.array((rs.randint(0, p, size=(c, dim)).astype(object) @ V.astype(object)) % p, dtype=numpy.int64)
        inc.add_rows(new)
        rows = numpy.concatenate([rows, new.astype(numpy.int64)])
        Ro, po = crref.rref_modp(rows.astype(numpy.uint64), p)
        assert inc.pivots == po and inc.rank == len(po)
        assert numpy.array_equal(inc.rref().astype(numpy.uint64), Ro[:len(po)].astype(numpy.uint64))
    assert inc.rows_added == sum(chunks)


def test_growing_loop_matches_reference_loop(gpu):
    """The same pivots and RREF as re-reducing the stacked matrix after every
    addition (the reference's loop, restated with the oracle)."""
    from linac_amd.incremental import pivots_of_growing_matrix
    p, cols, dim = 2 ** 31 - 1, 260, 97
    rs = numpy.random.RandomState(5)
    V = rs.randint(0, p, size=(dim, cols), dtype=numpy.int64)
    C = rs.randint(0, p, size=(400, dim), dtype=numpy.int64)
    S = modp.matmul_modp(C, V, p)  # sample row i = C[i] V

    def sample_rows(lo, hi):
        return S[lo:hi]

    R, piv = pivots_of_growing_matrix(sample_rows, cols, p, iteration_start=20, iteration_step=20, max_iteration=50)
    # reference loop: A grows by 20 rows until the last row of its RREF is zero
    n = 20
    while True:
        Ro, po = crref.rref_modp(S[:n].astype(numpy.uint64), p)
        if not Ro[-1].any():
            break
        n += 20
    assert piv == po and numpy.array_equal(R.astype(numpy.uint64), Ro[:len(po)].astype(numpy.uint64))
    with pytest.raises(Exception, match="exceeded dimension 60"):
        pivots_of_growing_matrix(sample_rows, cols, p, iteration_start=20, iteration_step=20, max_iteration=2)


P_LIMB_MAX = 4278124237  # the largest prime the int8-limb GEMM accepts (<= 4278124287)


def _from_digits(d, p):
    """Canonical value in [0, p) whose centred representative has the
    balanced base-256 digits d (low digit first)."""
    c = sum(v * 256 ** i for i, v in enumerate(d))
    assert -(p // 2) <= c <= p // 2
    return c % p


@pytest.mark.parametrize("k", [512, 1100])
@pytest.mark.parametrize("pattern", ["same_sign", "mixed"])
def test_gemm_sub_extreme_digits(gpu, k, pattern):
    """Operands whose every balanced digit sits at the end of its range, so
    that every int32 diagonal sum S_s reaches its largest magnitude: +2^14 per
    digit product (both digits -128) or -16256 (digits -128 and 127), at the
    single-chunk depth 512 and the chunked depth 1100; against Python
    integers."""
    from linac_amd.incremental import gemm_sub
    p = P_LIMB_MAX
    lo = _from_digits([-128, -128, -128, -126], p)
    hi = _from_digits([127, 127, 127, 126], p)
    m, n = 64, 72
    rs = numpy.random.RandomState(k)
    if pattern == "same_sign":
        A = numpy.full((m, k), lo, dtype=numpy.int64)
        B = numpy.full((k, n), lo, dtype=numpy.int64)
        B[:, n // 2:] = hi  # half the columns at the negative extreme
    else:
        A = numpy.where(rs.randint(0, 2, size=(m, k)) == 1, lo, hi).astype(numpy.int64)
        B = numpy.where(rs.randint(0, 2, size=(k, n)) == 1, lo, hi).astype(numpy.int64)
        A[: m // 4] = lo  # rows and columns at one extreme throughout as well
        B[:, : n // 4] = lo
    C = numpy.where(rs.randint(0, 2, size=(m, n)) == 1, 0, p - 1).astype(numpy.int64)
    dC = _dev(C)
    gemm_sub(dC, _dev(A), _dev(B), p, k=k)
    want = (C.astype(object) - A.astype(object) @ B.astype(object)) % p
    assert numpy.array_equal(_host(dC).astype(object), want)
