"""Batched GF(p) elimination: every result equals the matrix solved alone and
the oracle, across uneven shapes, different primes and joins between steps."""
import numpy
import pytest

from oracle import crref, modp

pytestmark = pytest.mark.gpu

PRIMES = [2 ** 31 - 1, 2 ** 31 - 19, 2147483629, 2147483587, 2 ** 31 + 11, 65537, 3, 4278124237]  # last: largest limb prime


def _mat(rows, cols, rank, p, seed):
    if p < 2 ** 31:
        return modp.low_rank(rows, cols, rank, p, seed).astype(numpy.uint32)
    rs = numpy.random.RandomState(seed)
    L = rs.randint(0, p, size=(rows, rank), dtype=numpy.int64).astype(object)
    U = rs.randint(0, p, size=(rank, cols), dtype=numpy.int64).astype(object)
    return numpy.array((L @ U) % p, dtype=numpy.uint32)


def test_mixed_shapes_and_primes(gpu):
    rs = numpy.random.RandomState(0)
    mats, primes = [], []
    for i in range(14):
        rows, cols = int(rs.randint(1, 700)), int(rs.randint(1, 700))
        rank = int(rs.randint(0, min(rows, cols) + 1))
        p = PRIMES[i % len(PRIMES)]
        A = _mat(rows, cols, max(rank, 1), p, i)
        if i % 3 == 0 and cols > 5:
            A[:, 2] = 0
        mats.append(A)
        primes.append(p)
    results = gpu.rref_batch(mats, primes)
    for A, p, (R, piv, rank) in zip(mats, primes, results):
        Ro, po = crref.rref_modp(A, p)
        assert piv == po and rank == len(po)
        assert numpy.array_equal(R.astype(numpy.uint64), Ro.astype(numpy.uint64))


def test_hundreds_of_items_in_one_batch(gpu):
    """300 matrices in one batch: the full-rank gate and its tile prefix sums
    run 256 matrices at a time, with rank-deficient blocks dropped on both
    sides of the boundary; U is built several column pairs per workgroup."""
    p = 2 ** 31 - 1
    rs = numpy.random.RandomState(5)
    mats = []
    for i in range(300):
        n = int(rs.randint(260, 330))
        rank = n if i % 7 else int(rs.randint(100, n))  # every 7th rank-deficient
        mats.append(_mat(n, n + int(rs.randint(0, 20)), rank, p, 1000 + i))
    results = gpu.rref_batch(mats, [p] * len(mats))
    for A, (R, piv, rank) in zip(mats, results):
        Ro, po = crref.rref_modp(A, p)
        assert piv == po and rank == len(po)
        assert numpy.array_equal(R, Ro)


def test_same_result_alone_and_in_the_mix(gpu):
    p = 2 ** 31 - 1
    rs = numpy.random.RandomState(1)
    mats = [_mat(n, n + 7, n - 3, p, 10 + k) for k, n in enumerate([300, 1200, 64, 900, 1500])]
    alone = [gpu.rref_with_pivots(A, p, method="blocked") for A in mats]
    mixed = gpu.rref_batch(mats, p)
    for (Ra, pa, ra), (Rm, pm, rm) in zip(alone, mixed):
        assert pa == pm and ra == rm and numpy.array_equal(Ra, Rm)


def test_continuous_joins_and_early_exits(gpu):
    p = 2 ** 31 - 19
    big = _mat(1300, 1300, 1250, p, 3)
    small = [_mat(150, 170, 120, p, 4 + k) for k in range(4)]
    with gpu.Batch() as b:
        ib = b.add(big, p)
        b.step()
        b.step()
        ids = [b.add(m, p) for m in small]  # join while the big one is half way
        left = b.step()
        assert all(b.done(i) for i in ids)   # 170 columns: one panel
        taken = {i: b.take(i) for i in ids}  # leave before the big one finishes
        assert not b.done(ib) and left >= 1
        b.run()
        Rb, pb, rb = b.take(ib)
    for m, i in zip(small, ids):
        Ro, po = crref.rref_modp(m, p)
        assert taken[i][1] == po and numpy.array_equal(taken[i][0], Ro)
    Ro, po = crref.rref_modp(big, p)
    assert pb == po and numpy.array_equal(Rb, Ro)


def test_admission_cap_and_errors(gpu):
    from linac_amd.batch import footprint
    fp = footprint(500, 500)
    with gpu.Batch(mem_cap_bytes=2 * fp + 1) as b:
        b.add(numpy.eye(500, dtype=numpy.uint32), 7)
        b.add(numpy.eye(500, dtype=numpy.uint32), 7)
        with pytest.raises(MemoryError):
            b.add(numpy.eye(500, dtype=numpy.uint32), 7)
        used, cap = b.memory()
        assert used == 2 * fp and cap == 2 * fp + 1
        with pytest.raises(ValueError):
            b.take(0)  # not finished yet
        with pytest.raises(ValueError):
            b.done(99)
        b.run()
        R, piv, rank = b.take(0)
        assert rank == 500 and numpy.array_equal(R, numpy.eye(500, dtype=numpy.uint32))
        b.add(numpy.eye(500, dtype=numpy.uint32), 7)  # memory released by take
    with pytest.raises(ValueError):
        gpu.rref_batch([numpy.eye(3, dtype=numpy.uint32)], [2 ** 32 - 5])  # beyond the limb limit


def test_loader_sweep_batched(gpu):
    """Loader output for a sweep of evaluation points row-reduced as one batch."""
    p = 2 ** 31 - 1
    rs = numpy.random.RandomState(7)
    S, rows, bl = 6, 400, 9
    stack = rs.randint(0, p, size=(S, rows, bl)).astype(numpy.uint32)
    lind = rs.randint(0, bl, size=(380, 6)).astype(numpy.uint32)
    mats = gpu.load_matrices_stacked(stack, lind, p)
    res = gpu.rref_batch(list(mats), p)
    for A, (R, piv, rank) in zip(mats, res):
        Ro, po = crref.rref_modp(A, p)
        assert piv == po and numpy.array_equal(R, Ro)


def test_edge_shapes_in_one_batch(gpu):
    """1x1, all-zero, single row, single column and late-pivot matrices
    together with a regular one: each equals the oracle."""
    rs = numpy.random.RandomState(12)
    p = 2 ** 31 - 19
    late = numpy.zeros((150, 600), dtype=numpy.int64)
    late[:, 520:] = rs.randint(0, p, size=(150, 80))
    mats = [numpy.array([[3]]), numpy.zeros((90, 310), dtype=numpy.int64), rs.randint(0, p, size=(1, 513)),
            rs.randint(0, p, size=(513, 1)), late, rs.randint(0, p, size=(700, 650))]
    results = gpu.rref_batch([m.astype(numpy.uint32) for m in mats], p)
    for A, (R, piv, rank) in zip(mats, results):
        Ro, po = crref.rref_modp(A, p)
        assert piv == po and rank == len(po)
        assert numpy.array_equal(R.astype(numpy.uint64), Ro.astype(numpy.uint64))


def test_take_many_order_subsets_and_errors(gpu):
    """take_many returns results in the order asked, leaves the others in the
    batch, and refuses duplicates and unfinished matrices."""
    p = 2 ** 31 - 1
    rs = numpy.random.RandomState(13)
    mats = [rs.randint(0, p, size=(n, n + 3)).astype(numpy.uint32) for n in (40, 300, 90, 520)]
    with gpu.Batch() as b:
        ids = [b.add(m, p) for m in mats]
        b.run()
        with pytest.raises(ValueError):
            b.take_many([ids[0], ids[0]])
        got = b.take_many([ids[3], ids[1]])
        rest = b.take_many([ids[2], ids[0]])
        with pytest.raises(ValueError):
            b.take(ids[1])  # already taken
    for A, (R, piv, rank) in zip([mats[3], mats[1], mats[2], mats[0]], got + rest):
        Ro, po = crref.rref_modp(A, p)
        assert piv == po and rank == len(po) and numpy.array_equal(R, Ro)
    with gpu.Batch() as b:
        big = b.add(rs.randint(0, p, size=(600, 600)).astype(numpy.uint32), p)
        small = b.add(rs.randint(0, p, size=(30, 30)).astype(numpy.uint32), p)
        b.step()
        assert b.done(small) and not b.done(big)
        with pytest.raises(ValueError):
            b.take_many([small, big])  # big is not finished
        assert b.take_many([small])[0][2] == 30


PRIMES_64 = [2 ** 32 + 15, 2 ** 61 - 1, 2 ** 62 - 57, 2 ** 63 - 25]


def _mat64(rows, cols, rank, p, seed):
    rs = numpy.random.RandomState(seed)
    L = numpy.array([[int(x) for x in rs.randint(0, 2 ** 62, size=rank, dtype=numpy.int64)] for _ in range(rows)],
                    dtype=object) % p
    U = numpy.array([[int(x) for x in rs.randint(0, 2 ** 62, size=cols, dtype=numpy.int64)] for _ in range(rank)],
                    dtype=object) % p
    return numpy.array((L @ U) % p, dtype=numpy.uint64)


def test_mixed_32_and_64_bit_primes(gpu):
    """One batch holding uint32 matrices (p < 2^32) and uint64 matrices
    (2^32 <= p < 2^63), joined at different steps, host and in-place device
    inputs, taken together: every result equals the oracle bit for bit."""
    import torch
    rs = numpy.random.RandomState(21)
    cases = []
    for i in range(10):
        rows, cols = int(rs.randint(1, 420)), int(rs.randint(1, 420))
        rank = max(1, int(rs.randint(0, min(rows, cols) + 1)))
        if i % 2:
            p = PRIMES_64[(i // 2) % len(PRIMES_64)]
            A = _mat64(rows, cols, rank, p, 100 + i)
        else:
            p = PRIMES[(i // 2) % len(PRIMES)]
            A = _mat(rows, cols, rank, p, 100 + i)
        if i % 3 == 0 and cols > 5:
            A[:, 1] = 0
        cases.append((A, p))
    dev = {}
    with gpu.Batch() as b:
        ids = []
        for i, (A, p) in enumerate(cases):
            if i == 6:
                b.step()  # the rest join a running batch
            if i % 4 == 1:  # uint64 (int64 bits) in HBM, reduced in place
                t = torch.from_numpy(A.view(numpy.int64).copy()).cuda() if A.dtype == numpy.uint64 else \
                    torch.from_numpy(A.view(numpy.int32).copy()).cuda()
                torch.cuda.synchronize()
                ids.append(b.add_device(t.data_ptr(), A.shape[0], A.shape[1], A.shape[1], p, in_place=True))
                dev[ids[-1]] = t
            else:
                ids.append(b.add(A, p))
        b.run()
        results = b.take_many(ids)
    torch.cuda.synchronize()
    for rid, (A, p), (R, piv, rank) in zip(ids, cases, results):
        Ro, po = crref.rref_modp(A, p)
        if R is None:
            R = dev[rid].cpu().numpy().view(numpy.uint64 if p >= 2 ** 32 else numpy.uint32)
        assert piv == po and rank == len(po), (A.shape, p)
        assert numpy.array_equal(R.astype(numpy.uint64), Ro.astype(numpy.uint64)), (A.shape, p)


def test_batch_u64_rejects_small_primes_and_wrong_width(gpu):
    import ctypes
    from linac_amd import _lib
    lib = _lib.load()
    A = numpy.ones((4, 4), dtype=numpy.uint64)
    with gpu.Batch() as b:
        rid = ctypes.c_int64(-1)
        with pytest.raises(ValueError, match="2\\^32"):
            _lib.check(lib.linac_batch_add_u64(b._h, A.ctypes.data_as(ctypes.c_void_p), 4, 4, 4, 2 ** 31 - 1, 0,
                                               ctypes.byref(rid)))
        with pytest.raises(ValueError):
            _lib.check(lib.linac_batch_add_u64(b._h, A.ctypes.data_as(ctypes.c_void_p), 4, 4, 4, 2 ** 61 + 1, 0,
                                               ctypes.byref(rid)))  # not prime
        i64 = b.add(A, 2 ** 61 - 1)
        b.run()
        # a 64-bit item is not taken through the 32-bit call
        out = numpy.zeros((4, 4), dtype=numpy.uint32)
        with pytest.raises(ValueError, match="32-bit"):
            _lib.check(lib.linac_batch_take_u32(b._h, i64, out.ctypes.data_as(ctypes.c_void_p), 4, 0, None, None))
        R, piv, rank = b.take(i64)
        assert rank == 1 and piv == [0] and R[0].tolist() == [1, 1, 1, 1]
