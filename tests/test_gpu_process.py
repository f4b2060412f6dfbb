"""Process and thread model of the drop-in calls on the GPU: spawned workers
after the parent used the GPU (reference tests/test_cuda_concurrency.py:9-15),
several threads on one device context, and release/re-initialisation."""
import multiprocessing
import threading

import numpy
import pytest

from oracle import crref, modp

pytestmark = pytest.mark.gpu

P = 2 ** 31 - 1


def _spawned_rref(A):
    import linac_amd
    return linac_amd.cuda_row_reduce(A, P)


def test_spawn_workers_after_gpu_use(gpu):
    A = modp.low_rank(300, 320, 280, P, 5)
    R0 = gpu.cuda_row_reduce(A, P)  # the parent has a live context now
    ctx = multiprocessing.get_context("spawn")
    with ctx.Pool(2) as pool:
        outs = pool.map(_spawned_rref, [A] * 4)
    for R in outs:
        assert numpy.array_equal(R, R0)
    Ro, _ = crref.rref_modp(A, P)
    assert numpy.array_equal(R0, Ro)


@pytest.mark.parametrize("method", ["plain", "blocked"])
def test_threads_share_one_context(gpu, method):
    """Entry points are serialised per device: concurrent calls from several
    threads give exactly the results of the same calls made one by one."""
    mats = [modp.low_rank(260 + 40 * i, 300, 200 + 10 * i, P, 20 + i) for i in range(6)]
    expect = [gpu.rref_with_pivots(m, P, method=method) for m in mats]
    got = [None] * len(mats)
    errors = []

    def work(i):
        try:
            for _ in range(3):
                got[i] = gpu.rref_with_pivots(mats[i], P, method=method)
        except Exception as e:  # noqa: BLE001
            errors.append(e)

    th = [threading.Thread(target=work, args=(i,)) for i in range(len(mats))]
    for t in th:
        t.start()
    for t in th:
        t.join(120)
    assert not errors, errors
    for (Re, pe, re_), (Rg, pg, rg) in zip(expect, got):
        assert pe == pg and re_ == rg and numpy.array_equal(Re, Rg)


def test_shutdown_releases_and_reinitialises(gpu):
    from linac_amd import _lib
    A = modp.low_rank(1100, 1200, 1000, P, 3)
    R0, p0, r0 = gpu.rref_with_pivots(A, P, method="blocked")
    b = gpu.Batch()
    bid = b.add(A, P)
    assert _lib.load().linac_shutdown() == 0  # releases contexts, cached engines and the open batch
    with pytest.raises(ValueError, match="released"):
        b.step()
    b.close()  # a released handle is accepted by destroy and ignored
    R1, p1, r1 = gpu.rref_with_pivots(A, P, method="blocked")  # a fresh context
    assert p1 == p0 and r1 == r0 and numpy.array_equal(R1, R0)
    del bid
