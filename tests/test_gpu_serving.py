"""Continuous-batching serving loop: every request's result is the same alone
and in the mix (bit-identical for GF(p), float64 too: a float request runs
as a job of its own, a few panels between batch steps), admission holds the memory cap, and requests that
arrive while others run join the batch."""
import numpy
import pytest

pytestmark = pytest.mark.gpu


def _alone(req, seed):
    from linac_amd import serving
    import torch
    serving.materialise(req, seed)
    if req.p == 0:
        A = req.tensor.cpu().numpy()
    else:
        A = req.tensor.cpu().numpy().view(numpy.uint32)
    req.tensor = None
    from linac_amd import rref_with_pivots
    # GF(p): any method gives the unique RREF; float: the serving loop's
    # float jobs run the blocked method
    return rref_with_pivots(A, req.p, method="blocked")


def test_results_same_alone_and_in_the_mix(gpu):
    from linac_amd import serving
    reqs = serving.make_requests(40, rate=400.0, seed=5, n_min=20, n_max=700, float_share=0.2)
    assert any(r.p == 0 for r in reqs) and any(r.p for r in reqs)
    results, wall, stats = serving.serve(reqs, seed=5, keep=True)
    assert [r.rid for r in results] == list(range(40))
    for req, res in zip(sorted(reqs, key=lambda r: r.rid), results):
        R, piv, rank = _alone(serving.Request(req.rid, req.rows, req.cols, req.p), 5)
        got = res.tensor.cpu().numpy()
        if req.p:
            got = got.view(numpy.uint32)
            assert numpy.array_equal(got, R), req
        else:
            assert numpy.array_equal(got, R), (req, float(numpy.max(numpy.abs(got - R))))
        assert res.pivots == piv and res.rank == rank
        assert res.finished >= res.arrival


def test_admission_holds_the_cap_and_queues(gpu):
    """A cap that fits about two of the requests: the loop still finishes
    every request, never exceeds the cap, and later arrivals wait."""
    from linac_amd import serving
    reqs = serving.make_requests(12, rate=0, seed=7, n_min=300, n_max=400, float_share=0.25)
    one = max(serving.Server(1)._bytes(r) for r in reqs)
    cap = int(2.2 * one)
    results, wall, stats = serving.serve(reqs, mem_cap_bytes=cap, seed=7)
    assert len(results) == 12 and stats["max_in_use_bytes"] <= cap
    assert stats["max_in_use_bytes"] >= one  # it did run requests side by side or back to back
    for r in results:
        assert r.rank > 0


def test_request_larger_than_the_cap_is_rejected(gpu):
    """A request whose footprint alone exceeds the cap can never be admitted:
    the server reports it instead of spinning."""
    from linac_amd import serving
    srv = serving.Server(mem_cap_bytes=1000)
    req = serving.materialise(serving.Request(0, 64, 64, serving.SERVING_PRIMES[0]))
    with pytest.raises(MemoryError):
        srv.submit(req)
    srv.close()


def test_served_results_equal_the_oracle(gpu):
    """Requests served in the mix against the C oracle (linac's row_reduce
    restated): GF(p) bit for bit with the same pivots, float64 within 1e-9
    with the same pivots."""
    from linac_amd import serving
    from oracle import crref
    reqs = serving.make_requests(24, rate=0, seed=11, n_min=40, n_max=300, float_share=0.25)
    results, wall, stats = serving.serve(reqs, seed=11, keep=True)
    assert any(r.p == 0 for r in reqs) and any(r.p for r in reqs)
    for req, res in zip(sorted(reqs, key=lambda r: r.rid), results):
        A = serving.materialise(serving.Request(req.rid, req.rows, req.cols, req.p), 11).tensor.cpu().numpy()
        got = res.tensor.cpu().numpy()
        if req.p:
            Ro, po = crref.rref_modp(A.view(numpy.uint32), req.p)
            assert numpy.array_equal(got.view(numpy.uint32), Ro), req
        else:
            Ro, po = crref.rref_f64(A)
            assert float(numpy.max(numpy.abs(got - Ro))) <= 1e-9, req
        assert res.pivots == po and res.rank == len(po), req


def test_serving_with_64_bit_primes(gpu):
    """Requests over primes that need 64-bit storage are served in the same
    batch as 32-bit ones; results equal the oracle bit for bit."""
    from linac_amd import serving
    from oracle import crref
    primes = serving.SERVING_PRIMES[:3] + serving.SERVING_PRIMES_64
    reqs = serving.make_requests(16, rate=0, seed=13, n_min=30, n_max=260, float_share=0.0, primes=primes)
    assert any(r.p >= 2 ** 32 for r in reqs) and any(r.p < 2 ** 32 for r in reqs)
    results, wall, stats = serving.serve(reqs, seed=13, keep=True)
    for req, res in zip(sorted(reqs, key=lambda r: r.rid), results):
        A = serving.materialise(serving.Request(req.rid, req.rows, req.cols, req.p), 13).tensor.cpu().numpy()
        wide = req.p >= 2 ** 32
        Ro, po = crref.rref_modp(A.view(numpy.uint64 if wide else numpy.uint32), req.p)
        got = res.tensor.cpu().numpy().view(numpy.uint64 if wide else numpy.uint32)
        assert numpy.array_equal(got, Ro), req
        assert res.pivots == po and res.rank == len(po), req
