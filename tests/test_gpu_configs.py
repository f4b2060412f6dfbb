"""Every GF(p) configuration of the benchmark at its stated size, each checked
by a certificate that shares nothing with the library's arithmetic (float64
limb products on rocBLAS, reduced mod p; see tests/certificates.py):

* planted systems [A | A x] with A uniform: the RREF is [I | x];
* rank 0.9 N as a seeded product L U: rank, pivots, the [I | X ; 0] shape and
  Freivalds' check U v == U[:, P] (R v) (mod p), which pins R completely
  because RREF(L U) = RREF(U) for L of full column rank;
* at N = 5000 (rank 4500) the whole output bit for bit against the C oracle
  (its pivots and RREF checksum, committed under tests/golden/);
* the 64-system loader sweep: rank, pivots and identity per system, and the
  loader's entries against the loader oracle on a sample.
Inputs are generated on the device from fixed seeds."""
import hashlib
import json
import os

import numpy
import pytest

import certificates as cert
from conftest import GOLDEN

pytestmark = pytest.mark.gpu

P31 = 2 ** 31 - 1
P61 = 2 ** 61 - 1


def _torch():
    import torch
    return torch


def _x(n, p, seed):
    torch = _torch()
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    return torch.randint(0, p, (n,), generator=g, device="cuda", dtype=torch.int64)


@pytest.mark.parametrize("n, method", [(10000, "blocked"), (20000, "blocked"), (10000, "plain"), (20000, "plain")])
def test_u32_planted_full_size(gpu, n, method):
    torch = _torch()
    ld = (n + 1 + 3) // 4 * 4  # 16-byte rows
    W = cert.uniform_u32(n, n, P31, seed=n, ld=ld)
    x = _x(n, P31, n + 1)
    W[:, n] = cert.matvec_mod(W, x, P31, cols=n).to(torch.int32)
    torch.cuda.synchronize()
    piv, rank = gpu.rref_device(W.data_ptr(), n, n + 1, ld, P31, numpy.uint32, method=method)
    assert rank == n and piv == list(range(n))
    assert torch.equal(cert.as_u64(W[:, n]), x), "solution column differs from the planted x"
    assert cert.check_identity_block(W, n)
    del W
    torch.cuda.empty_cache()


def _lu_input(n, r, p, seed):
    torch = _torch()
    L = cert.as_u64(cert.uniform_u32(n, r, p, seed=seed))
    U = cert.as_u64(cert.uniform_u32(r, n, p, seed=seed + 1))
    A = cert.matmul_mod(L, U, p).to(torch.int32)
    del L
    return A, U


def _check_low_rank(A, U, n, r, p, piv, rank, trials=2):
    torch = _torch()
    assert rank == r and piv == list(range(r))
    assert cert.check_identity_block(A[:r], r)
    assert int(A[r:].count_nonzero().item()) == 0
    for t in range(trials):
        v = _x(n, p, 1000 + t)
        lhs = cert.matvec_mod(U, v, p)
        w = cert.matvec_mod(A[:r], v, p)
        rhs = cert.matvec_mod(U, w, p, cols=r)
        assert torch.equal(lhs, rhs), "Freivalds: U v != U[:, P] (R v) (mod p)"


@pytest.mark.parametrize("method", ["blocked"])
def test_u32_rank_09n_lu_20000(gpu, method):
    torch = _torch()
    n, r = 20000, 18000
    A, U = _lu_input(n, r, P31, seed=7)
    torch.cuda.synchronize()
    piv, rank = gpu.rref_device(A.data_ptr(), n, n, n, P31, numpy.uint32, method=method)
    _check_low_rank(A, U, n, r, P31, piv, rank)
    del A, U
    torch.cuda.empty_cache()


@pytest.fixture(scope="module")
def lu5000():
    """The N = 5000, rank 4500 input (host-built, seeded) and the C oracle's
    pivots and RREF checksum for it (tests/golden/make_config_fixtures.py)."""
    from oracle import modp
    cfg = json.load(open(os.path.join(GOLDEN, "lu5000.json")))
    A = modp.low_rank(5000, 5000, 4500, P31, cfg["seed"]).astype(numpy.uint32)
    assert hashlib.sha256(A.tobytes()).hexdigest() == cfg["input_sha256"]
    return A, cfg


@pytest.mark.parametrize("method", ["blocked", "plain"])
def test_u32_5000_rank_4500_bit_exact_vs_oracle(gpu, lu5000, method):
    A, cfg = lu5000
    R, piv, rank = gpu.rref_with_pivots(A, P31, method=method)
    assert rank == cfg["rank"] == 4500 and piv == cfg["pivots"]
    assert hashlib.sha256(numpy.ascontiguousarray(R).tobytes()).hexdigest() == cfg["rref_sha256"]


@pytest.mark.parametrize("method", ["blocked", "plain"])
def test_u64_planted_10000(gpu, method):
    """uint64 storage, p = 2^61 - 1."""
    torch = _torch()
    n = 10000
    ld = n + 2  # 16-byte rows of uint64
    g = torch.Generator(device="cuda")
    g.manual_seed(61)
    W = torch.zeros((n, ld), dtype=torch.int64, device="cuda")
    for r0 in range(0, n, 2000):
        W[r0:r0 + 2000, :n] = torch.randint(0, P61, (min(2000, n - r0), n), generator=g, device="cuda",
                                            dtype=torch.int64)
    x = [int(v) for v in torch.randint(0, P61, (n,), generator=g, device="cuda", dtype=torch.int64).cpu()]
    b = cert.matvec_mod_u64(W, x, P61, cols=n)
    W[:, n] = torch.tensor(b, dtype=torch.int64, device="cuda")
    torch.cuda.synchronize()
    piv, rank = gpu.rref_device(W.data_ptr(), n, n + 1, ld, P61, numpy.uint64, method=method)
    assert rank == n and piv == list(range(n))
    assert W[:, n].cpu().tolist() == x
    assert cert.check_identity_block(W, n)
    del W
    torch.cuda.empty_cache()


def test_u32_100000_planted(gpu):
    """The ~40 GB configuration: [A | A x] with N = 100000, blocked."""
    torch = _torch()
    n = 100000
    ld = n + 4
    W = cert.uniform_u32(n, n, P31, seed=100, ld=ld)
    x = _x(n, P31, 101)
    W[:, n] = cert.matvec_mod(W, x, P31, cols=n, chunk_rows=2048).to(torch.int32)
    torch.cuda.synchronize()
    piv, rank = gpu.rref_device(W.data_ptr(), n, n + 1, ld, P31, numpy.uint32, method="blocked")
    assert rank == n and piv == list(range(n))
    assert torch.equal(cert.as_u64(W[:, n]), x)
    assert cert.check_identity_block(W, n)
    del W
    torch.cuda.empty_cache()


def test_sweep_64_systems_3000(gpu):
    """The batched configuration: 8 variables, 3000 monomials of degree <= 6
    times a prefactor, 3000 sample points per system, 64 systems built by the
    loader in HBM and reduced in place by one batch."""
    torch = _torch()
    from linac_amd import sweep
    from oracle import loader
    bases, lind, _ = sweep.sweep_problem(64, 8, 6, 3000, 3000, P31, 0)
    res = sweep.reduce_sweep(bases, lind, P31)
    assert res.ranks == [3000] * 64
    assert all(pv == list(range(3000)) for pv in res.pivots)
    m = res.matrices
    assert bool(m.diagonal(dim1=1, dim2=2).eq(1).all().item())
    assert int(m.count_nonzero().item()) == 64 * 3000
    del res, m
    torch.cuda.empty_cache()
    # the loader's entries (before the reduction) against the Python-int oracle
    rs = numpy.random.RandomState(5)
    for s in (0, 21, 63):
        A = gpu.cuda_load_matrix(bases[s], lind, field_characteristic=P31)
        assert A.shape == (3000, 3000) and A.dtype == numpy.uint32
        rows = rs.randint(0, 3000, size=40)
        ref = loader.load_reference(bases[s][rows], lind, P31)
        assert numpy.array_equal(A[rows], ref)
