"""The autotuner measures on the device, sets and saves its choices, and the
knobs it moves leave GF(p) results bit-exact."""
import json

import numpy
import pytest

from oracle import crref, modp

pytestmark = pytest.mark.gpu


def test_quick_tune_sets_and_saves(gpu, tmp_path):
    from linac_amd import autotune
    before = autotune.current()
    try:
        data = autotune.tune(quick=True, path=str(tmp_path / "tune.json"))
        saved = json.load(open(data["file"]))
        assert saved["knobs"] == data["knobs"] == {k: autotune.get(k) for k in data["knobs"]}
        assert data["measurements"]["gf_crossover"] and data["measurements"]["float_crossover"]
        assert "gfx" in data["device"]
        for c in ("s", "m"):  # the GEMM shapes were chosen per size class
            assert data["knobs"][f"gemm_group.{c}"] in (4, 8, 16)
            assert data["knobs"][f"inner_split.{c}"] in (4, 8, 16)
            assert data["knobs"][f"tile64.{c}"] in (1, 2)
            assert data["measurements"]["by_class"][c]
    finally:
        autotune.apply(before)


@pytest.mark.parametrize("group, lookahead", [(1, 0), (4, 1), (16, 0), (64, 1)])
def test_knobs_keep_gf_results_exact(gpu, group, lookahead):
    from linac_amd import autotune
    before = autotune.current()
    p = 2 ** 31 - 1
    A = modp.low_rank(700, 900, 500, p, 21)
    try:
        autotune.set_knob("gemm_group", group)
        autotune.set_knob("lookahead", lookahead)
        R, piv, rank = gpu.rref_with_pivots(A, p, method="blocked")
    finally:
        autotune.apply(before)
    Ro, po = crref.rref_modp(A, p)
    assert piv == po and numpy.array_equal(R, Ro)


@pytest.mark.parametrize("n, rank, p", [(1200, 1000, 2 ** 31 - 1), (1200, None, 4278124237),
                                        (1100, None, 2 ** 61 - 1)])
def test_time_sliced_trailing_parts_stay_exact(gpu, n, rank, p):
    """With time slices the trailing update beyond the next panel runs in tile
    windows beside later inner steps; every window is applied once, in order."""
    from linac_amd import autotune
    before = autotune.current()
    rs = numpy.random.RandomState(23)
    if rank is not None:
        A = modp.low_rank(n, n + 64, rank, p, 23)
    else:  # uniform rows, full rank with overwhelming probability
        A = rs.randint(0, 2 ** 62, size=(n, n + 64), dtype=numpy.int64).astype(numpy.uint64) % numpy.uint64(p)
    try:
        autotune.set_knob("timeslice", 1)
        autotune.set_knob("lookahead", 1)
        R, piv, r = gpu.rref_with_pivots(A, p, method="blocked")
        big = rs.randint(0, p, size=(3000, 3000), dtype=numpy.int64)
        Rb = gpu.cuda_row_reduce(big, p, method="blocked")
    finally:
        autotune.apply(before)
    Ro, po = crref.rref_modp(A, p)
    assert piv == po and r == len(po) and numpy.array_equal(R, Ro)
    assert numpy.array_equal(Rb, numpy.identity(3000, dtype=Rb.dtype))


@pytest.mark.parametrize("n, rank, p", [(1500, 1300, 2 ** 31 - 1), (1300, None, 4278124237)])
def test_direct_to_lds_trailing_gemm_matches_register_staged(gpu, n, rank, p):
    """The trailing GEMM staged by direct-to-LDS loads (knob gemm_lds = 1) and
    the register-staged one (0) give the oracle's RREF, including the
    composite updates of a rank-deficient input, and the same standalone
    modular GEMM on 16-byte aligned rows."""
    import torch
    from linac_amd import autotune
    from linac_amd.incremental import gemm_sub
    before = autotune.current()
    rs = numpy.random.RandomState(31)
    if rank is not None:
        A = modp.low_rank(n, n + 200, rank, p, 31)
    else:
        A = rs.randint(0, 2 ** 62, size=(n, n + 200), dtype=numpy.int64).astype(numpy.uint64) % numpy.uint64(p)
    C = rs.randint(0, p, size=(700, 384), dtype=numpy.int64)
    L = rs.randint(0, p, size=(700, 450), dtype=numpy.int64)
    Bm = rs.randint(0, p, size=(450, 384), dtype=numpy.int64)
    dev = lambda a: torch.from_numpy(a.astype(numpy.uint32).view(numpy.int32)).cuda()
    out = {}
    try:
        for v in (0, 1):
            autotune.set_knob("gemm_lds", v)
            R, piv, r = gpu.rref_with_pivots(A, p, method="blocked")
            dC = dev(C)
            gemm_sub(dC, dev(L), dev(Bm), p, k=450)
            out[v] = (R, piv, r, dC.cpu().numpy().view(numpy.uint32).copy())
    finally:
        autotune.apply(before)
    Ro, po = crref.rref_modp(A, p)
    want = (C.astype(object) - (L.astype(object) @ Bm.astype(object)) % p) % p
    for v in (0, 1):
        R, piv, r, G = out[v]
        assert piv == po and r == len(po) and numpy.array_equal(R, Ro), v
        assert numpy.array_equal(G.astype(object), want), v


def test_size_class_knobs_keep_results_exact(gpu):
    """Per size class GEMM shapes (tile order, inner-GEMM split, the 64-bit
    64 x 64 tile) change no result: small matrices against the oracle, a
    medium-class 64-bit one against the same solve with the global knobs."""
    import torch
    from linac_amd import autotune
    before = autotune.current()
    p32, p64 = 2 ** 31 - 1, 2 ** 61 - 1
    A = modp.low_rank(1300, 1400, 1200, p32, 5)
    rs = numpy.random.RandomState(6)
    B = rs.randint(0, 2 ** 62, size=(1100, 1200), dtype=numpy.int64).astype(numpy.uint64) % numpy.uint64(p64)
    B[700:] = (B[:400] + B[300:700]) % numpy.uint64(p64)
    n = 3200
    g = torch.Generator(device="cuda")
    g.manual_seed(7)
    src = torch.randint(0, p64, (n, n + 64), generator=g, device="cuda", dtype=torch.int64)
    src[:, n:] = src[:, :64]  # non-pivot columns
    try:
        ref = src.clone()
        piv_ref, rank_ref = gpu.rref_device(ref.data_ptr(), n, n + 64, n + 64, p64, numpy.uint64, method="blocked")
        for c, v in (("s", 1), ("s", 2), ("m", 2)):
            autotune.set_knob(f"gemm_group.{c}", 1 if v == 1 else 16)
            autotune.set_knob(f"inner_split.{c}", 2 if v == 1 else 8)
            autotune.set_knob(f"tile64.{c}", v)
        Ra, pa, ra = gpu.rref_with_pivots(A, p32, method="blocked")
        Rb, pb, rb = gpu.rref_with_pivots(B, p64, method="blocked")
        work = src.clone()
        piv, rank = gpu.rref_device(work.data_ptr(), n, n + 64, n + 64, p64, numpy.uint64, method="blocked")
    finally:
        autotune.apply(before)
    Ro, po = crref.rref_modp(A, p32)
    assert pa == po and numpy.array_equal(Ra, Ro)
    Ro, po = crref.rref_modp(B, p64)
    assert pb == po and numpy.array_equal(Rb.astype(numpy.uint64), Ro.astype(numpy.uint64))
    assert piv == piv_ref and rank == rank_ref == n
    assert torch.equal(work, ref)


@pytest.mark.parametrize("threads,compact", [(256, 0), (512, 0), (512, 1), (256, 1)])
def test_inner_factor_workgroup_sizes_exact(gpu, threads, compact):
    """The 32-bit inner factorisation with 4 or 8 waves (knob inner_threads),
    in either form, gives the oracle's RREF: full-rank panels, rank-deficient
    ones (whose first free candidate row is often zero in the pivot column)
    and a matrix with zero columns."""
    from linac_amd import autotune
    p = 2 ** 31 - 1
    A = modp.low_rank(1300, 1400, 1200, p, 11)
    rs = numpy.random.RandomState(12)
    B = rs.randint(0, p, size=(1100, 1200)).astype(numpy.uint32)
    B[:, 100:140] = 0
    B[600:] = (B[:500].astype(numpy.uint64) * 3 % p).astype(numpy.uint32)
    try:
        autotune.set_knob("inner_threads", threads)
        autotune.set_knob("inner_compact", compact)
        out = [gpu.rref_with_pivots(M, p, method="blocked") for M in (A, B)]
    finally:
        autotune.set_knob("inner_threads", 0)
        autotune.set_knob("inner_compact", 0)
    for M, (R, piv, rank) in zip((A, B), out):
        Ro, po = crref.rref_modp(M, p)
        assert piv == po and rank == len(po) and numpy.array_equal(R, Ro)


@pytest.mark.parametrize("threads", [256, 1024])
def test_inner_factor_workgroup_sizes_exact_u64(gpu, threads):
    """The same for 64-bit storage (compact form) with 4 or 16 waves."""
    from linac_amd import autotune
    p = 2 ** 61 - 1
    rs = numpy.random.RandomState(13)
    B = rs.randint(0, 2 ** 62, size=(1100, 1200), dtype=numpy.int64).astype(numpy.uint64) % numpy.uint64(p)
    B[:, 200:230] = 0
    B[700:] = (B[:400] + B[300:700]) % numpy.uint64(p)
    try:
        autotune.set_knob("inner_threads", threads)
        R, piv, rank = gpu.rref_with_pivots(B, p, method="blocked")
    finally:
        autotune.set_knob("inner_threads", 0)
    Ro, po = crref.rref_modp(B, p)
    assert piv == po and rank == len(po) and numpy.array_equal(R.astype(numpy.uint64), Ro.astype(numpy.uint64))


def test_float_panel_rows_per_wave_same_result(gpu):
    """The float64 panel kernel with 16 or 32 rows per wave (knob
    fpanel_rows, per size class) computes the same RREF."""
    from linac_amd import autotune
    before = autotune.current()
    rs = numpy.random.RandomState(8)
    A = rs.rand(900, 901)
    out = []
    try:
        for v in (16, 32):
            autotune.set_knob("fpanel_rows.s", v)
            R, piv, rank = gpu.rref_with_pivots(A, method="blocked")
            out.append((R, piv, rank))
    finally:
        autotune.apply(before)
    (R1, p1, r1), (R2, p2, r2) = out
    assert p1 == p2 and r1 == r2 == 900
    assert numpy.allclose(R1, R2, rtol=0, atol=1e-12)
    x = numpy.linalg.solve(A[:, :900], A[:, 900])
    assert numpy.allclose(R2[:, 900], x, rtol=1e-9, atol=1e-9)
