"""The C ABI library: symbols, host-side validation, process model.  CPU only
(no compute call reaches a GPU here)."""
import ctypes
import multiprocessing
import os
import re

import numpy
import pytest

from conftest import ROOT

import linac_amd
from linac_amd import _lib
from linac_amd.hip_row_reduce import field_dtype, to_field_array


def header_symbols():
    text = open(os.path.join(ROOT, "include", "linac_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(linac_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    syms = header_symbols()
    assert "linac_rref_u32" in syms
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing
    # and the ctypes table covers them all
    assert set(syms) <= set(_lib.SIGNATURES)


def test_primality_matches_sympy():
    from sympy import isprime
    lib = _lib.load()
    for n in [0, 1, 2, 3, 4, 65537, 2 ** 31 - 1, 2 ** 31 - 19, 2 ** 31 + 11, 2 ** 32 - 5, 2 ** 61 - 1, 2 ** 61 + 15,
              2 ** 61 + 1, 3215031751, 2 ** 62 - 57, 1000000007 * 3]:
        assert lib.linac_is_prime(n) == int(isprime(n)), n


def test_invalid_arguments_raise_before_gpu_work():
    with pytest.raises(ValueError, match="not prime"):
        linac_amd.cuda_row_reduce(numpy.eye(4, dtype=numpy.uint32), 15)
    with pytest.raises(ValueError):
        linac_amd.cuda_row_reduce(numpy.eye(4), -3)
    lib = _lib.load()
    a = numpy.zeros((4, 4), dtype=numpy.uint32)
    rc = lib.linac_rref_u32(a.ctypes.data_as(ctypes.c_void_p), 4, 4, 3, 7, 0, None, None, None)
    assert rc == _lib.ERR_INVALID and "ld" in _lib.last_error()
    rc = lib.linac_rref_u64(a.ctypes.data_as(ctypes.c_void_p), 2, 2, 2, 2 ** 63 + 29, 0, None, None, None)
    assert rc == _lib.ERR_INVALID
    # empty input is valid and needs no device
    assert lib.linac_rref_u32(None, 0, 5, 5, 7, 0, None, None, None) == 0


def test_dtype_rules_follow_reference():
    """pycuda_row_reduce.py:41-48."""
    assert field_dtype(2 ** 31 - 1) == numpy.uint32
    assert field_dtype(2 ** 32 - 5) == numpy.uint32
    assert field_dtype(2 ** 32 + 15) == numpy.uint64
    assert field_dtype(7, _mod64=True) == numpy.uint64
    src = numpy.array([[-1, 5], [2 ** 40, 3]], dtype=numpy.int64)
    a = to_field_array(src, 7, numpy.dtype(numpy.uint32))
    assert a.dtype == numpy.uint32 and a.tolist() == [[6, 5], [2 ** 40 % 7, 3]]
    assert src[0, 0] == -1  # caller's array untouched
    o = numpy.array([[10 ** 30, 1]], dtype=object)
    assert to_field_array(o, 2 ** 61 - 1, numpy.dtype(numpy.uint64)).tolist() == [[10 ** 30 % (2 ** 61 - 1), 1]]


def _child_call(q):
    try:
        linac_amd.cuda_row_reduce(numpy.eye(3, dtype=numpy.uint32), 7)
        q.put("ok")
    except Exception as e:  # noqa: BLE001
        q.put(f"{type(e).__name__}: {e}")


def test_fork_after_runtime_init_fails_cleanly():
    """Reference tests/test_cuda_concurrency.py:18-27: a child forked after the
    runtime was touched cannot use it; the error says 'initialization error'."""
    _lib.device_count()  # touches the runtime in this process
    ctx = multiprocessing.get_context("fork")
    q = ctx.Queue()
    pr = ctx.Process(target=_child_call, args=(q,))
    pr.start()
    pr.join(60)
    msg = q.get(timeout=5)
    assert "initialization error" in msg.lower() and msg.startswith("RuntimeError")


def test_pivot_tools_match_reference_restatement():
    from oracle import crref, modp, pivots
    p = 2 ** 31 - 1
    A = modp.low_rank(20, 30, 9, p, 1)
    R, piv = crref.rref_modp(A, p)
    R2 = linac_amd.drop_bottom_zero_rows(R)
    assert R2.shape == pivots.drop_bottom_zero_rows(R).shape == (9, 30)
    assert linac_amd.pivot_columns_from_row_reduced_echelon_form(R2) == pivots.pivot_columns(R2) == piv
    F = numpy.zeros((3, 5))
    F[0, 1] = 1.0
    F[1, 3] = 1.0 + 1e-12
    assert linac_amd.pivot_columns_from_row_reduced_echelon_form(linac_amd.drop_bottom_zero_rows(F)) == [1, 3]


def test_loader_validation_before_gpu_work():
    bases = numpy.arange(12, dtype=numpy.uint32).reshape(3, 4)
    with pytest.raises(ValueError, match="lindices"):
        linac_amd.cuda_load_matrix(bases, numpy.array([[0, 4]], dtype=numpy.uint32), field_characteristic=7)
    with pytest.raises(ValueError, match="degree"):
        linac_amd.cuda_load_matrix(bases, numpy.zeros((2, 33), dtype=numpy.uint32), field_characteristic=7)
    with pytest.raises(ValueError, match="not prime"):
        linac_amd.cuda_load_matrix(bases, numpy.zeros((2, 2), dtype=numpy.uint32), field_characteristic=9)
    with pytest.raises(AssertionError):
        linac_amd.cuda_load_matrix(bases, numpy.zeros((2, 2), dtype=numpy.uint32), (3, 3), field_characteristic=7)
    with pytest.raises(NotImplementedError):
        linac_amd.load_matrices([], [], {'field': None}, use_cuda=False)


def test_loader_dtype_rules_follow_reference():
    """pycuda_matrix_loader.py:61-76 (note the 2**31-1 threshold)."""
    from linac_amd.hip_matrix_loader import loader_dtype
    assert loader_dtype(2 ** 31 - 1) == numpy.uint32
    assert loader_dtype(2 ** 31 + 11) == numpy.uint64
    assert loader_dtype(7, _mod64=True) == numpy.uint64
    assert loader_dtype(0, numpy.ones((2, 2))) == numpy.float64
    assert loader_dtype(0, numpy.ones((2, 2)) * 1j) == numpy.complex128


def test_tuning_knobs_set_get_and_validation():
    """Knob API (host only): defaults, round trip, range and key checks."""
    from linac_amd import autotune
    before = autotune.current()
    assert set(before) == set(autotune.KNOBS)
    try:
        autotune.set_knob("gemm_group", 4)
        assert autotune.get("gemm_group") == 4
        autotune.set_knob("lookahead", 0)
        assert autotune.get("lookahead") == 0
        with pytest.raises(ValueError):
            autotune.set_knob("gemm_group", 0)
        with pytest.raises(ValueError):
            autotune.set_knob("lookahead", 2)
        autotune.set_knob("timeslice", 1)
        assert autotune.get("timeslice") == 1
        with pytest.raises(ValueError):
            autotune.set_knob("timeslice", 2)
        autotune.set_knob("gemm_lds", 0)
        assert autotune.get("gemm_lds") == 0
        with pytest.raises(ValueError):
            autotune.set_knob("gemm_lds", 2)
        for v in (0, 256, 512, 1024):
            autotune.set_knob("inner_threads", v)
            assert autotune.get("inner_threads") == v
        for v in (128, 300, 2048):
            with pytest.raises(ValueError):
                autotune.set_knob("inner_threads", v)
        for v in (0, 1, 2, 4, 8):
            autotune.set_knob("u_passes", v)
            assert autotune.get("u_passes") == v
        for v in (3, 6, 16):
            with pytest.raises(ValueError):
                autotune.set_knob("u_passes", v)
        with pytest.raises(ValueError):
            autotune.set_knob("no_such_knob", 1)
        with pytest.raises(ValueError):
            autotune.get("no_such_knob")
    finally:
        autotune.apply(before)
        autotune.set_knob("inner_threads", 0)
        autotune.set_knob("u_passes", 0)
    assert autotune.current() == before


def test_tuning_file_round_trip(tmp_path):
    from linac_amd import autotune
    before = autotune.current()
    path = tmp_path / "t.json"
    try:
        autotune.save({"knobs": dict(before, gf_blocked_min_rows=777)}, str(path))
        data = autotune.load(str(path))
        assert data["knobs"]["gf_blocked_min_rows"] == 777 and autotune.get("gf_blocked_min_rows") == 777
        assert autotune.load(str(tmp_path / "missing.json")) is None
    finally:
        autotune.apply(before)


def test_tuning_file_without_float_crossovers(tmp_path):
    """The tuning the package applies by itself leaves the float crossovers
    (which can change float results in the last bits) alone."""
    from linac_amd import autotune
    before = autotune.current()
    path = tmp_path / "t.json"
    try:
        autotune.save({"knobs": dict(before, gf_blocked_min_rows=555,
                                     float_blocked_min_rows=before["float_blocked_min_rows"] + 7)}, str(path))
        autotune.load(str(path), float_crossovers=False)
        assert autotune.get("gf_blocked_min_rows") == 555
        assert autotune.get("float_blocked_min_rows") == before["float_blocked_min_rows"]
        autotune.load(str(path))
        assert autotune.get("float_blocked_min_rows") == before["float_blocked_min_rows"] + 7
    finally:
        autotune.apply(before)
