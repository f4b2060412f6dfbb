"""Drop-in ``cuda_row_reduce`` on the gfx950 library.

Same signature, defaults, dtype rules and return value as
``linac/pycuda_row_reduce.py:18-151``:

* ``field_characteristic > 0``: uint64 storage when ``_mod64`` or (``_mod64`` is
  None and p > 2**32 - 1), else uint32 (``:41-48``);
* ``field_characteristic == 0``: float64 when ``_real`` or every imaginary part
  is 0, else complex128 (``:49-56``);
* the caller's array is never modified; a new C-contiguous array of that dtype
  holding the RREF (non-zero rows first, zero rows at the bottom) is returned
  (``:44,47,52,55,151``).

Differences that only affect inputs the reference mishandles: integer and
object inputs are reduced modulo p before the cast (the reference's
``astype`` wraps negative or >= 2**32 values), a composite p raises
``ValueError`` (the reference returns a wrong matrix), and the pivot search
covers every row (the reference's covers a 2048-row window).

``rref_with_pivots`` additionally returns the pivot columns and the rank that
the library computes on the device, so callers need not re-scan R for them.
"""
import ctypes
import time

import numpy

from . import _lib
from .timeit_decorator import timeit


def field_dtype(field_characteristic, _mod64=None):
    """numpy dtype the reference uses for a prime (``pycuda_row_reduce.py:41-48``)."""
    if _mod64 or (_mod64 is None and field_characteristic > 2 ** 32 - 1):
        return numpy.dtype(numpy.uint64)
    return numpy.dtype(numpy.uint32)


def to_field_array(matrix, p, dtype):
    """A new C-contiguous array of ``dtype`` holding ``matrix`` mod p (canonical or not)."""
    m = numpy.asarray(matrix)
    if m.ndim != 2:
        raise ValueError(f"expected a 2-D matrix, got shape {m.shape}")
    if m.dtype == object:
        m = numpy.vectorize(lambda x: int(x) % p, otypes=[object])(m) if m.size else m
        return numpy.ascontiguousarray(m.astype(dtype))
    if m.dtype.kind == 'i':
        return numpy.ascontiguousarray(numpy.mod(m.astype(numpy.int64), p).astype(dtype))
    if m.dtype.kind == 'u':
        if m.dtype.itemsize > dtype.itemsize:
            m = numpy.mod(m, p)
        return numpy.array(m, dtype=dtype, order='C', copy=True)
    if m.dtype.kind == 'b':
        return numpy.array(m, dtype=dtype, order='C', copy=True)
    # floats: the reference's astype semantics
    return numpy.array(m, dtype=dtype, order='C', copy=True)


def _rref_modp_host(a, p, method, want_pivots):
    """In-place RREF of the C-contiguous uint32/uint64 host array ``a``."""
    lib = _lib.load()
    rows, cols = a.shape
    piv = numpy.zeros(max(1, min(rows, cols)), dtype=numpy.int64)
    rank = ctypes.c_int64(0)
    flags = _lib.METHODS[method]
    fn = lib.linac_rref_u32 if a.dtype == numpy.uint32 else lib.linac_rref_u64
    if a.dtype == numpy.uint32 and p > 2 ** 32 - 1:
        raise ValueError(f"p = {p} does not fit uint32 storage; pass _mod64=True")
    rc = fn(a.ctypes.data_as(ctypes.c_void_p), rows, cols, cols, p, flags, None,
            piv.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), ctypes.byref(rank))
    _lib.check(rc)
    return piv[:rank.value].tolist(), rank.value


def _prepare(matrix, field_characteristic, _real, _mod64):
    p = int(field_characteristic)
    if p < 0:
        raise ValueError("field_characteristic must be >= 0")
    if p > 0:
        dtype = field_dtype(p, _mod64)
        return p, to_field_array(matrix, p, dtype)
    m = numpy.asarray(matrix)
    if _real or (_real is None and (m.dtype.kind != 'c' and m.dtype != object or
                                    (m.size and numpy.max(numpy.abs(numpy.imag(m.astype(complex)))) == 0))):
        return 0, numpy.array(numpy.real(m) if m.dtype.kind == 'c' else m, dtype=numpy.float64, order='C', copy=True)
    return 0, numpy.array(m, dtype=numpy.complex128, order='C', copy=True)


@timeit
def cuda_row_reduce(matrix, field_characteristic=0, verbose=False, _real=None, _mod64=None, method="auto",
                    refine=False):
    r"""
    Args:
        matrix (2D numpy.ndarray):
            The matrix to be cast to row-reduced echelon form.
        field_characteristic (int, optional, default: 0):
            The characteristic p of the field, or 0 for real and complex matrices.
        verbose (bool, optional, default: False):
            If True, prints additional information.
        method ("auto" | "plain" | "blocked" | "mixed"):
            Elimination method on the GPU (extension; default picks).
            "mixed" (float64 only): FP32 trailing updates plus refinement.
        refine (bool):
            Real and complex matrices only: residual-based refinement of the
            non-pivot columns (extension; blocked method).
    """
    t0 = time.time()
    p, a = _prepare(matrix, field_characteristic, _real, _mod64)
    t1 = time.time()
    if p > 0:
        if refine:
            raise ValueError("refine applies to real and complex matrices")
        _rref_modp_host(a, p, method, False)
    else:
        from .hip_float import rref_float_host
        rref_float_host(a, method, refine=refine)
    if verbose:
        print(f"\rconversion {t1 - t0:.3f} s, device call {time.time() - t1:.3f} s", end="\n")
    return a


def rref_with_pivots(matrix, field_characteristic=0, _real=None, _mod64=None, method="auto", refine=False):
    """(R, pivot_columns, rank): ``cuda_row_reduce`` plus the device's pivot list."""
    p, a = _prepare(matrix, field_characteristic, _real, _mod64)
    if p > 0:
        if refine:
            raise ValueError("refine applies to real and complex matrices")
        piv, rank = _rref_modp_host(a, p, method, True)
    else:
        from .hip_float import rref_float_host
        piv, rank = rref_float_host(a, method, refine=refine)
    return a, piv, rank


def rref_device(ptr, rows, cols, ld, field_characteristic, dtype, method="auto", stream=None, finalize=True,
                profile=False, refine=False):
    """In-place RREF of a matrix already in HBM (``ptr`` = device address).

    ``dtype``: numpy.uint32 / numpy.uint64 (field_characteristic = p, rows
    16-byte aligned) or numpy.float64 / numpy.complex128 (field_characteristic
    0, pivot tolerance 1e-9).  Returns (pivot_columns, rank).
    """
    lib = _lib.load()
    p = int(field_characteristic)
    dt = numpy.dtype(dtype)
    piv = numpy.zeros(max(1, min(rows, cols)), dtype=numpy.int64)
    rank = ctypes.c_int64(0)
    flags = _lib.DEVICE_PTR | _lib.METHODS[method] | (0 if finalize else _lib.NO_FINALIZE) | \
        (_lib.PROFILE if profile else 0) | (_lib.REFINE if refine else 0)
    st = ctypes.c_void_p(stream) if stream else None
    out = (piv.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), ctypes.byref(rank))
    if dt.kind in 'fc':
        if p != 0:
            raise ValueError("float matrices need field_characteristic 0")
        fn = lib.linac_rref_f64 if dt == numpy.float64 else lib.linac_rref_c128
        rc = fn(ctypes.c_void_p(ptr), rows, cols, ld, 1e-9, flags, st, *out)
    else:
        fn = lib.linac_rref_u32 if dt == numpy.uint32 else lib.linac_rref_u64
        rc = fn(ctypes.c_void_p(ptr), rows, cols, ld, p, flags, st, *out)
    _lib.check(rc)
    return piv[:rank.value].tolist(), rank.value
