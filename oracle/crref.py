"""ctypes binding of ``oracle/rref.c`` (test infrastructure only).

Builds ``liboracle.so`` with gcc on first use when it is missing.
"""
import ctypes
import os
import subprocess

import numpy

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "liboracle.so")
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE, "liboracle.so"], check=True)


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        lib = ctypes.CDLL(_SO)
        i64p = ctypes.POINTER(ctypes.c_int64)
        lib.oracle_rref_modp_u64.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_uint64, i64p, i64p]
        lib.oracle_rref_modp_u32.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_uint64, i64p, i64p]
        lib.oracle_rref_f64.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_double, i64p, i64p]
        lib.oracle_rref_c128.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_double, i64p, i64p]
        lib.oracle_load_matrix_modp_u64.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                                    ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, ctypes.c_uint64]
        lib.oracle_load_matrix_modp_u64.restype = None
        _lib = lib
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def rref_modp(matrix, p):
    """Returns (R, pivot column list); R is uint32 when p < 2**32 else uint64."""
    lib = _load()
    m = numpy.asarray(matrix)
    if m.dtype == object:
        m = m % p
    dt = numpy.uint32 if p < 2 ** 32 else numpy.uint64
    if m.dtype.kind == 'i':
        m = numpy.mod(m.astype(numpy.int64), p)
    a = numpy.ascontiguousarray(m).astype(dt).copy()
    rows, cols = a.shape
    piv = numpy.zeros(max(min(rows, cols), 1), dtype=numpy.int64)
    rank = ctypes.c_int64(0)
    fn = lib.oracle_rref_modp_u32 if dt == numpy.uint32 else lib.oracle_rref_modp_u64
    rc = fn(_ptr(a), rows, cols, int(p), piv.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), ctypes.byref(rank))
    if rc != 0:
        raise ZeroDivisionError(f"non-invertible head modulo {p}; is it prime?")
    return a, piv[:rank.value].tolist()


def rref_f64(matrix, threshold=1e-9):
    lib = _load()
    a = numpy.ascontiguousarray(matrix, dtype=numpy.float64).copy()
    rows, cols = a.shape
    piv = numpy.zeros(max(min(rows, cols), 1), dtype=numpy.int64)
    rank = ctypes.c_int64(0)
    lib.oracle_rref_f64(_ptr(a), rows, cols, threshold, piv.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), ctypes.byref(rank))
    return a, piv[:rank.value].tolist()


def rref_c128(matrix, threshold=1e-9):
    lib = _load()
    a = numpy.ascontiguousarray(matrix, dtype=numpy.complex128).copy()
    rows, cols = a.shape
    piv = numpy.zeros(max(min(rows, cols), 1), dtype=numpy.int64)
    rank = ctypes.c_int64(0)
    lib.oracle_rref_c128(_ptr(a), rows, cols, threshold, piv.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), ctypes.byref(rank))
    return a, piv[:rank.value].tolist()


def load_matrix_modp(bases, lindices, p):
    lib = _load()
    b = numpy.ascontiguousarray(bases, dtype=numpy.uint64)
    li = numpy.ascontiguousarray(lindices, dtype=numpy.uint32)
    rows, basis_len = b.shape
    cols, degree = li.shape
    out = numpy.zeros((rows, cols), dtype=numpy.uint64)
    lib.oracle_load_matrix_modp_u64(_ptr(out), _ptr(b), _ptr(li), rows, cols, basis_len, degree, int(p))
    return out
