"""float64 / complex128 RREF through the library (``field_characteristic == 0``).

Semantics of ``cuda_row_reduce`` for real and complex matrices
(``linac/pycuda_row_reduce.py:49-56,87-131``, ``linac/row_reduce.cu:55``):
one-shot row scaling by the largest |x|, partial pivoting on |x|, heads
accepted iff |x| > 1e-9.
"""
import ctypes

import numpy

from . import _lib

TOLERANCE = 1e-9


def rref_float_host(a, method="auto", tol=TOLERANCE, refine=False):
    """In-place RREF of the C-contiguous float64 / complex128 host array ``a``.

    ``refine``: residual-based refinement of the non-pivot columns on the
    blocked method (``_lib.refine_last()`` then holds its statistics).
    Returns (pivot_columns, rank).
    """
    lib = _lib.load()
    rows, cols = a.shape
    piv = numpy.zeros(max(1, min(rows, cols)), dtype=numpy.int64)
    rank = ctypes.c_int64(0)
    fn = lib.linac_rref_f64 if a.dtype == numpy.float64 else lib.linac_rref_c128
    flags = _lib.METHODS[method] | (_lib.REFINE if refine else 0)
    rc = fn(a.ctypes.data_as(ctypes.c_void_p), rows, cols, cols, tol, flags, None,
            piv.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), ctypes.byref(rank))
    _lib.check(rc)
    return piv[:rank.value].tolist(), rank.value
