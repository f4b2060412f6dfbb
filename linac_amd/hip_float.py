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


def float_job_workspace(rows, cols, dtype=numpy.float64):
    """Device bytes a FloatJob of this shape needs for its own workspace."""
    n = ctypes.c_int64(0)
    cplx = 1 if numpy.dtype(dtype) == numpy.complex128 else 0
    _lib.check(_lib.load().linac_float_job_workspace(rows, cols, cplx, ctypes.byref(n)))
    return n.value


class FloatJob:
    """A blocked float RREF of a device matrix, advanced a few 64-column
    panels per ``step`` (``linac_float_job_*``), so that a long solve can be
    interleaved with other work.  ``workspace``: a device tensor of at least
    ``float_job_workspace`` bytes that the job uses (and the caller keeps
    alive) until it ends; without one the device context's workspace is used
    and only one such job may be open per device."""

    def __init__(self, ptr, rows, cols, ld, dtype, tol=TOLERANCE, workspace=None):
        self._lib = _lib.load()
        self.rows, self.cols = rows, cols
        h = ctypes.c_void_p()
        cplx = 1 if numpy.dtype(dtype) == numpy.complex128 else 0
        ws, nbytes = (None, 0) if workspace is None else \
            (ctypes.c_void_p(workspace.data_ptr()), workspace.numel() * workspace.element_size())
        _lib.check(self._lib.linac_float_job_create(ctypes.byref(h), ctypes.c_void_p(ptr), rows, cols, ld, cplx, tol,
                                                    _lib.DEVICE_PTR, ws, nbytes))
        self._h = h
        self._ws = workspace
        self.remaining = None

    def step(self, max_panels=1):
        """Run up to ``max_panels`` panels; returns how many remain."""
        rem = ctypes.c_int64(0)
        _lib.check(self._lib.linac_float_job_step(self._h, int(max_panels), ctypes.byref(rem)))
        self.remaining = rem.value
        return rem.value

    def finish(self):
        """(pivot_columns, rank); the RREF is in the device matrix."""
        piv = numpy.zeros(max(1, min(self.rows, self.cols)), dtype=numpy.int64)
        rank = ctypes.c_int64(0)
        h, self._h = self._h, None
        _lib.check(self._lib.linac_float_job_finish(h, piv.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                                    ctypes.byref(rank)))
        self._ws = None
        return piv[:rank.value].tolist(), rank.value

    def close(self):
        if self._h:
            self._lib.linac_float_job_destroy(self._h)
            self._h = None
        self._ws = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001 - interpreter shutdown
            pass
