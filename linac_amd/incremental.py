"""Resident, incremental row reduction over GF(p) for callers that grow a
matrix and reduce it again after every addition.

The reference's callers re-run ``cuda_row_reduce`` on the whole stacked
matrix each time rows are added: ``VectorSpaceOfFunctions.__get_pivots__``
(``linac/vector_spaces/over_function_fields.py:74-101``) and the drop-and-retry
loop of ``iterative_gaussian_solver`` (``linac/linear_system_solver.py:24-79``).
Here the RREF R of the rows so far stays in HBM. New rows B are folded into it
with two exact modular GEMMs (``linac_gemm_sub_u32``, int8 limbs on the
matrix cores) and one RREF of the new rows alone:

1. E = B - B[:, P] R. This removes the resident pivot columns P from the new
   rows; E[:, P] = 0 because R[:, P] = I.
2. E is reduced on the device. Its pivots P_E are new columns, and its rows
   are zero in the columns P.
3. R <- R - R[:, P_E] E_r eliminates the new pivot columns from the old rows.
   The rows [R; E_r] then hold the identity in P + P_E, which makes them the
   RREF of all rows so far (the RREF over a field is unique).

The cost is that of the new rows, not of the whole stack.
"""
import ctypes

import numpy

from . import _lib
from .hip_row_reduce import rref_device

LIMB_P_MAX = 4278124287


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _idx(a):
    if a is None:
        return None, None
    arr = numpy.ascontiguousarray(numpy.asarray(a, dtype=numpy.int64))
    return arr, arr.ctypes.data_as(ctypes.c_void_p)


def gemm_sub(C, A, B, p, k=None, a_cols=None, b_rows=None):
    """C <- (C - A[:, a_cols] B[b_rows, :]) mod p on device tensors holding uint32
    bits (torch.int32, row-major, canonical entries); a_cols / b_rows default
    to 0..k-1.  C may be A or B itself."""
    import torch
    for t in (C, A, B):
        if t.dtype != torch.int32 or not t.is_cuda or t.dim() != 2 or t.stride(1) != 1:
            raise ValueError("gemm_sub takes 2-D row-major int32 device tensors")
    m, n = C.shape
    if A.shape[0] != m or B.shape[1] != n:
        raise ValueError(f"shapes do not chain: C {tuple(C.shape)}, A {tuple(A.shape)}, B {tuple(B.shape)}")
    k = int(k if k is not None else (len(a_cols) if a_cols is not None else len(b_rows)))
    ac, acp = _idx(a_cols)
    br, brp = _idx(b_rows)
    torch.cuda.current_stream().synchronize()  # the library works on its own stream
    _lib.check(_lib.load().linac_gemm_sub_u32(_ptr(C), C.stride(0), _ptr(A), A.stride(0), A.shape[1], acp,
                                              _ptr(B), B.stride(0), B.shape[0], brp, m, n, k, int(p),
                                              _lib.DEVICE_PTR))
    return C


class ResidentRREF:
    """RREF of a growing set of rows over GF(p), 2 < p <= 4278124287, kept in HBM.

    ``add_rows(rows)`` takes a (k, cols) array or a device int32 tensor of
    canonical residues (host arrays are reduced modulo p first) and returns
    the number of new pivots; ``rank``, ``pivots`` and ``rref()`` read the
    state.  Resident rows are kept in insertion order; ``rref()`` gathers them
    in pivot order."""

    def __init__(self, cols, field_characteristic):
        import torch
        p = int(field_characteristic)
        if not (2 < p <= LIMB_P_MAX):
            raise ValueError(f"ResidentRREF needs an odd prime p <= {LIMB_P_MAX}, got {p}")
        self.p, self.cols = p, int(cols)
        self.ld = (self.cols + 3) // 4 * 4  # 16-byte rows for the device RREF
        self._R = torch.zeros((min(256, max(1, self.cols)), self.ld), dtype=torch.int32, device="cuda")
        self._piv = []        # pivot column of each resident row (insertion order)
        self.rows_added = 0

    @property
    def rank(self):
        return len(self._piv)

    @property
    def pivots(self):
        return sorted(self._piv)

    def _device_rows(self, rows):
        import torch
        k = rows.shape[0]
        E = torch.zeros((k, self.ld), dtype=torch.int32, device="cuda")
        if isinstance(rows, torch.Tensor):
            if rows.shape[1] != self.cols:
                raise ValueError(f"rows have {rows.shape[1]} columns, expected {self.cols}")
            E[:, :self.cols] = rows.to(device="cuda", dtype=torch.int32)
        else:
            a = numpy.asarray(rows)
            if a.ndim != 2 or a.shape[1] != self.cols:
                raise ValueError(f"rows must be a (k, {self.cols}) array")
            a = numpy.mod(a.astype(object) if a.dtype == object else a.astype(numpy.int64), self.p)
            E[:, :self.cols] = torch.from_numpy(numpy.asarray(a, dtype=numpy.uint32).view(numpy.int32)).cuda()
        return E

    def add_rows(self, rows):
        import torch
        k = int(rows.shape[0])
        self.rows_added += k
        if k == 0:
            return 0
        E = self._device_rows(rows)
        r = self.rank
        if r:
            gemm_sub(E, E, self._R, self.p, k=r, a_cols=self._piv)  # E <- B - B[:, P] R
        torch.cuda.synchronize()
        # the blocked method stops after the panel where every new row has become a pivot row
        piv_e, rank_e = rref_device(E.data_ptr(), k, self.cols, self.ld, self.p, numpy.uint32, method="blocked")
        if rank_e == 0:
            return 0
        if r:
            gemm_sub(self._R[:r], self._R[:r], E, self.p, k=rank_e, a_cols=piv_e)  # R <- R - R[:, P_E] E_r
        if r + rank_e > self._R.shape[0]:
            grown = torch.zeros((min(self.cols, max(2 * self._R.shape[0], r + rank_e)), self.ld), dtype=torch.int32,
                                device="cuda")
            grown[:r] = self._R[:r]
            self._R = grown
        self._R[r:r + rank_e] = E[:rank_e]
        self._piv.extend(int(c) for c in piv_e)
        torch.cuda.synchronize()
        return rank_e

    def rref(self, device=False):
        """The rank x cols RREF (rows in pivot order): uint32 numpy array, or
        the int32 device tensor with ``device=True``."""
        import torch
        order = numpy.argsort(numpy.asarray(self._piv, dtype=numpy.int64), kind="stable")
        idx = torch.from_numpy(order).to("cuda")
        R = self._R[:self.rank].index_select(0, idx)[:, :self.cols].contiguous() if self.rank else \
            torch.zeros((0, self.cols), dtype=torch.int32, device="cuda")
        return R if device else R.cpu().numpy().view(numpy.uint32)


def pivots_of_growing_matrix(sample_rows, cols, field_characteristic, iteration_start=20, iteration_step=20,
                             max_iteration=50):
    """The row-growing loop of ``VectorSpaceOfFunctions.__get_pivots__``
    (``linac/vector_spaces/over_function_fields.py:74-101``) on a resident
    matrix. ``sample_rows(start, stop)`` returns the rows of sample points
    start..stop-1.

    Rows are added ``iteration_step`` at a time while every row so far is
    independent. That is the reference's test that the last row of the full
    RREF is non-zero. Returns ``(rref, pivots)``, the RREF without its zero
    rows, like the reference's ``self.rref`` and ``self.pivots``, and raises
    the reference's exception past ``max_iteration`` additions."""
    inc = ResidentRREF(cols, field_characteristic)
    inc.add_rows(sample_rows(0, iteration_start))
    total, iteration = iteration_start, 0
    while inc.rank == total:
        if iteration >= max_iteration:
            raise Exception(f"Vector space exceeded dimension {iteration_start + iteration_step * max_iteration}")
        lo = iteration_start + iteration * iteration_step
        inc.add_rows(sample_rows(lo, lo + iteration_step))
        total += iteration_step
        iteration += 1
    return inc.rref(), inc.pivots
