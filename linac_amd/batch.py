"""Batched GF(p) row reduction (no counterpart in the reference).

Many matrices -- any shapes, any primes 2 < p < 2^63 (uint32 storage up to
4278124287, uint64 storage from 2^32) -- advance together, one 256-column
panel per step, in shared kernel launches.  Matrices
can be added between steps and taken out as soon as they finish, which is
what the serving loop (``linac_amd.serving``) builds on.  Each result equals
``rref_with_pivots`` of that matrix alone.
"""
import ctypes

import numpy

from . import _lib
from .hip_row_reduce import to_field_array


WIDE_P_MIN = 2 ** 32  # primes from here on use 64-bit storage


def _wide(p):
    return int(p) >= WIDE_P_MIN


def footprint(rows, cols, p=2):
    """Device bytes one matrix occupies inside a batch (matrix + workspace);
    64-bit storage for p >= 2^32."""
    out = ctypes.c_uint64(0)
    lib = _lib.load()
    fn = lib.linac_batch_footprint_u64 if _wide(p) else lib.linac_batch_footprint_u32
    _lib.check(fn(rows, cols, ctypes.byref(out)))
    return out.value


class Batch:
    """A set of GF(p) matrices eliminated together.  Use from one thread."""

    def __init__(self, mem_cap_bytes=0):
        self._lib = _lib.load()
        h = ctypes.c_void_p()
        _lib.check(self._lib.linac_batch_create(ctypes.byref(h), int(mem_cap_bytes)))
        self._h = h
        self._shapes = {}

    def close(self):
        if self._h:
            self._lib.linac_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001 - interpreter shutdown
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def add(self, matrix, p):
        """Copy a matrix into the batch; returns its id.  MemoryError when the
        batch's memory cap would be exceeded (admission control)."""
        wide = _wide(p)
        a = to_field_array(matrix, int(p), numpy.dtype(numpy.uint64 if wide else numpy.uint32))
        rows, cols = a.shape
        rid = ctypes.c_int64(-1)
        fn = self._lib.linac_batch_add_u64 if wide else self._lib.linac_batch_add_u32
        _lib.check(fn(self._h, a.ctypes.data_as(ctypes.c_void_p), rows, cols, cols, int(p), 0, ctypes.byref(rid)))
        self._shapes[rid.value] = (rows, cols, False, wide)
        return rid.value

    def add_device(self, ptr, rows, cols, ld, p, in_place=False):
        """Add a matrix already in HBM (device address ``ptr``, row stride
        ``ld`` elements; uint32 entries, uint64 for p >= 2^32).  ``in_place``: reduce it where it lies (no
        copy); the buffer must outlive the matrix's stay in the batch and
        holds its RREF once taken."""
        rid = ctypes.c_int64(-1)
        flags = _lib.DEVICE_PTR | (_lib.IN_PLACE if in_place else 0)
        wide = _wide(p)
        fn = self._lib.linac_batch_add_u64 if wide else self._lib.linac_batch_add_u32
        _lib.check(fn(self._h, ctypes.c_void_p(ptr), rows, cols, ld, int(p), flags, ctypes.byref(rid)))
        self._shapes[rid.value] = (rows, cols, bool(in_place), wide)
        return rid.value

    def step(self):
        """Advance every unfinished matrix by one panel; returns how many remain."""
        n = ctypes.c_int64(0)
        _lib.check(self._lib.linac_batch_step(self._h, ctypes.byref(n)))
        return n.value

    def run(self):
        _lib.check(self._lib.linac_batch_run(self._h))

    def done(self, rid):
        rc = self._lib.linac_batch_status(self._h, int(rid))
        if rc < 0:
            _lib.check(rc)
        return rc == 1

    def memory(self):
        used, cap = ctypes.c_uint64(0), ctypes.c_uint64(0)
        _lib.check(self._lib.linac_batch_memory(self._h, ctypes.byref(used), ctypes.byref(cap)))
        return used.value, cap.value

    def take(self, rid):
        """(R, pivot_columns, rank) of a finished matrix; releases its memory.
        For a matrix added in place, R is None: the RREF is in the caller's
        device buffer."""
        if rid not in self._shapes:
            raise ValueError(f"unknown or already taken id {rid}")
        rows, cols, in_place, wide = self._shapes[rid]
        out = None if in_place else numpy.empty((rows, cols), dtype=numpy.uint64 if wide else numpy.uint32)
        piv = numpy.zeros(max(1, min(rows, cols)), dtype=numpy.int64)
        rank = ctypes.c_int64(0)
        fn = self._lib.linac_batch_take_u64 if wide else self._lib.linac_batch_take_u32
        _lib.check(fn(self._h, int(rid), None if out is None else out.ctypes.data_as(ctypes.c_void_p),
                      cols, 0, piv.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), ctypes.byref(rank)))
        del self._shapes[rid]
        return out, piv[:rank.value].tolist(), rank.value

    def take_many(self, rids):
        """take() for several finished matrices at once (two host round trips
        per storage width); returns a list in the order of ``rids``."""
        rids = [int(r) for r in rids]
        for rid in rids:
            if rid not in self._shapes:
                raise ValueError(f"unknown or already taken id {rid}")
        got = {}
        groups = {w: [r for r in rids if self._shapes[r][3] == w] for w in (False, True)}
        for wide, group in groups.items():
            if group:
                got.update(zip(group, self._take_group(group, wide)))
        return [got[r] for r in rids]

    def _take_group(self, rids, wide):
        n = len(rids)
        outs, piv = [], []
        for rid in rids:
            rows, cols, in_place, _ = self._shapes[rid]
            outs.append(None if in_place else numpy.empty((rows, cols), dtype=numpy.uint64 if wide else numpy.uint32))
            piv.append(numpy.zeros(max(1, min(rows, cols)), dtype=numpy.int64))
        ids = (ctypes.c_int64 * n)(*rids)
        out_p = (ctypes.c_void_p * n)(*[None if o is None else o.ctypes.data for o in outs])
        lds = (ctypes.c_int64 * n)(*[self._shapes[r][1] for r in rids])
        piv_p = (ctypes.c_void_p * n)(*[x.ctypes.data for x in piv])
        ranks = (ctypes.c_int64 * n)()
        fn = self._lib.linac_batch_take_many_u64 if wide else self._lib.linac_batch_take_many_u32
        _lib.check(fn(self._h, ids, n, out_p, lds, 0, piv_p, ranks))
        res = []
        for rid, o, x, r in zip(rids, outs, piv, ranks):
            del self._shapes[rid]
            res.append((o, x[:r].tolist(), int(r)))
        return res


def rref_batch(matrices, primes):
    """RREF of every matrix (each with its own prime) in one batch.

    Returns a list of (R, pivot_columns, rank).
    """
    if numpy.isscalar(primes):
        primes = [primes] * len(matrices)
    with Batch() as b:
        ids = [b.add(m, p) for m, p in zip(matrices, primes)]
        b.run()
        return b.take_many(ids)
