"""linac_rref_u32_noalloc (no allocation: scratch in a static device array)
against linac_rref_u32 with the plain method, through the C ABI, on device
matrices: two calls in a row with the second smaller, pivot rows that are not
rows 0..rank-1 (the gather through the scratch), and the size limit."""
import ctypes

import numpy
import pytest

from oracle import modp

pytestmark = pytest.mark.gpu

P = 2 ** 31 - 1


def _dev(A, ld):
    import torch
    buf = torch.zeros((A.shape[0], ld), dtype=torch.int32, device="cuda")
    buf[:, :A.shape[1]] = torch.from_numpy(numpy.ascontiguousarray(A, dtype=numpy.uint32).view(numpy.int32)).cuda()
    torch.cuda.synchronize()
    return buf


def _run(fn, buf, rows, cols, ld, p, flags):
    import torch
    piv = numpy.zeros(max(1, min(rows, cols)), dtype=numpy.int64)
    rank = ctypes.c_int64(-1)
    rc = fn(ctypes.c_void_p(buf.data_ptr()), rows, cols, ld, p, flags, None,
            piv.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), ctypes.byref(rank))
    torch.cuda.synchronize()
    return rc, piv[:max(rank.value, 0)].tolist(), rank.value


def _compare(lib, A, p, ld):
    from linac_amd import _lib
    rows, cols = A.shape
    fn = lib.linac_rref_u32_noalloc
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, ctypes.c_uint32, ctypes.c_uint32,
                   ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)]
    b0, b1 = _dev(A, ld), _dev(A, ld)
    rc0, piv0, r0 = _run(lib.linac_rref_u32, b0, rows, cols, ld, p, _lib.DEVICE_PTR | _lib.METHOD_PLAIN)
    rc1, piv1, r1 = _run(fn, b1, rows, cols, ld, p, _lib.DEVICE_PTR | _lib.METHOD_PLAIN)
    assert rc0 == 0 and rc1 == 0, _lib.last_error()
    assert piv1 == piv0 and r1 == r0
    assert bool((b0 == b1).all().item()), "no-allocation result differs from linac_rref_u32"
    return piv0, r0


def test_noalloc_matches_linac_rref_u32(gpu):
    from linac_amd import _lib
    lib = _lib.load()
    assert lib.linac_init(0) == 0
    # 1: rank-deficient with its rows shuffled, so that pivot rows are not
    #    rows 0..rank-1 and the result is gathered through the scratch array
    A = modp.low_rank(900, 700, 500, P, 3)[numpy.random.RandomState(1).permutation(900)]
    piv, rank = _compare(lib, A, P, 704)
    assert rank == 500
    # 2: smaller, full rank (pivot rows 0..n-1: in place); the scratch keeps the
    #    previous call's contents, which must not leak into this one
    B = numpy.random.RandomState(2).randint(0, 65537, size=(300, 340)).astype(numpy.uint32)
    piv, rank = _compare(lib, B, 65537, 340)
    assert rank == 300 and piv == list(range(300))


def test_noalloc_refuses_beyond_its_limit(gpu):
    from linac_amd import _lib
    import torch
    lib = _lib.load()
    assert lib.linac_init(0) == 0
    fn = lib.linac_rref_u32_noalloc
    fn.restype = ctypes.c_int
    big = torch.zeros((4096, 4096), dtype=torch.int32, device="cuda")  # 64 MiB > the 32 MiB scratch
    rc = fn(ctypes.c_void_p(big.data_ptr()), ctypes.c_int64(4096), ctypes.c_int64(4096), ctypes.c_int64(4096),
            ctypes.c_uint32(P), ctypes.c_uint32(_lib.DEVICE_PTR), None, None, None)
    assert rc == _lib.ERR_INVALID and "no-allocation limit" in _lib.last_error()
    host = numpy.zeros((4, 4), dtype=numpy.uint32)
    rc = fn(ctypes.c_void_p(host.ctypes.data), ctypes.c_int64(4), ctypes.c_int64(4), ctypes.c_int64(4),
            ctypes.c_uint32(P), ctypes.c_uint32(0), None, None, None)
    assert rc == _lib.ERR_INVALID
