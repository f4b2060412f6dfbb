"""COO files into HBM over GF(p) (linac_coo_scatter_u32/u64) and the device
RREF of what they hold, against host residues and the oracle (bit-exact)."""
import os
from fractions import Fraction

import numpy
import pytest

from conftest import GOLDEN
from linac_amd import sparse_matrix_tools as smt
from oracle import crref

pytestmark = pytest.mark.gpu

COO_JSON = os.path.join(GOLDEN, "test_coo.json")
COO_LARGE = os.path.join(GOLDEN, "test_coo_large_fractions.json")


def _host_residues(coo, p, shape):
    """Dense residues of a COO dictionary, each checked as value * den == num (mod p)."""
    M = numpy.zeros(shape, dtype=object)
    for (r, c), v in coo.items():
        f = Fraction(v)
        x = f.numerator % p * pow(f.denominator % p, p - 2, p) % p
        assert x * (f.denominator % p) % p == f.numerator % p
        M[r, c] = x
    return M


def _to_numpy(t, p):
    a = t.cpu().numpy()
    return a.view(numpy.uint32) if p <= 2 ** 32 - 1 else a.view(numpy.uint64)


@pytest.mark.parametrize("p", [2 ** 31 - 1, 2 ** 31 - 19, 2 ** 61 - 1])
@pytest.mark.parametrize("fmt", ["json", "txt"])
def test_coo_fixture_into_hbm_then_rref(gpu, tmp_path, p, fmt):
    import torch
    from linac_amd import rref_device
    coo = smt.coo_from_json_coo(COO_JSON)
    path = COO_JSON
    if fmt == "txt":
        path = str(tmp_path / "m.txt")
        smt.plain_txt_coo_from_coo(coo, path)
    assert smt.device_matrix_from_coo(path, p).shape == (86, 55)  # the fixture's own shape
    dev = smt.device_matrix_from_coo(path, p, shape=(86, 56))  # 16-byte rows for rref_device
    want = _host_residues(coo, p, dev.shape)
    got = _to_numpy(dev, p)
    assert numpy.array_equal(got.astype(object), want)
    # the device matrix goes straight into the device RREF
    rows, cols = dev.shape
    dt = numpy.uint32 if p <= 2 ** 32 - 1 else numpy.uint64
    piv, rank = rref_device(dev.data_ptr(), rows, cols, cols, p, dt)
    torch.cuda.synchronize()
    Ro, po = crref.rref_modp(numpy.array(want, dtype=numpy.uint64), p)
    assert piv == po and rank == len(po)
    assert numpy.array_equal(_to_numpy(dev, p).astype(numpy.uint64), Ro.astype(numpy.uint64))


def test_large_fraction_rref_stays_rref_mod_p(gpu):
    """The reference's large-fraction fixture is an RREF over Q; reduced modulo p
    it is still one, so the device RREF of it is itself."""
    import torch
    from linac_amd import rref_device
    p = 2 ** 31 - 1
    coo = smt.coo_from_json_coo(COO_LARGE)
    rows = max(r for r, c in coo) + 1
    cols = (max(c for r, c in coo) + 1 + 3) // 4 * 4  # 16-byte rows (extra zero columns change nothing)
    dev = smt.device_matrix_from_coo(COO_LARGE, p, shape=(rows, cols))
    before = _to_numpy(dev, p).copy()
    assert numpy.array_equal(before.astype(object), _host_residues(coo, p, (rows, cols)))
    piv, rank = rref_device(dev.data_ptr(), rows, cols, cols, p, numpy.uint32)
    torch.cuda.synchronize()
    assert rank == rows
    assert numpy.array_equal(_to_numpy(dev, p), before)


def test_coo_dict_shape_and_errors(gpu):
    p = 65537
    coo = {(0, 0): 3, (2, 5): Fraction(-1, 2), (1, 1): "7"}
    dev = smt.device_matrix_from_coo(coo, p, shape=(4, 8))
    got = _to_numpy(dev, p)
    assert got.shape == (4, 8) and got[0, 0] == 3 and got[1, 1] == 7
    assert got[2, 5] * 2 % p == p - 1 and numpy.count_nonzero(got) == 3
    with pytest.raises(ValueError):
        smt.device_matrix_from_coo(coo, p, shape=(2, 8))  # row 2 outside two rows
    with pytest.raises(ValueError):
        smt.device_matrix_from_coo({}, p)
    with pytest.raises(ValueError):
        smt.device_matrix_from_coo({(0, 0): "1/65537"}, p)  # denominator divisible by p


def test_coo_scatter_c_abi_rejects_bad_index(gpu):
    """The C ABI checks host index arrays before any GPU work."""
    import ctypes
    import torch
    from linac_amd import _lib
    out = torch.zeros((3, 4), dtype=torch.int32, device="cuda")
    ri = numpy.array([0, 3], dtype=numpy.int64)
    ci = numpy.array([0, 1], dtype=numpy.int64)
    v = numpy.array([1, 2], dtype=numpy.uint32)
    rc = _lib.load().linac_coo_scatter_u32(ctypes.c_void_p(out.data_ptr()), 4, 3, 4, ri.ctypes.data_as(ctypes.c_void_p),
                                           ci.ctypes.data_as(ctypes.c_void_p), v.ctypes.data_as(ctypes.c_void_p), 2, 0)
    assert rc == -1 and "outside" in _lib.last_error()
