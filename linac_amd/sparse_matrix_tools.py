"""Coordinate-list (COO) matrices on disk, and COO files loaded into HBM
(``linac/sparse_matrix_tools.py:13-99``).

The host functions keep the reference's names, file formats and results:
* JSON: ``{"(row, column)": "value"}``;
* plain text: one ``row column value`` line per entry;
* values as strings (Fractions by default);
* a COO dictionary that lists the non-zero entries in row-major order, plus
  the bottom-right entry whenever that is needed to keep the matrix's shape.

They run on whole arrays, and keys are parsed rather than evaluated.

``device_matrix_from_coo`` is the route to the GPU path. It reduces a COO
file or dictionary modulo p on the host (integers directly, fractions through
the inverse of the denominator), then scatters the entries into a zeroed
row-major device matrix (``linac_coo_scatter_u32/u64``). The result feeds
``rref_device`` without a dense host copy.
"""
import ctypes
import json
import re
from fractions import Fraction

import numpy

from . import _lib

_KEY = re.compile(r"\(\s*(-?\d+)\s*,\s*(-?\d+)\s*\)")


def sparsity(matrix):
    """String representing the % of zeros in matrix."""
    matrix = numpy.atleast_2d(matrix)
    return "%.3f" % ((1 - numpy.count_nonzero(matrix) / (matrix.shape[0] * matrix.shape[1])) * 100) + "%"


def coo_from_matrix(matrix):
    """COO (coordinate list) dictionary {(row, column): value} of the non-zero
    entries, row-major, plus the bottom-right entry when the non-zeros alone do
    not span the shape."""
    matrix = numpy.asarray(matrix)
    rows, columns = matrix.shape
    ri, ci = numpy.nonzero(matrix != 0)
    coo = dict(zip(zip(ri.tolist(), ci.tolist()), list(matrix[ri, ci])))
    if len(coo) == 0 or ri.max() + 1 != rows or ci.max() + 1 != columns:
        coo[(rows - 1, columns - 1)] = matrix[rows - 1, columns - 1]
    return coo


def _coo_arrays(coo):
    keys = numpy.array(list(coo.keys()), dtype=numpy.int64).reshape(-1, 2)
    return keys[:, 0], keys[:, 1]


def matrix_from_coo(coo):
    """Object array of shape (max row + 1, max column + 1); absent entries are 0."""
    ri, ci = _coo_arrays(coo)
    matrix = numpy.zeros((int(ri.max()) + 1, int(ci.max()) + 1), dtype=object)
    vals = numpy.empty(len(coo), dtype=object)
    vals[:] = list(coo.values())
    matrix[ri, ci] = vals
    return matrix


def _parse_key(key):
    m = _KEY.fullmatch(key.strip())
    if m is None:
        raise ValueError(f"not a COO key: {key!r}")
    return int(m.group(1)), int(m.group(2))


def _parse_keys(keys):
    """Row and column arrays of '(row, column)' keys, parsed as whole arrays
    when every key has that exact form, else key by key."""
    if not keys:
        return numpy.zeros(0, dtype=numpy.int64), numpy.zeros(0, dtype=numpy.int64)
    a = numpy.char.strip(numpy.asarray(keys))
    if numpy.all(numpy.char.startswith(a, "(")) and numpy.all(numpy.char.endswith(a, ")")):
        body = numpy.char.partition(numpy.char.strip(a, "()"), ",")
        r, c = _int_strings(numpy.char.strip(body[:, 0])), _int_strings(numpy.char.strip(body[:, 2]))
        if r is not None and c is not None and numpy.all(body[:, 1] == ","):
            return r, c
    pairs = [_parse_key(k) for k in keys]
    return (numpy.array([k[0] for k in pairs], dtype=numpy.int64), numpy.array([k[1] for k in pairs], dtype=numpy.int64))


def coo_from_json_coo(file_name, dtype=Fraction):
    """COO dictionary from a JSON file {"(row, column)": "value"}."""
    with open(file_name, 'r') as f:
        data = json.load(f)
    return {_parse_key(k): dtype(v) for k, v in data.items()}


def matrix_from_json_coo(file_name, dtype=Fraction):
    return matrix_from_coo(coo_from_json_coo(file_name, dtype=dtype))


def json_coo_from_coo(coo, file_name):
    with open(file_name, 'w') as f:
        json.dump({str(key): str(val) for key, val in coo.items()}, f)


def json_coo_from_matrix(matrix, file_name):
    json_coo_from_coo(coo_from_matrix(matrix), file_name)


def plain_txt_coo_from_coo(coo, file_name):
    with open(file_name, "w+") as f:
        f.write("".join(f"{key[0]} {key[1]} {val}\n" for key, val in coo.items()))


def plain_txt_coo_from_matrix(matrix, file_name):
    plain_txt_coo_from_coo(coo_from_matrix(matrix), file_name)


def _read_plain_txt(file_name):
    """(row strings, column strings, value strings) of a plain-text COO file,
    tokenised as a whole; lines with other than three fields are an error."""
    with open(file_name, 'r') as f:
        text = f.read()
    tok = text.split()
    lines = [line for line in text.splitlines() if line.strip()]
    if len(tok) != 3 * len(lines):
        for line in lines:
            if len(line.split()) != 3:
                raise ValueError(f"a COO line needs 'row column value', got {line!r}")
    return tok[0::3], tok[1::3], tok[2::3]


def coo_from_plain_txt_coo(file_name, dtype=Fraction):
    """COO dictionary from lines 'row column value'."""
    r, c, v = _read_plain_txt(file_name)
    return {(int(a), int(b)): dtype(x) for a, b, x in zip(r, c, v)}


def matrix_from_plain_txt_coo(file_name, dtype=Fraction):
    return matrix_from_coo(coo_from_plain_txt_coo(file_name, dtype=dtype))


# ---------------------------------------------------------------------------
# COO into HBM over GF(p)
# ---------------------------------------------------------------------------

def _modpow_vec(b, e, p):
    """b^e mod p elementwise for int64 arrays with p^2 < 2^63."""
    r = numpy.ones_like(b)
    b = b % p
    while e:
        if e & 1:
            r = r * b % p
        b = b * b % p
        e >>= 1
    return r


def _int_strings(a):
    """int64 array of decimal strings of at most 18 characters, else None."""
    if a.size and numpy.char.str_len(a).max() > 18:
        return None
    try:
        return a.astype(numpy.int64)
    except ValueError:
        return None


def residues_mod_p(values, p):
    """Values (ints, Fractions, or their strings) reduced modulo the prime p as
    a uint64 array; a fraction's denominator must not be divisible by p.
    Decimal strings of integers and of fractions a/b whose parts fit int64 are
    reduced as whole arrays (the inverse of b by Fermat, vectorised, when
    p^2 < 2^63); everything else entry by entry with exact integers."""
    p = int(p)
    values = list(values)
    out = numpy.zeros(len(values), dtype=numpy.uint64)
    done = numpy.zeros(len(values), dtype=bool)
    if values and p < 2 ** 62 and all(isinstance(v, str) for v in values):
        arr = numpy.asarray(values)
        isfrac = numpy.char.find(arr, "/") >= 0
        idx = numpy.nonzero(~isfrac)[0]
        ints = _int_strings(arr[idx])
        if ints is not None:
            out[idx] = numpy.mod(ints, p).astype(numpy.uint64)
            done[idx] = True
        idx = numpy.nonzero(isfrac)[0]
        if idx.size:
            parts = numpy.char.partition(arr[idx], "/")
            num, den = _int_strings(parts[:, 0]), _int_strings(parts[:, 2])
            if num is not None and den is not None:
                dm = numpy.mod(den, p)
                if (dm == 0).any():
                    i = int(idx[numpy.nonzero(dm == 0)[0][0]])
                    raise ValueError(f"entry {i}: denominator {values[i].split('/')[1]} is divisible by p = {p}")
                if p < 3037000499:  # p^2 < 2^63: int64 products
                    inv = _modpow_vec(dm, p - 2, p)
                    out[idx] = (numpy.mod(num, p) * inv % p).astype(numpy.uint64)
                else:
                    out[idx] = [int(a) % p * pow(int(b), p - 2, p) % p for a, b in zip(num.tolist(), dm.tolist())]
                done[idx] = True
    for i in numpy.nonzero(~done)[0].tolist():
        v = values[i]
        f = v if isinstance(v, Fraction) else Fraction(v)
        den = f.denominator % p
        if den == 0:
            raise ValueError(f"entry {i}: denominator {f.denominator} is divisible by p = {p}")
        out[i] = (f.numerator % p) * pow(den, p - 2, p) % p
    return out


def _storage(p):
    return numpy.uint32 if p <= 2 ** 32 - 1 else numpy.uint64


def device_matrix_from_coo(coo_or_file, field_characteristic, shape=None, fmt=None):
    """Dense row-major device matrix (torch tensor on cuda: int32 holding uint32
    bits for p < 2^32, else int64 holding uint64) of a COO dictionary or file
    ('.json' or plain text, or ``fmt`` = "json" / "txt"), entries reduced mod p.
    ``shape`` defaults to (max row + 1, max column + 1)."""
    import torch
    p = int(field_characteristic)
    if p < 2:
        raise ValueError("device_matrix_from_coo needs a prime field characteristic")
    if isinstance(coo_or_file, dict):
        ri, ci = _coo_arrays(coo_or_file)
        vals = list(coo_or_file.values())
    else:
        name = str(coo_or_file)
        kind = fmt or ("json" if name.endswith(".json") else "txt")
        if kind == "json":
            with open(name, 'r') as f:
                data = json.load(f)
            ri, ci = _parse_keys(list(data.keys()))
            vals = list(data.values())
        elif kind == "txt":
            r, c, vals = _read_plain_txt(name)
            ri = numpy.asarray(r).astype(numpy.int64)
            ci = numpy.asarray(c).astype(numpy.int64)
        else:
            raise ValueError(f"unknown COO format {kind!r}")
    if len(ri) == 0:
        raise ValueError("empty COO input")
    rows, cols = (int(ri.max()) + 1, int(ci.max()) + 1) if shape is None else (int(shape[0]), int(shape[1]))
    if ri.min() < 0 or ci.min() < 0 or ri.max() >= rows or ci.max() >= cols:
        raise ValueError(f"COO index outside the shape ({rows}, {cols})")
    res = residues_mod_p(vals, p)
    dt = _storage(p)
    lib = _lib.load()
    if dt == numpy.uint32:
        out = torch.empty((rows, cols), dtype=torch.int32, device="cuda")
        v = numpy.ascontiguousarray(res.astype(numpy.uint32))
        fn = lib.linac_coo_scatter_u32
    else:
        out = torch.empty((rows, cols), dtype=torch.int64, device="cuda")
        v = numpy.ascontiguousarray(res)
        fn = lib.linac_coo_scatter_u64
    ri = numpy.ascontiguousarray(ri)
    ci = numpy.ascontiguousarray(ci)
    torch.cuda.current_stream().synchronize()  # the library works on its own stream
    _lib.check(fn(ctypes.c_void_p(out.data_ptr()), cols, rows, cols, ri.ctypes.data_as(ctypes.c_void_p),
                  ci.ctypes.data_as(ctypes.c_void_p), v.ctypes.data_as(ctypes.c_void_p), len(ri), 0))
    return out
