"""Drop-in ``cuda_load_matrix`` and ``load_matrices`` on the gfx950 library.

Same signatures, defaults, dtype rules and return values as
``linac/pycuda_matrix_loader.py:28-230``:

* ``cuda_load_matrix``: uint64 when ``_mod64`` or (``_mod64`` is None and
  p > 2**31 - 1), else uint32 (``:61-68``; note the threshold differs from
  ``cuda_row_reduce``'s, as in the reference); float64 when ``_real`` or the
  first row of ``bases`` is real, else complex128 (``:69-76``);
* ``load_matrices``: the field comes from ``points['field']`` (dict of value
  columns) or ``points[0].field`` (point objects called with a variable name,
  ``:148-151``); monomials are padded with ``'1'`` (``:166-171``); the
  variables are ``'1'`` followed by the distinct names ordered by length
  (``:174``; ties broken by name here, which changes nothing in the output);
  one prefactor column is appended per system (``:184-190``) and each system's
  index table points at it (``:197-202``); an empty ansatz gives an empty
  uint32 matrix (``:226-227``).

The host preparation is whole-array numpy (no per-entry Python loops for
dict-form points).  ``use_cuda=False`` is not offered: this package computes
on the GPU only.  p-adic fields (which need pyadic's residue map,
``:158-161``) are not supported.
"""
import ctypes

import numpy

from . import _lib
from .timeit_decorator import timeit


def loader_dtype(field_characteristic, bases=None, _real=None, _mod64=None):
    p = int(field_characteristic)
    if p > 0:
        if _mod64 or (_mod64 is None and p > 2 ** 31 - 1):
            return numpy.dtype(numpy.uint64)
        return numpy.dtype(numpy.uint32)
    if _real:
        return numpy.dtype(numpy.float64)
    if _real is None and bases is not None:
        b0 = numpy.asarray(bases)[0] if len(bases) else numpy.zeros(1)
        if numpy.max(numpy.abs(numpy.imag(numpy.asarray(b0, dtype=complex))), initial=0) == 0:
            return numpy.dtype(numpy.float64)
    return numpy.dtype(numpy.complex128)


def _field_values(values, p, dtype):
    """Entries mod p as a C-contiguous array of dtype (object, signed or unsigned input)."""
    v = numpy.asarray(values)
    if v.dtype == object:
        v = numpy.vectorize(lambda x: int(x) % p, otypes=[object])(v) if v.size else v
        return numpy.ascontiguousarray(v.astype(dtype))
    if v.dtype.kind in 'iu':
        if v.dtype.kind == 'i' or v.dtype.itemsize > dtype.itemsize:
            v = numpy.mod(v.astype(numpy.int64) if v.dtype.kind == 'i' else v, p)
        return numpy.ascontiguousarray(v.astype(dtype))
    return numpy.ascontiguousarray(v.astype(dtype))


def _load_call(out, bases, lindices, p, dtype, flags=0, stream=None):
    lib = _lib.load()
    rows, basis_len = bases.shape
    cols, degree = lindices.shape
    args = (out.ctypes.data_as(ctypes.c_void_p) if isinstance(out, numpy.ndarray) else ctypes.c_void_p(out),
            bases.ctypes.data_as(ctypes.c_void_p) if isinstance(bases, numpy.ndarray) else ctypes.c_void_p(bases),
            lindices.ctypes.data_as(ctypes.c_void_p))
    if dtype == numpy.uint32:
        rc = lib.linac_load_matrix_u32(*args, rows, cols, basis_len, degree, p, flags, stream)
    elif dtype == numpy.uint64:
        rc = lib.linac_load_matrix_u64(*args, rows, cols, basis_len, degree, p, flags, stream)
    elif dtype == numpy.float64:
        rc = lib.linac_load_matrix_f64(*args, rows, cols, basis_len, degree, flags, stream)
    else:
        rc = lib.linac_load_matrix_c128(*args, rows, cols, basis_len, degree, flags, stream)
    _lib.check(rc)


@timeit
def cuda_load_matrix(bases, lindices, shape=None, field_characteristic=0, _real=None, _mod64=None, verbose=False):
    """
    Construct a dense matrix on the GPU by evaluating monomials as products of precomputed base values.

    Parameters
    ----------
    bases : (n_rows, basis_length) array
        Variable values per row/point (optionally including a prefactor column).
    lindices : (n_cols, degree) uint32 array
        Each row specifies a monomial as indices into `bases` (uniform degree).
    field_characteristic : int, default 0
        Modulus p for finite fields; 0 selects real/complex arithmetic.

    Returns
    -------
    A : (n_rows, n_cols) ndarray
    """
    bases = numpy.atleast_2d(numpy.asarray(bases))
    lindices = numpy.ascontiguousarray(numpy.atleast_2d(numpy.asarray(lindices)), dtype=numpy.uint32)
    n_rows, n_cols = bases.shape[0], lindices.shape[0]
    if shape is not None:
        assert tuple(shape) == (n_rows, n_cols), f"Shape mismatch: provided {shape}, inferred {(n_rows, n_cols)}"
    p = int(field_characteristic)
    dtype = loader_dtype(p, bases, _real, _mod64)
    if p > 0:
        if dtype == numpy.uint32 and p > 2 ** 32 - 1:
            raise ValueError(f"p = {p} does not fit uint32 storage; pass _mod64=True")
        b = _field_values(bases, p, dtype)
    else:
        src = numpy.real(bases) if dtype == numpy.float64 and numpy.iscomplexobj(bases) else bases
        b = numpy.ascontiguousarray(numpy.asarray(src).astype(dtype))
    out = numpy.zeros((n_rows, n_cols), dtype=dtype)
    if n_rows and n_cols:
        _load_call(out, b, lindices, p, dtype)
    return out


def load_matrices_stacked(bases_stack, lindices, field_characteristic, _real=None, _mod64=None):
    """Many systems sharing one index table in one launch.

    ``bases_stack``: (S, rows, basis_len); returns (S, rows, cols).  This is
    the sweep form of the loader: one system per evaluation point.
    """
    bs = numpy.asarray(bases_stack)
    S, rows, bl = bs.shape
    flat = cuda_load_matrix(bs.reshape(S * rows, bl), lindices, None, field_characteristic, _real, _mod64)
    return flat.reshape(S, rows, -1)


def _native(field):
    if field.characteristic == 0:
        return float if field.name in ['mpf', 'R'] else complex
    if field.name in ["padic", "Qp"]:
        raise NotImplementedError("p-adic fields need pyadic's residue map; pass finite-field values instead")
    return int


def _flatten(nested):
    out = []
    for x in nested:
        if isinstance(x, (list, tuple)):
            out.extend(_flatten(x))
        else:
            out.append(x)
    return out


def load_matrices(prefactors, ansatze, points, use_cuda=True, _use_galois=True, verbose=False):
    """
    Build one or more matrices representing polynomial linear systems evaluated at numerical points.

    Parameters
    ----------
    prefactors : list
        One prefactor per system (string key or callable).
    ansatze : list
        One ansatz per system; each ansatz is a list of monomials (lists of variable names).
    points : list or dict
        Evaluation points: a dict of value columns with a 'field' entry, or a list of point
        objects with a ``field`` attribute that return a value when called with a name.
    use_cuda : bool, default True
        Must be True: the matrices are built on the GPU.

    Returns
    -------
    As : list of ndarrays, one per system.
    """
    if not use_cuda:
        raise NotImplementedError("linac_amd builds matrices on the GPU only (use_cuda=True)")
    field = points['field'] if isinstance(points, dict) else points[0].field
    native = _native(field)
    p = int(field.characteristic)

    padded = []
    for ansatz in ansatze:
        width = max([len(m) for m in ansatz], default=0)
        padded.append([list(m) + ['1'] * (width - len(m)) for m in ansatz])
    names = sorted(set(_flatten(padded)) - {'1'}, key=lambda x: (len(x), x))
    variables = ['1'] + names
    col_of = {v: i for i, v in enumerate(variables)}

    def column(values):
        v = numpy.asarray(values)
        if p > 0:
            return v if v.dtype.kind in 'iu' else numpy.vectorize(lambda x: int(x) % p, otypes=[object])(v)
        return v.astype(complex if native is complex else float) if v.dtype != object else \
            numpy.vectorize(native, otypes=[object])(v)

    if isinstance(points, dict):
        cols = [column(points[v]) for v in variables]
        n = len(cols[0])
    else:
        n = len(points)
        cols = [column([pt(v) for pt in points]) for v in variables]

    As = []
    result_is_vector = False
    for i, prefactor in enumerate(prefactors):
        if isinstance(points, dict):
            pref = column(points[prefactor])
        else:
            pref = column([pt(prefactor) if isinstance(prefactor, str)
                           else (1 if prefactor(pt) == 1 else prefactor(pt)) for pt in points])
        base_cols = cols + [pref]
        if p > 0:
            dtype = loader_dtype(p)
            bases = numpy.stack([_field_values(c, p, dtype) for c in base_cols], axis=1)
        else:
            bases = numpy.stack([numpy.asarray(c) for c in base_cols], axis=1)
        if bases.ndim < 2:
            result_is_vector = True
        bases = numpy.atleast_2d(bases)
        width = len(padded[i][0]) if len(padded[i]) > 0 else 0
        lindices = numpy.zeros((len(padded[i]), width + 1), dtype=numpy.uint32)
        if len(padded[i]):
            lindices[:, :width] = numpy.array([[col_of[v] for v in m] for m in padded[i]], dtype=numpy.uint32)
            lindices[:, width] = bases.shape[1] - 1
            As.append(cuda_load_matrix(bases, lindices, (bases.shape[0], lindices.shape[0]),
                                       field_characteristic=p, _real=(native is float), verbose=verbose))
        else:
            As.append(numpy.empty((n, 0), dtype='uint32'))
    if result_is_vector:
        return [A[0] for A in As]
    return As
