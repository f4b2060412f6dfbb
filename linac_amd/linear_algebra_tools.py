"""Pivots, kernels and echelon forms of row-reduced matrices
(``linac/linear_algebra_tools.py:10-90``).

Same results as the reference's functions for the matrices the GPU path
returns, computed with whole-array numpy operations instead of Python loops
over rows.  For GF(p) output ``rref_with_pivots`` returns the pivot list
directly from the device, and the kernel constructions accept it through
``pivots=`` so that no scan runs at all.
"""
import numpy


def _ones_mask(A, target):
    """Elementwise ``A == target`` with the reference's tolerance rules: exact
    for integer arrays, numpy.isclose for float/complex, exact for objects that
    do not convert (huge integers, Fractions outside float range)."""
    A = numpy.asarray(A)
    if A.dtype.kind in 'ui':
        return A == target
    try:
        return numpy.isclose(A.astype(complex), target)
    except (TypeError, OverflowError):
        return A == target


def pivot_columns_from_row_reduced_echelon_form(row_reduced_echelon_form):
    """First column in each row whose entry is (close to) 1; IndexError if a row has none."""
    R = numpy.asarray(row_reduced_echelon_form)
    if R.shape[0] == 0:
        return []
    ones = _ones_mask(R, 1)
    if not ones.any(axis=1).all():
        raise IndexError("index 0 is out of bounds for axis 0 with size 0")
    return numpy.argmax(ones, axis=1).tolist()


def non_pivot_columns_from_row_reduced_echelon_form(row_reduced_echelon_form, pivots=None):
    """Columns that hold no pivot, in increasing order."""
    R = numpy.asarray(row_reduced_echelon_form)
    piv = pivot_columns_from_row_reduced_echelon_form(R) if pivots is None else list(pivots)
    keep = numpy.ones(R.shape[1], dtype=bool)
    keep[piv] = False
    return numpy.nonzero(keep)[0].tolist()


def non_pivot_columns_from_canonical_kernel(canonical_kernel):
    """For each kernel column, the last row holding (close to) -1."""
    K = numpy.asarray(canonical_kernel)
    if K.shape[1] == 0:
        return []
    neg = _ones_mask(K, -1)
    if not neg.any(axis=0).all():
        raise IndexError("index -1 is out of bounds for axis 0 with size 0")
    return (K.shape[0] - 1 - numpy.argmax(neg[::-1, :], axis=0)).tolist()


def pivot_columns_from_canonical_kernel(canonical_kernel):
    K = numpy.asarray(canonical_kernel)
    keep = numpy.ones(K.shape[0], dtype=bool)
    keep[non_pivot_columns_from_canonical_kernel(K)] = False
    return numpy.nonzero(keep)[0].tolist()


def isclose_or_greater(x, y, *args, **kwargs):
    return numpy.isclose(x, y, *args, **kwargs) or x > y


def drop_bottom_zero_rows(row_reduced_matrix):
    """Drop trailing rows whose max |x| is neither close to nor above 1."""
    R = row_reduced_matrix
    if len(R) == 0:
        return R
    mags = numpy.abs(numpy.asarray(R)).max(axis=1).astype(float)
    keep = numpy.isclose(mags, 1) | (mags > 1)
    n = len(R)
    while n > 0 and not keep[n - 1]:
        n -= 1
    return R[:n]


def canonical_kernel_from_row_reduced_echelon_form(row_reduced_echelon_form, pivots=None):
    """Kernel (null space) in canonical form of a matrix given in reduced echelon
    form with no zero rows: the non-pivot columns of R stacked on -I, with the
    rows of -I interleaved at the non-pivot positions.  The result has R's
    dtype promoted with int (the reference concatenates R with an int -I)."""
    R = numpy.asarray(row_reduced_echelon_form)
    r, m = R.shape
    piv = pivot_columns_from_row_reduced_echelon_form(R) if pivots is None else list(pivots)
    nonpiv = non_pivot_columns_from_row_reduced_echelon_form(R, piv)
    identity = numpy.identity(m - r, dtype=int)
    stacked = numpy.concatenate((R[:, nonpiv], -identity), axis=0)
    # row i of the kernel is row order.index(i) of the stacked block, order = pivots + non-pivots
    where = numpy.empty(m, dtype=numpy.int64)
    where[numpy.asarray(piv + nonpiv, dtype=numpy.int64)] = numpy.arange(m)
    return stacked[where, :]


def row_reduced_echelon_form_from_canonical_kernel(canonical_kernel):
    """Inverse of canonical_kernel_from_row_reduced_echelon_form."""
    K = numpy.asarray(canonical_kernel)
    m, k = K.shape
    nonpiv = non_pivot_columns_from_canonical_kernel(K)
    piv = pivot_columns_from_canonical_kernel(K)
    identity = numpy.identity(m - k, dtype=int)
    stacked = numpy.concatenate((identity, K[piv, :]), axis=1)
    where = numpy.empty(m, dtype=numpy.int64)
    where[numpy.asarray(piv + nonpiv, dtype=numpy.int64)] = numpy.arange(m)
    return stacked[:, where]


def invert_permutation(permutation):
    """Inverse of a permutation given as a sequence of indices (returned as a tuple)."""
    return tuple(numpy.argsort(permutation))


def permutation_matrix_from_permutation(permutation):
    """Permutation (as list of indices) to permutation matrix: P[i, permutation[i]] = 1."""
    n = len(permutation)
    P = numpy.zeros((n, n), dtype=int)
    if n:
        P[numpy.arange(n), numpy.asarray(permutation)] = 1
    return P
