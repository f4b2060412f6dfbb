"""Pivot extraction on a row-reduced matrix (``linac/linear_algebra_tools.py:10-16,48-51``).

Same results as the reference's functions for the matrices the GPU path
returns, computed with whole-array numpy operations instead of a Python loop
over rows.  For GF(p) output ``rref_with_pivots`` returns the same pivot list
directly from the device, with no scan at all.
"""
import numpy


def pivot_columns_from_row_reduced_echelon_form(row_reduced_echelon_form):
    """First column in each row whose entry is (close to) 1; IndexError if a row has none."""
    R = numpy.asarray(row_reduced_echelon_form)
    if R.shape[0] == 0:
        return []
    if R.dtype.kind in 'ui':
        ones = R == 1
    else:
        try:
            ones = numpy.isclose(R.astype(complex), 1)
        except (TypeError, OverflowError):
            ones = R == 1
    has = ones.any(axis=1)
    if not has.all():
        raise IndexError("index 0 is out of bounds for axis 0 with size 0")
    return numpy.argmax(ones, axis=1).tolist()


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
