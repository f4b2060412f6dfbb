"""Restatement of linac's pivot extraction (test infrastructure only).

``linac/linear_algebra_tools.py:10-16``: the pivot column of row i is the first
column whose entry is (close to) 1; ``:48-51``: trailing rows whose max |x| is
not close to or above 1 are dropped.
"""
import numpy


def pivot_columns(rref):
    try:
        return [int(numpy.where(numpy.isclose(rref[i, :].astype('complex'), 1))[0][0]) for i in range(rref.shape[0])]
    except TypeError:
        return [int(numpy.where(numpy.isclose(rref[i, :].astype('int64'), 1))[0][0]) for i in range(rref.shape[0])]


def drop_bottom_zero_rows(rref):
    def isclose_or_greater(x, y):
        return numpy.isclose(x, y) or x > y
    while len(rref) > 0 and not isclose_or_greater(float(abs(rref[-1]).max()), 1):
        rref = rref[:-1]
    return rref
