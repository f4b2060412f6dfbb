"""Restatement of linac's pivot extraction (test infrastructure only).

``linac/linear_algebra_tools.py:10-19``: the pivot column of row i is the first
column whose entry is (close to) 1; ``:48-51``: trailing rows whose max |x| is
not close to or above 1 are dropped; ``:54-78``: canonical kernels.
"""
import numpy


def pivot_columns(rref):
    """``linac/linear_algebra_tools.py:10-19``."""
    try:
        return [int(numpy.where(numpy.isclose(rref[i, :].astype('complex'), 1))[0][0]) for i in range(rref.shape[0])]
    except TypeError:
        return [int(numpy.where(numpy.isclose(rref[i, :].astype('int64'), 1))[0][0]) for i in range(rref.shape[0])]
    except OverflowError:
        return [int(numpy.where(rref[i, :] == 1)[0][0]) for i in range(rref.shape[0])]


def drop_bottom_zero_rows(rref):
    def isclose_or_greater(x, y):
        return numpy.isclose(x, y) or x > y
    while len(rref) > 0 and not isclose_or_greater(float(abs(rref[-1]).max()), 1):
        rref = rref[:-1]
    return rref


def non_pivot_columns(rref):
    """``linac/linear_algebra_tools.py:22-24``."""
    piv = pivot_columns(rref)
    return [j for j in range(rref.shape[1]) if j not in piv]


def canonical_kernel(rref):
    """``linac/linear_algebra_tools.py:54-63``: the non-pivot columns of R on top
    of -I, rows interleaved so that row i of the kernel belongs to column i."""
    identity = numpy.identity(rref.shape[1] - rref.shape[0], dtype=int)
    nonpiv, piv = non_pivot_columns(rref), pivot_columns(rref)
    incomplete = numpy.delete(rref, piv, axis=1)
    scrambled = numpy.concatenate((incomplete, -identity), axis=0)
    order = piv + nonpiv
    return scrambled[[order.index(i) for i in range(len(order))], :]


def kernel_non_pivot_rows(kernel):
    """``linac/linear_algebra_tools.py:27-36``: last row holding -1 in each column."""
    try:
        return [int(numpy.where(numpy.isclose(kernel[:, i].astype('complex'), -1))[0][-1])
                for i in range(kernel.shape[1])]
    except TypeError:
        return [int(numpy.where(numpy.isclose(kernel[:, i].astype('int64'), -1))[0][-1])
                for i in range(kernel.shape[1])]
    except OverflowError:
        return [int(numpy.where(kernel[:, i] == -1)[0][-1]) for i in range(kernel.shape[1])]


def rref_from_canonical_kernel(kernel):
    """``linac/linear_algebra_tools.py:66-78``."""
    identity = numpy.identity(kernel.shape[0] - kernel.shape[1], dtype=int)
    nonpiv = kernel_non_pivot_rows(kernel)
    piv = [i for i in range(kernel.shape[0]) if i not in nonpiv]
    incomplete = numpy.delete(kernel, nonpiv, axis=0)
    scrambled = numpy.concatenate((identity, incomplete), axis=1)
    order = piv + nonpiv
    return scrambled[:, [order.index(i) for i in range(len(order))]]
