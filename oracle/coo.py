"""Restatement of linac's COO helpers (test infrastructure only).

``linac/sparse_matrix_tools.py:20-42``: the COO dictionary lists the non-zero
entries row by row; when they do not reach the last row and column, the
bottom-right entry is added so that the shape survives the round trip.  The
dense matrix of a COO dictionary is an object array of zeros with the entries
written in.
"""
import numpy


def coo_from_matrix(matrix):
    rows, columns = matrix.shape
    coo = {}
    for r in range(rows):
        for c in range(columns):
            if matrix[r, c] != 0:
                coo[(r, c)] = matrix[r, c]
    if (coo == {} or rows != max(r for r, c in coo) + 1 or columns != max(c for r, c in coo) + 1):
        coo[(rows - 1, columns - 1)] = matrix[rows - 1, columns - 1]
    return coo


def matrix_from_coo(coo):
    rows = max(r for r, c in coo) + 1
    columns = max(c for r, c in coo) + 1
    matrix = numpy.zeros((rows, columns), dtype=object)
    for key, val in coo.items():
        matrix[key] = val
    return matrix
