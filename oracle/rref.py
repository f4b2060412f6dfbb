"""numpy restatement of linac's CPU ``row_reduce`` (test infrastructure only).

Follows ``linac/row_reduce.py:16-78`` step by step:

* ``prime == 0`` is treated as ``None`` (``row_reduce.py:21-22``);
* ``pivoting`` must be 0..3 else ``ValueError`` (``:24-26``);
* ``scaling=True`` divides every row by its max ``|x|`` once, before the loop
  (``:32-34``) -- note this turns integer input into floats, exactly like the
  reference;
* with ``prime`` the matrix is computed in int64 modulo ``prime`` (``:35-38``);
* the ``while i < rows and j < cols`` loop (``:40``) with partial pivoting
  ``argmax |A[i:, j]|`` (``:45-47``, first maximum wins, as numpy's argmax),
  rook / total pivoting with column permutations tracked in
  ``variable_order`` (``:49-58``);
* pivot accepted iff ``|A[i, j]| > threshold`` (``:60``), row normalised by the
  modular inverse (extended Euclid, ``ZeroDivisionError`` if gcd != 1,
  ``:61-65``) or by the head (``:66-67``);
* elimination below (``:68``) and, for reduced echelon form, above (``:69-70``);
* the trailing block reduced modulo prime after every step (``:75-76``).

Returns ``(matrix, variable_order)`` like the reference.
"""
import fractions

import numpy


def extended_euclid(a, b):
    """Return (s, t, g) with s*a + t*b == g == gcd(a, b).

    Stands in for ``pyadic.finite_field.extended_euclidean_algorithm`` as used
    at ``linac/row_reduce.py:37,62``; only ``s`` (the inverse of a mod b when
    g == 1) and ``g`` are consumed there.
    """
    a, b = int(a), int(b)
    old_r, r = a, b
    old_s, s = 1, 0
    old_t, t = 0, 1
    while r != 0:
        q = old_r // r
        old_r, r = r, old_r - q * r
        old_s, s = s, old_s - q * s
        old_t, t = t, old_t - q * t
    if old_r < 0:
        old_r, old_s, old_t = -old_r, -old_s, -old_t
    return old_s, old_t, old_r


def row_reduce(matrix, pivoting=1, scaling=True, reduced_echelon=True, threshold=10 ** -9, prime=None, progress=None):
    """Row reduction to (reduced) echelon form; see module docstring.
    ``progress`` (not in the reference): called with the column index every
    512 columns, for long timed runs."""
    if prime == 0:
        prime = None
    if pivoting not in [0, 1, 2, 3]:
        raise ValueError("Invalid pivoting.")

    i, j = 0, 0
    variable_order = list(range(matrix.shape[1]))

    if scaling is True:
        row_scales = abs(matrix).max(axis=1, keepdims=True)
        matrix = matrix / row_scales
    if prime is not None:
        if matrix.dtype is numpy.dtype(object):
            matrix = numpy.vectorize(
                lambda x: (x.numerator * extended_euclid(x.denominator, prime)[0]) % prime
                if isinstance(x, fractions.Fraction) else x)(matrix)
        matrix = matrix.astype('int64')
    else:
        matrix = matrix.copy()

    while i < matrix.shape[0] and j < matrix.shape[1]:
        if pivoting == 1:
            pivot_row = numpy.argmax(numpy.abs(matrix[i:, j])) + i
            matrix[[i, pivot_row], :] = matrix[[pivot_row, i], :]
        elif pivoting > 1:
            sub = numpy.abs(matrix[i:, j:])
            if pivoting == 2:
                flat = numpy.argmax(sub > 0.1)
            else:
                flat = numpy.argmax(sub)
            pivot_row, pivot_column = numpy.unravel_index(flat, sub.shape) + numpy.array([i, j])
            matrix[[i, pivot_row], :] = matrix[[pivot_row, i], :]
            matrix[:, [j, pivot_column]] = matrix[:, [pivot_column, j]]
            variable_order[j], variable_order[pivot_column] = variable_order[pivot_column], variable_order[j]

        if abs(matrix[i][j]) > threshold:
            if prime is not None:
                s, _, gcd = extended_euclid(matrix[i][j], prime)
                if gcd != 1:
                    raise ZeroDivisionError(
                        "Inverse of {} mod {} does not exist. Are you sure {} is prime?".format(matrix[i][j], prime, prime))
                matrix[i, :] = matrix[i, :] * (s % prime) % prime
            else:
                matrix[i, :] = matrix[i, :] / matrix[i][j]
            matrix[i + 1:, j:] = matrix[i + 1:, j:] - matrix[i + 1:, j: j + 1] * matrix[i: i + 1, j:]
            if reduced_echelon is True:
                matrix[:i, j:] = matrix[:i, j:] - matrix[:i, j: j + 1] * matrix[i: i + 1, j:]
            i, j = i + 1, j + 1
        else:
            j = j + 1

        if prime is not None:
            matrix[:, j:] = matrix[:, j:] % prime
        if progress is not None and j % 512 == 0:
            progress(j)

    return matrix, variable_order


def rref_modp(matrix, prime, progress=None):
    """The prime path the reference's own callers use.

    ``row_reduce(A, pivoting=1, scaling=False, threshold=0, prime=p)`` as at
    ``linac/vector_spaces/over_number_fields.py:109`` and
    ``linac/vector_spaces/over_function_fields.py:83``.  Requires
    ``prime < 2**31.5`` so int64 products cannot overflow (the reference has the
    same limit).
    """
    if prime * prime >= 2 ** 63:
        raise ValueError("int64 restatement needs prime**2 < 2**63; use the C oracle")
    matrix = numpy.asarray(matrix)
    if matrix.dtype != object:
        matrix = numpy.mod(matrix.astype(numpy.int64) if matrix.dtype != numpy.uint64 else matrix, prime).astype(numpy.int64)
    return row_reduce(matrix, pivoting=1, scaling=False, threshold=0, prime=prime, progress=progress)[0]


def rref_float(matrix, threshold=10 ** -9):
    """The float / complex path: partial pivoting with one-shot row scaling.

    ``cuda_row_reduce`` with ``field_characteristic=0`` pre-scales rows by their
    max ``|x|`` (``linac/row_reduce.cu:201-270``), searches ``argmax |x|`` and
    accepts a head iff ``|x| > 1e-9`` (``row_reduce.cu:404-418``): the same as
    ``row_reduce(A, pivoting=1, scaling=True, threshold=1e-9)``.
    """
    return row_reduce(numpy.asarray(matrix), pivoting=1, scaling=True, threshold=threshold)[0]
