"""Restatement of linac's matrix loader (test infrastructure only).

``build_systems`` follows ``linac/pycuda_matrix_loader.py:148-202``: the field
from ``points['field']`` or ``points[0].field`` (``:148-151``), the native type
(``:152-163``; p-adic residues need pyadic and are not restated), padding of
every monomial with ``'1'`` (``:166-171``), ``variables = ['1'] + sorted(set(...),
key=len)`` (``:174``), the numerical bases per point (``:175-178``), a prefactor
column appended per system (``:184-190``) and ``lindices`` (``:197-202``).

``load_reference`` is the Python-int fallback of ``:221-225`` -- the arithmetic
of the ``LoadMatrix`` kernel (``linac/matrix_loader.cu:88-112``):
``A[r, c] = prod_t bases[r, lindices[c, t]]`` (mod p), with dtype uint64 when
p > 2**32 - 1 else uint32 (``:223``).
"""
from functools import reduce
from operator import mul

import numpy


def _flatten(nested):
    out = []
    for x in nested:
        if isinstance(x, (list, tuple)):
            out.extend(_flatten(x))
        else:
            out.append(x)
    return out


def build_systems(prefactors, ansatze, points):
    """Returns (field, native_type, [(bases, lindices), ...])."""
    field = points['field'] if isinstance(points, dict) else points[0].field
    if field.characteristic == 0:
        native = float if field.name in ['mpf', 'R'] else complex
    else:
        native = int
    padded = []
    for ansatz in ansatze:
        max_length = max([len(s) for s in ansatz], default=0)
        padded.append([list(s) + ['1'] * (max_length - len(s)) for s in ansatz])
    variables = ['1'] + sorted(list(set(_flatten(padded))), key=lambda x: len(x))
    if isinstance(points, dict):
        bases_monomials = numpy.vectorize(native, otypes='O')(numpy.array([points[v] for v in variables]).T.copy())
    else:
        bases_monomials = numpy.array([[native(pt(v)) for v in variables] for pt in points], dtype=object)
    systems = []
    for i, prefactor in enumerate(prefactors):
        if isinstance(points, dict):
            col = numpy.vectorize(native, otypes='O')(numpy.array([points[prefactor]]).T.copy())
        else:
            col = numpy.array([[native(pt(prefactor)) if isinstance(prefactor, str)
                                else 1 if prefactor(pt) == 1 else native(prefactor(pt))] for pt in points], dtype=object)
        bases = numpy.atleast_2d(numpy.block([bases_monomials, col]))
        width = len(padded[i][0]) if len(padded[i]) > 0 else 0
        lindices = numpy.zeros((len(padded[i]), width + 1), dtype='uint32')
        for j, monomial in enumerate(padded[i]):
            k = -1
            for k, var in enumerate(monomial):
                lindices[j, k] = variables.index(var)
            lindices[j, k + 1] = len(bases[0]) - 1
        systems.append((bases, lindices))
    return field, native, systems


def load_reference(bases, lindices, p):
    """Python-int (p > 0) or float/complex (p == 0) evaluation of every entry."""
    if p > 0:
        return numpy.array([[reduce(lambda a, b: int(a) * int(b) % p, [base[k] for k in idx]) for idx in lindices]
                            for base in bases], dtype='uint64' if p > 2 ** 32 - 1 else 'uint32').reshape(len(bases), len(lindices))
    return numpy.array([[reduce(mul, [base[k] for k in idx]) for idx in lindices] for base in bases]).reshape(len(bases), len(lindices))


def load_matrices_reference(prefactors, ansatze, points):
    field, native, systems = build_systems(prefactors, ansatze, points)
    out = []
    for bases, lindices in systems:
        if len(lindices) == 0:
            out.append(numpy.empty((bases.shape[0], 0), dtype='uint32'))
        else:
            out.append(load_reference(bases, lindices, field.characteristic))
    return out
