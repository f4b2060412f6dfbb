"""Restatement of linac's ``iterative_gaussian_solver`` (test infrastructure only).

``linac/linear_system_solver.py:24-79`` with ``use_gpu=True``: the square
system ``[A | b]`` (n x (n + 1)) is row reduced; unknowns whose columns are not
pivots are dropped as redundant (the augmented column never is, ``:47-49``),
unknowns whose solution entry is close to 0 (``numpy.isclose(complex(v), 0)``,
``:51-52``) are dropped as zero, and the next iteration reduces the original
system restricted to the surviving unknowns -- rows and columns of the same
indices, plus the augmented column (``:34-38``).  The loop stops when an
iteration drops nothing new or every unknown is dropped (``:74-75``), or after
``max_iterations``.  The result lists one value per unknown, 0 for dropped
ones, the surviving pivots' solution values in order otherwise (``:68-69``).

This restatement tracks the surviving unknowns as a sorted list of original
indices instead of re-indexing the shrinking matrix; the two are equivalent
because rows and columns are always removed together.  The RREF is the C
oracle's (``crref``), which restates the GPU path's arithmetic.
"""
import numpy

from . import crref, pivots


def _rref(matrix, p):
    if p:
        return crref.rref_modp(matrix, p)[0]
    m = numpy.asarray(matrix)
    if m.dtype.kind == 'c' and numpy.any(numpy.imag(m) != 0):
        return crref.rref_c128(m)[0]
    return crref.rref_f64(numpy.real(m))[0]


def iterative_gaussian_solver(matrix, max_iterations=1, field_characteristic=0):
    """Returns (solution, history) with history = [(dropped_redundant,
    dropped_zero)] after each iteration (sorted original indices)."""
    A = numpy.asarray(matrix)
    n = A.shape[0]
    assert A.shape[1] == n + 1
    p = int(field_characteristic)
    live = list(range(n))
    red, zero = set(), set()
    solution, history = None, []
    for _ in range(max_iterations):
        before = red | zero
        m = len(live)
        R = pivots.drop_bottom_zero_rows(_rref(A[numpy.ix_(live, live + [n])], p))
        local_red = [j for j in pivots.non_pivot_columns(R) if j != m] if len(R) else list(range(m))
        col = R[:, -1].tolist() if len(R) else []
        zero_pos = [i for i, v in enumerate(col) if numpy.isclose(complex(v), 0)]
        pivot_unknowns = [live[j] for j in range(m) if j not in local_red]
        red |= {live[j] for j in local_red}
        zero |= {pivot_unknowns[i] for i in zero_pos}
        values = [v for i, v in enumerate(col) if i not in zero_pos]
        dropped = red | zero
        it = iter(values)
        solution = [0 if i in dropped else next(it) for i in range(n)]
        history.append((sorted(red), sorted(zero)))
        live = [u for u in live if u not in dropped]
        if dropped == before or len(dropped) == n:
            break
    return solution, history
