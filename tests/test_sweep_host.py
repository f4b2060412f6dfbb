"""Host-side construction of a sweep (exponent vectors, loader index table)."""
from math import comb

import numpy

from linac_amd.sweep import exponent_vectors, monomial_indices, sweep_problem


def test_exponent_vectors_count_and_order():
    ev = exponent_vectors(8, 6)
    assert len(ev) == comb(14, 6) == 3003 and len(set(ev)) == 3003
    assert ev[0] == (0,) * 8 and all(sum(a) <= sum(b) for a, b in zip(ev, ev[1:]))
    assert all(max(e) <= 6 and min(e) >= 0 for e in ev)


def test_monomial_indices_evaluate_the_monomial():
    """Evaluating the index table on a basis row [1, x..., w] gives w * prod x^e."""
    p = 10007
    mons = [(0, 0, 0), (2, 0, 1), (0, 3, 0), (1, 1, 1)]
    lind = monomial_indices(mons, 3, 3)
    assert lind.shape == (4, 4) and (lind[:, 3] == 4).all()
    row = [1, 3, 5, 7, 11]
    for e, idx in zip(mons, lind):
        got = 1
        for k in idx:
            got = got * row[k] % p
        assert got == 11 * 3 ** e[0] * 5 ** e[1] * 7 ** e[2] % p


def test_sweep_problem_shapes_and_range():
    p = 2 ** 31 - 1
    bases, lind, mons = sweep_problem(3, 8, 6, 3000, 17, p, 0)
    assert bases.shape == (3, 17, 10) and bases.dtype == numpy.uint32 and lind.shape == (3000, 7)
    assert (bases[:, :, 0] == 1).all() and int(bases.max()) < p
    assert len(set(mons)) == 3000
    again = sweep_problem(3, 8, 6, 3000, 17, p, 0)
    assert numpy.array_equal(again[0], bases) and numpy.array_equal(again[1], lind)
