"""The algebra of the blocked engine's echelon form, restated in numpy on
small matrices: panels of width w eliminate with Gauss-Jordan on the rows
from r0 on (r0 = the panel's first unused row rounded down to a multiple of
`align`, or the previous panel's r0 for a panel that composes with a deferred
one); the rows above keep their entries.  The result is then completed by the
block back substitution in result order,
    G[0:s_l, F] -= A[piv_rows[0:s_l], P_l] G[s_l:s_l+q_l, F]   (last panel first),
and written as identity in the pivot columns.  It must equal the oracle's
RREF for every width, alignment and composition pattern, including
rank-deficient, zero-column and duplicate-column inputs."""
import numpy
import pytest

from oracle import crref


def _forward(A, p, W, align, compose):
    A = A.copy() % p
    N, M = A.shape
    used = numpy.zeros(N, bool)
    prow, pcol, panels = [], [], []
    r0 = 0
    deferred = False
    for J in range(0, M, W):
        lo = int(numpy.argmin(used)) if not used.all() else N
        if not (compose and deferred):
            r0 = lo // align * align
        X = A.copy()
        q = 0
        for c in range(J, min(J + W, M)):
            cand = [i for i in range(N) if not used[i] and X[i, c]]
            if not cand:
                continue
            i = cand[0]
            used[i] = True
            X[i] = X[i] * pow(int(X[i, c]), p - 2, p) % p
            for k in range(N):
                if k != i and X[k, c]:
                    X[k] = (X[k] - X[k, c] * X[i]) % p
            prow.append(i)
            pcol.append(c)
            q += 1
        A[r0:] = X[r0:]  # rows above r0 take no part of this panel's update
        if r0 > 0 and q > 0:
            panels.append((J, min(W, M - J), r0))
        deferred = (not deferred) and q > 0
        if used.all():
            break
    return A, prow, pcol, panels


def _complete(A, p, prow, pcol, panels):
    N, M = A.shape
    order = numpy.argsort(pcol)
    rows = [prow[k] for k in order]
    cols = [pcol[k] for k in order]
    rank = len(cols)
    F = [c for c in range(M) if c not in set(cols)]
    G = A[numpy.ix_(rows, F)].astype(numpy.int64) if rank and F else numpy.zeros((rank, len(F)), numpy.int64)
    for J, w, r0 in reversed(panels):
        sl = int(numpy.searchsorted(cols, J))
        el = int(numpy.searchsorted(cols, J + w))
        f0 = int(numpy.searchsorted(F, J))
        if el == sl or sl == 0 or f0 >= len(F):
            continue
        L = A[numpy.ix_(rows[:sl], cols[sl:el])].astype(object)
        G[:sl, f0:] = numpy.array((G[:sl, f0:].astype(object) - L @ G[sl:el, f0:].astype(object)) % p,
                                  dtype=numpy.int64)
    R = numpy.zeros((N, M), numpy.int64)
    for k in range(rank):
        R[k, cols[k]] = 1
        R[k, F] = G[k]
    return R, cols


def _inputs(seed, p):
    rs = numpy.random.RandomState(seed)
    out = []
    for t in range(24):
        N, M = int(rs.randint(2, 36)), int(rs.randint(2, 36))
        kind = t % 4
        if kind == 0:
            A = rs.randint(0, p, (N, M))
        elif kind == 1:
            r = int(rs.randint(1, min(N, M) + 1))
            A = (rs.randint(0, p, (N, r)).astype(object) @ rs.randint(0, p, (r, M)).astype(object)) % p
            A = numpy.array(A, dtype=numpy.int64)
        elif kind == 2:
            A = rs.randint(0, p, (N, M)) * (rs.rand(N, M) < 0.15)
        else:
            A = rs.randint(0, p, (N, M))
            A[:, rs.rand(M) < 0.3] = 0
            A[rs.rand(N) < 0.3] = 0
            for c in range(1, M, 5):
                A[:, c] = A[:, c - 1]
        out.append(A.astype(numpy.int64))
    return out


@pytest.mark.parametrize("p", [101, 2 ** 31 - 1])
@pytest.mark.parametrize("W, align, compose", [(1, 1, False), (3, 2, True), (4, 4, True), (8, 4, False), (5, 8, True)])
def test_echelon_completion_equals_rref(p, W, align, compose):
    for A in _inputs(W * 100 + align, p):
        E, prow, pcol, panels = _forward(A, p, W, align, compose)
        R, cols = _complete(E, p, prow, pcol, panels)
        want, wp = _oracle(A, p)
        assert cols == wp
        assert numpy.array_equal(R % p, want % p)


def _oracle(A, p):
    R, piv = crref.rref_modp(A.astype(numpy.uint64), p)
    return R.astype(numpy.int64), list(piv)


def test_a_dropped_correction_is_caught():
    """Skipping the back substitution leaves stale entries whenever r0 > 0:
    the restated completion above is what removes them."""
    p = 101
    A = numpy.random.RandomState(5).randint(0, p, (20, 26)).astype(numpy.int64)
    E, prow, pcol, panels = _forward(A, p, 4, 4, True)
    assert panels
    R, cols = _complete(E, p, prow, pcol, [])
    want, wp = _oracle(A, p)
    assert cols == wp and not numpy.array_equal(R % p, want % p)
