"""Exact modular helpers for building and certifying test inputs (test infrastructure only).

None of this restates reference code; it is textbook arithmetic used to pin the
oracle and the GPU results independently of any elimination order:

* ``matmul_modp``    exact A @ B mod p through 16-bit limb splitting on float64
                     BLAS (every partial sum is an integer < 2**53);
* ``freivalds``      A == A[:, pivots] @ R (mod p) checked on random vectors,
                     which holds iff every row of A lies in the row space of R;
* ``planted``        [A | A x] whose RREF must be [I | x] for invertible A;
* ``brute_rref``     Python-int Gauss-Jordan with first-non-zero pivoting
                     (a different pivot rule from the oracle: the RREF over a
                     field is unique, so both must agree exactly);
* ``sympy_rref``     sympy's ``DomainMatrix(..., GF(p)).rref()``;
* ``check_rref_shape`` the defining properties of a reduced row-echelon form.
"""
import numpy


def matmul_modp(A, B, p):
    """Exact (A @ B) mod p for p < 2**31 (uint/int arrays with entries in [0, p))."""
    A = numpy.asarray(A, dtype=numpy.int64) % p
    B = numpy.asarray(B, dtype=numpy.int64) % p
    if p >= 2 ** 31:
        Ao, Bo = A.astype(object), B.astype(object)
        return numpy.asarray((Ao @ Bo) % p, dtype=numpy.int64)
    K = A.shape[1]
    out = numpy.zeros((A.shape[0], B.shape[1]), dtype=numpy.int64)
    step = 1 << 20  # K * 2**32 < 2**53
    c16, c32 = (1 << 16) % p, (1 << 32) % p
    for k0 in range(0, K, step):
        a = A[:, k0:k0 + step]
        b = B[k0:k0 + step, :]
        a1, a0 = (a >> 16).astype(numpy.float64), (a & 0xFFFF).astype(numpy.float64)
        b1, b0 = (b >> 16).astype(numpy.float64), (b & 0xFFFF).astype(numpy.float64)
        x11 = (a1 @ b1).astype(numpy.int64) % p
        x10 = ((a1 @ b0).astype(numpy.int64) + (a0 @ b1).astype(numpy.int64)) % p
        x00 = (a0 @ b0).astype(numpy.int64) % p
        out = (out + x11 * c32 % p + x10 * c16 % p + x00) % p
    return out


def random_matrix(rows, cols, p, seed):
    return numpy.random.RandomState(seed).randint(p, size=(rows, cols), dtype=numpy.int64 if p >= 2 ** 31 else numpy.uint32)


def low_rank(rows, cols, rank, p, seed):
    """Seeded mod-p product of a rows x rank and a rank x cols uniform matrix."""
    rs = numpy.random.RandomState(seed)
    L = rs.randint(p, size=(rows, rank), dtype=numpy.int64)
    U = rs.randint(p, size=(rank, cols), dtype=numpy.int64)
    return matmul_modp(L, U, p)


def freivalds(A, R, pivots, p, trials=3, seed=1234):
    """True iff A == A[:, pivots] @ R[:len(pivots)] (mod p) on ``trials`` random vectors."""
    A = numpy.asarray(A, dtype=numpy.int64) % p
    R = numpy.asarray(R, dtype=numpy.int64)[:len(pivots)] % p
    rs = numpy.random.RandomState(seed)
    for _ in range(trials):
        v = rs.randint(p, size=(A.shape[1], 1), dtype=numpy.int64)
        lhs = matmul_modp(A, v, p)
        rhs = matmul_modp(A[:, pivots], matmul_modp(R, v, p), p) if len(pivots) else numpy.zeros_like(lhs)
        if not numpy.array_equal(lhs, rhs):
            return False
    return True


def check_rref_shape(R, pivots, p=None):
    """Defining properties of an RREF with the given pivot columns.

    Rows < rank have a 1 at their pivot column, zeros before it and zeros in
    every other pivot column; pivot columns strictly increase; rows >= rank are
    zero.  With ``p`` the entries must be canonical in [0, p).
    """
    R = numpy.asarray(R)
    r = len(pivots)
    if any(b <= a for a, b in zip(pivots, pivots[1:])):
        return False
    if p is not None and (R.astype(object) >= p).any():
        return False
    for k, c in enumerate(pivots):
        if R[k, c] != 1 or numpy.any(R[k, :c] != 0):
            return False
        col = R[:, c].copy()
        col[k] = 0
        if numpy.any(col != 0):
            return False
    return not numpy.any(R[r:] != 0)


def planted(n, p, seed, extra=1):
    """([A | A @ X] mod p, X) with A a seeded uniform n x n matrix."""
    rs = numpy.random.RandomState(seed)
    A = rs.randint(p, size=(n, n), dtype=numpy.int64)
    X = rs.randint(p, size=(n, extra), dtype=numpy.int64)
    return numpy.concatenate([A, matmul_modp(A, X, p)], axis=1), X


def brute_rref(A, p):
    """Python-int Gauss-Jordan, first non-zero row as pivot (tiny inputs only)."""
    M = [[int(x) % p for x in row] for row in numpy.asarray(A).tolist()]
    rows = len(M)
    cols = len(M[0]) if rows else 0
    i = 0
    pivots = []
    for j in range(cols):
        pr = next((r for r in range(i, rows) if M[r][j] != 0), None)
        if pr is None:
            continue
        M[i], M[pr] = M[pr], M[i]
        inv = pow(M[i][j], p - 2, p)
        M[i] = [(x * inv) % p for x in M[i]]
        for r in range(rows):
            if r != i and M[r][j] != 0:
                f = M[r][j]
                M[r] = [(x - f * y) % p for x, y in zip(M[r], M[i])]
        pivots.append(j)
        i += 1
        if i == rows:
            break
    return numpy.array(M, dtype=object).reshape(rows, cols), pivots


def sympy_rref(A, p):
    """sympy's GF(p) RREF: returns (R as object array, pivot columns)."""
    from sympy import GF
    from sympy.polys.matrices import DomainMatrix
    K = GF(p)
    rows = [[K(int(x)) for x in row] for row in numpy.asarray(A).tolist()]
    dm = DomainMatrix(rows, numpy.asarray(A).shape, K)
    R, piv = dm.rref()
    out = numpy.array([[int(K.to_int(x)) % p for x in row] for row in R.to_list()], dtype=object)
    return out, list(piv)
