"""Certificates for GF(p) row reductions at full size, computed on the device
by a route that shares nothing with the library: torch float64 products of
16-bit limbs (rocBLAS), which are exact because every partial sum is an
integer below 2**53, followed by int64 reductions mod p.

Also the seeded inputs of the configurations (device-generated uniform
matrices, L*U products of given rank, float systems)."""
import numpy

LIMB = 1 << 16


def _torch():
    import torch
    return torch


def uniform_u32(rows, cols, p, seed, ld=None, chunk_elems=1 << 28):
    """rows x ld int32 tensor (uint32 bits) on the device, columns [0, cols)
    uniform in [0, p) from a seeded device generator, the rest zero."""
    torch = _torch()
    ld = cols if ld is None else ld
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    out = torch.zeros((rows, ld), dtype=torch.int32, device="cuda")
    step = max(1, chunk_elems // max(1, cols))
    for r0 in range(0, rows, step):
        r1 = min(rows, r0 + step)
        out[r0:r1, :cols] = torch.randint(0, p, (r1 - r0, cols), generator=g, device="cuda",
                                          dtype=torch.int64).to(torch.int32)
    return out


def as_u64(t):
    """uint32 bits held in an int32 tensor -> int64 tensor of the values."""
    return t.to(_torch().int64) & 0xFFFFFFFF


def _limbs(t):
    """int64 tensor of values < 2**32 -> (low, high) 16-bit limbs as float64."""
    torch = _torch()
    return (t & 0xFFFF).to(torch.float64), (t >> 16).to(torch.float64)


def _combine(d00, d01, d10, d11, p):
    """(d00 + 2**16 (d01 + d10) + 2**32 d11) mod p from exact float64 limb
    products (each below 2**53); intermediates stay below 2**63."""
    torch = _torch()
    q = [x.to(torch.int64) % p for x in (d00, d01, d10, d11)]
    mid = (q[1] + q[2]) % p
    hi = (q[3] * LIMB) % p
    hi = (hi * LIMB) % p
    return (q[0] + (mid * LIMB) % p + hi) % p


def matmul_mod(A, B, p, chunk_rows=4096):
    """(A @ B) mod p for int64 device tensors with entries in [0, p), p < 2**32,
    inner dimension below 2**21 (so that limb sums stay exact in float64)."""
    torch = _torch()
    assert A.shape[1] == B.shape[0] and A.shape[1] < (1 << 21)
    b0, b1 = _limbs(B)
    out = torch.empty((A.shape[0], B.shape[1]), dtype=torch.int64, device=A.device)
    for r0 in range(0, A.shape[0], chunk_rows):
        a0, a1 = _limbs(A[r0:r0 + chunk_rows])
        out[r0:r0 + chunk_rows] = _combine(a0 @ b0, a0 @ b1, a1 @ b0, a1 @ b1, p)
        del a0, a1
    return out


def matvec_mod(A32, v, p, cols=None, chunk_rows=4096):
    """(A @ v) mod p where A32 is an int32 tensor of uint32 values (only its
    first `cols` columns are used) and v an int64 vector in [0, p)."""
    torch = _torch()
    cols = A32.shape[1] if cols is None else cols
    V = torch.stack(_limbs(v.to(torch.int64)), dim=1)  # cols x 2
    out = torch.empty(A32.shape[0], dtype=torch.int64, device=A32.device)
    for r0 in range(0, A32.shape[0], chunk_rows):
        a = as_u64(A32[r0:r0 + chunk_rows, :cols])
        a0, a1 = _limbs(a)
        del a
        d0 = a0 @ V
        d1 = a1 @ V
        out[r0:r0 + chunk_rows] = _combine(d0[:, 0], d0[:, 1], d1[:, 0], d1[:, 1], p)
        del a0, a1, d0, d1
    return out


def matvec_mod_u64(A64, x, p, cols=None, chunk_rows=2048):
    """(A @ x) mod p for 2**32 <= p < 2**63: A64 an int64 device tensor with
    entries in [0, p), x a list of Python ints.  Four 16-bit limbs per operand,
    exact float64 limb products on the device, combined with Python integers.
    Returns a list of Python ints."""
    torch = _torch()
    cols = A64.shape[1] if cols is None else cols
    xv = torch.tensor([int(v) for v in x], dtype=torch.int64, device=A64.device)
    X = torch.stack([((xv >> (16 * j)) & 0xFFFF).to(torch.float64) for j in range(4)], dim=1)
    D = torch.empty((A64.shape[0], 4, 4), dtype=torch.float64, device=A64.device)
    for r0 in range(0, A64.shape[0], chunk_rows):
        a = A64[r0:r0 + chunk_rows, :cols]
        for i in range(4):
            D[r0:r0 + chunk_rows, i] = ((a >> (16 * i)) & 0xFFFF).to(torch.float64) @ X
    Dh = D.to(torch.int64).cpu().numpy()
    w = [[pow(2, 16 * (i + j), p) for j in range(4)] for i in range(4)]
    return [sum(int(Dh[r, i, j]) * w[i][j] for i in range(4) for j in range(4)) % p for r in range(Dh.shape[0])]


def check_identity_block(W, n, chunk_rows=8192):
    """True when W[:, :n] (int tensor, n x >=n) is the n x n identity."""
    torch = _torch()
    for r0 in range(0, n, chunk_rows):
        blk = W[r0:r0 + chunk_rows, :n]
        rows = blk.shape[0]
        if int(blk.count_nonzero().item()) != rows:
            return False
        idx = torch.arange(rows, device=blk.device)
        if not bool((blk[idx, idx + r0] == 1).all().item()):
            return False
    return True


# ---- float systems of the precision configurations ------------------------

def float_system(kind, n, seed=0):
    """[A | b] (n x (n + 1)) of a float configuration: 'f64' uniform [0, 1);
    'c128' real and imaginary parts N(0, 1); 'illcond' U diag(logspace(0, -8))
    V^T with Haar U, V (QR of Gaussians, sign-fixed) and b uniform [0, 1)."""
    rs = numpy.random.RandomState(seed)
    if kind == "f64":
        return rs.random_sample((n, n + 1))
    if kind == "c128":
        return rs.standard_normal((n, n + 1)) + 1j * rs.standard_normal((n, n + 1))
    if kind != "illcond":
        raise ValueError(kind)

    def haar(m):
        q, r = numpy.linalg.qr(rs.standard_normal((m, m)))
        return q * numpy.sign(numpy.diag(r))

    A = numpy.empty((n, n + 1))
    A[:, :n] = (haar(n) * numpy.logspace(0, -8, n)) @ haar(n).T
    A[:, n] = rs.random_sample(n)
    return A


def refine_solution(A, B, X, iters=4):
    """X (solutions of A X = B, square A) refined with residuals in 80-bit long
    double and LAPACK LU corrections: accurate to about cond(A) 2^-64 + 2^-53."""
    import scipy.linalg
    lu = scipy.linalg.lu_factor(A)
    ext = numpy.clongdouble if numpy.iscomplexobj(A) or numpy.iscomplexobj(B) else numpy.longdouble
    Al, Bl, Xl = A.astype(ext), B.astype(ext), X.astype(ext)
    for _ in range(iters):
        R = Bl - Al @ Xl
        Xl = Xl + scipy.linalg.lu_solve(lu, R.astype(A.dtype)).astype(ext)
    return Xl.astype(A.dtype)
