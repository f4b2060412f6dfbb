"""A functional-reconstruction sweep on the device (no counterpart in the reference).

One polynomial ansatz is evaluated at many outer points; each outer point
gives one linear system whose rows are inner sample points and whose columns
are the ansatz monomials times a per-point prefactor -- the matrices
``load_matrices`` builds (``linac/pycuda_matrix_loader.py:127-230``), one per
system.  Here all systems are built by one launch of the monomial loader
straight into HBM and row-reduced there by one batch, in place: no matrix
crosses PCIe, and only pivot columns and ranks come back unless asked for.

Basis layout of every system (the ``load_matrices`` convention,
``pycuda_matrix_loader.py:174-202``): column 0 is the constant 1, columns
1..n_vars the variables, the last column the prefactor; a monomial of degree
d < max_degree is padded with column 0.
"""
import ctypes
import itertools
import time

import numpy

from . import _lib
from .batch import Batch


def exponent_vectors(n_vars, max_degree):
    """Every exponent vector of total degree <= max_degree, by degree, then
    lexicographically in the variable indices (C(n_vars + d, d) of them)."""
    out = []
    for d in range(max_degree + 1):
        for combo in itertools.combinations_with_replacement(range(n_vars), d):
            e = [0] * n_vars
            for v in combo:
                e[v] += 1
            out.append(tuple(e))
    return out


def monomial_indices(monomials, n_vars, max_degree):
    """Loader index table (n_monomials, max_degree + 1): each variable's
    basis column repeated by its exponent, padding 0 (the constant), and the
    prefactor column n_vars + 1 last."""
    lind = numpy.zeros((len(monomials), max_degree + 1), dtype=numpy.uint32)
    for m, e in enumerate(monomials):
        idx = [v + 1 for v in range(n_vars) for _ in range(e[v])]
        lind[m, :len(idx)] = idx
    lind[:, max_degree] = n_vars + 1
    return lind


def sweep_problem(n_systems=64, n_vars=8, max_degree=6, n_monomials=3000, n_samples=3000, p=2 ** 31 - 1, seed=0):
    """Seeded synthetic sweep.

    ``n_monomials`` distinct exponent vectors drawn from all of total degree
    <= max_degree (a seeded permutation); per system, ``n_samples`` points
    uniform in GF(p)^n_vars and one prefactor per point uniform in GF(p).
    Returns ``(bases, lindices, monomials)`` with bases of shape
    (n_systems, n_samples, n_vars + 2), uint32.
    """
    rs = numpy.random.RandomState(seed)
    exps = exponent_vectors(n_vars, max_degree)
    if n_monomials > len(exps):
        raise ValueError(f"only {len(exps)} monomials of degree <= {max_degree} in {n_vars} variables")
    monomials = [exps[i] for i in rs.permutation(len(exps))[:n_monomials]]
    bases = numpy.empty((n_systems, n_samples, n_vars + 2), dtype=numpy.uint32)
    bases[:, :, 0] = 1
    bases[:, :, 1:n_vars + 1] = rs.randint(0, p, size=(n_systems, n_samples, n_vars), dtype=numpy.int64)
    bases[:, :, n_vars + 1] = rs.randint(0, p, size=(n_systems, n_samples), dtype=numpy.int64)
    return bases, monomial_indices(monomials, n_vars, max_degree), monomials


def load_matrices_arguments(bases, monomials, system):
    """The same system in ``load_matrices`` form: (prefactors, ansatze, points)
    with dict-form points (variable names x1..xn, prefactor 'w')."""
    n_vars = bases.shape[2] - 2
    names = [f"x{v + 1}" for v in range(n_vars)]
    ansatz = [[names[v] for v in range(n_vars) for _ in range(e[v])] for e in monomials]
    points = {name: bases[system, :, v + 1].astype(numpy.int64) for v, name in enumerate(names)}
    points['1'] = numpy.ones(bases.shape[1], dtype=numpy.int64)
    points['w'] = bases[system, :, n_vars + 1].astype(numpy.int64)
    return ['w'], [ansatz], points


class SweepResult:
    """Device-resident RREFs of a sweep plus their pivot columns and ranks."""

    def __init__(self, matrices, pivots, ranks, load_seconds, reduce_seconds):
        self.matrices = matrices          # torch int32 tensor (S, rows, cols) on the device, uint32 bits
        self.pivots = pivots              # list of lists
        self.ranks = ranks
        self.load_seconds = load_seconds
        self.reduce_seconds = reduce_seconds

    def host(self, system):
        """RREF of one system as a uint32 numpy array."""
        return self.matrices[system].cpu().numpy().view(numpy.uint32)


def reduce_sweep(bases, lindices, p, mem_cap_bytes=0):
    """Load every system with one loader launch into HBM and row-reduce them
    all in one batch, in place.  ``bases``: (S, rows, basis_len) uint32 in
    [0, p); ``lindices``: (cols, degree) uint32.  Returns a SweepResult.

    Device buffers are torch tensors (the library only sees their addresses);
    the library runs on its own stream, so host-visible ordering is kept by
    synchronising torch's stream before and after."""
    import torch

    p = int(p)
    if not 2 < p <= 4278124287:
        raise ValueError("the batched path needs 2 < p <= 4278124287")
    bases = numpy.ascontiguousarray(bases, dtype=numpy.uint32)
    lindices = numpy.ascontiguousarray(lindices, dtype=numpy.uint32)
    S, rows, basis_len = bases.shape
    cols = lindices.shape[0]
    dev = torch.device("cuda")
    d_bases = torch.from_numpy(bases.reshape(S * rows, basis_len).view(numpy.int32)).to(dev)
    d_lind = torch.from_numpy(lindices.view(numpy.int32)).to(dev)
    d_out = torch.empty((S, rows, cols), dtype=torch.int32, device=dev)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _lib.check(_lib.load().linac_load_matrix_u32(ctypes.c_void_p(d_out.data_ptr()), ctypes.c_void_p(d_bases.data_ptr()),
                                                 ctypes.c_void_p(d_lind.data_ptr()), S * rows, cols, basis_len,
                                                 lindices.shape[1], p, _lib.DEVICE_PTR, None))
    t1 = time.perf_counter()
    pivots, ranks = [], []
    with Batch(mem_cap_bytes) as b:
        item = rows * cols * 4
        ids = [b.add_device(d_out.data_ptr() + s * item, rows, cols, cols, p, in_place=True) for s in range(S)]
        b.run()
        for _, piv, rank in b.take_many(ids):
            pivots.append(piv)
            ranks.append(rank)
    t2 = time.perf_counter()
    del d_bases, d_lind
    return SweepResult(d_out, pivots, ranks, t1 - t0, t2 - t1)
