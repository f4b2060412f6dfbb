"""A sweep of systems built by the loader in HBM and row-reduced in place by
one batch, against the loader oracle and the RREF oracle per system."""
import numpy
import pytest

from oracle import crref, loader
from test_loader_oracle import Field

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n_systems, n_vars, max_degree, n_monomials, n_samples, p, seed", [
    (5, 3, 4, 30, 40, 2 ** 31 - 1, 0),      # tall: full column rank
    (4, 4, 3, 33, 25, 2 ** 31 - 19, 1),     # wide: rank limited by the samples
    (3, 2, 9, 50, 50, 65537, 2),            # square, small prime
])
def test_sweep_matches_oracle(gpu, n_systems, n_vars, max_degree, n_monomials, n_samples, p, seed):
    from linac_amd import sweep
    bases, lind, monomials = sweep.sweep_problem(n_systems, n_vars, max_degree, n_monomials, n_samples, p, seed)
    res = sweep.reduce_sweep(bases, lind, p)
    assert len(res.ranks) == n_systems
    for s in range(n_systems):
        A = loader.load_reference(bases[s], lind, p)
        Ro, po = crref.rref_modp(A, p)
        assert res.pivots[s] == po and res.ranks[s] == len(po)
        assert numpy.array_equal(res.host(s).astype(numpy.uint64), Ro.astype(numpy.uint64))


def test_sweep_system_equals_load_matrices(gpu):
    """The sweep's basis/index layout is the one load_matrices builds from
    dict-form points (pycuda_matrix_loader.py:174-202)."""
    from linac_amd import sweep
    p = 2 ** 31 - 1
    bases, lind, monomials = sweep.sweep_problem(2, 3, 4, 34, 20, p, 5)
    for s in range(2):
        pref, ans, pts = sweep.load_matrices_arguments(bases, monomials, s)
        pts['field'] = Field("finite field", p, 0)
        A = gpu.load_matrices(pref, ans, pts)[0]
        assert numpy.array_equal(A, loader.load_reference(bases[s], lind, p))


def test_sweep_in_place_admission(gpu):
    """An in-place matrix is admitted on its workspace alone."""
    from linac_amd.batch import Batch, footprint
    import torch
    rows, cols, p = 300, 280, 2 ** 31 - 1
    full = footprint(rows, cols)
    mat = rows * ((cols + 63) // 64 * 64) * 4
    ws = full - ((mat + 255) // 256 * 256)
    A = numpy.random.RandomState(3).randint(0, p, size=(rows, cols), dtype=numpy.int64).astype(numpy.uint32)
    d = torch.from_numpy(A.view(numpy.int32)).cuda()
    torch.cuda.synchronize()
    with Batch(ws) as b:
        with pytest.raises(MemoryError):
            b.add(A, p)  # a copy needs the matrix's bytes too
        rid = b.add_device(d.data_ptr(), rows, cols, cols, p, in_place=True)
        assert b.memory() == (ws, ws)
        b.run()
        R, piv, rank = b.take(rid)
    assert R is None
    Ro, po = crref.rref_modp(A, p)
    assert piv == po and numpy.array_equal(d.cpu().numpy().view(numpy.uint32), Ro)
