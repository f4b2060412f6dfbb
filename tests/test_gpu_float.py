"""float64 / complex128 RREF on the GPU: reference tolerance checks
(reference tests/test_row_reduce_random_cuda.py:23-49, complex cases), solves
against LAPACK, and the float oracle on rank-deficient inputs."""
import numpy
import pytest

from oracle import crref
from test_pivot_structure import KNOWN_MEDIUM_PIVOTS, planted_pivot_matrix

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [1024, 2048])
def test_random_complex_gives_identity(gpu, n):
    rs = numpy.random.RandomState(n)
    A = rs.rand(n, n) + 1j * rs.rand(n, n)
    R = gpu.cuda_row_reduce(A, 0)
    assert R.dtype == numpy.complex128
    assert numpy.all(numpy.isclose(R - numpy.identity(n), numpy.zeros((n, n))))


def test_real_input_uses_float64(gpu):
    rs = numpy.random.RandomState(1)
    A = rs.rand(300, 300)
    R = gpu.cuda_row_reduce(A.astype(complex), 0)
    assert R.dtype == numpy.float64 and numpy.allclose(R, numpy.identity(300))


@pytest.mark.parametrize("method", ["plain", "blocked"])
@pytest.mark.parametrize("n, kind", [(1000, "f64"), (700, "c128"), (1500, "f64")])
def test_augmented_solution_matches_lapack(gpu, n, kind, method):
    rs = numpy.random.RandomState(n)
    A = rs.random_sample((n, n))
    b = rs.random_sample((n, 3))
    if kind == "c128":
        A = A + 1j * rs.standard_normal((n, n))
        b = b + 1j * rs.standard_normal((n, 3))
    R = gpu.cuda_row_reduce(numpy.concatenate([A, b], axis=1), 0, method=method)
    x = numpy.linalg.solve(A, b)
    rel = numpy.max(numpy.abs(R[:, n:] - x)) / numpy.max(numpy.abs(x))
    assert rel < 1e-10, rel
    assert numpy.allclose(R[:, :n], numpy.identity(n), atol=1e-12)


@pytest.mark.parametrize("method", ["plain", "blocked"])
@pytest.mark.parametrize("shape, rank, seed, cplx", [
    ((60, 80), 35, 0, False), ((120, 70), 50, 1, True), ((200, 200), 199, 2, False),
    ((700, 900), 400, 3, False),   # several panels, rank-deficient panels
    ((333, 517), 333, 4, True),    # ragged sizes (not multiples of 64)
    ((3000, 150), 140, 5, False),  # tall: many workgroups in the panel kernel
    ((90, 1500), 90, 6, False),    # wide: many panels, few rows
])
def test_rank_deficient_matches_float_oracle(gpu, shape, rank, seed, cplx, method):
    rs = numpy.random.RandomState(seed)
    L = rs.standard_normal((shape[0], rank))
    U = rs.standard_normal((rank, shape[1]))
    if cplx:
        L = L + 1j * rs.standard_normal(L.shape)
        U = U + 1j * rs.standard_normal(U.shape)
    A = L @ U
    A[:, 5] = 0
    R, piv, rk = gpu.rref_with_pivots(A, 0, method=method)
    Ro, po = (crref.rref_c128 if cplx else crref.rref_f64)(A)
    assert piv == po and rk == len(po)
    assert numpy.max(numpy.abs(R - Ro)) < 1e-9 * max(1.0, numpy.max(numpy.abs(Ro)))


@pytest.mark.parametrize("field", [0, 2 ** 31 - 1])
def test_known_medium_pivot_structure(gpu, field):
    """Pivots [0..8, 10, 12..27, 38..41, 59] of a 32 x 86 rank-31 matrix
    (reference tests/test_null_space.py:26-62; its fixture is a pickle, so the
    matrix is rebuilt synthetically with the same RREF structure)."""
    A = planted_pivot_matrix(field, seed=3)
    R = gpu.cuda_row_reduce(A, field)
    R = gpu.drop_bottom_zero_rows(R)
    assert R.shape == (31, 86)
    assert gpu.pivot_columns_from_row_reduced_echelon_form(R) == KNOWN_MEDIUM_PIVOTS


@pytest.mark.parametrize("method", ["plain", "blocked"])
def test_integer_valued_ties_and_zero_columns(gpu, method):
    """Small integers make equal |x| common: the smallest row wins ties, as in
    the oracle; whole zero columns are skipped.  (No zero rows here: the
    oracle keeps the reference's division of a zero row by its zero maximum,
    whose NaNs stop the elimination; this library leaves such rows at zero.)"""
    rs = numpy.random.RandomState(7)
    A = rs.randint(-2, 3, size=(400, 300)).astype(numpy.float64)
    A[:, 10:20] = 0
    R, piv, rk = gpu.rref_with_pivots(A, 0, method=method)
    Ro, po = crref.rref_f64(A)
    assert piv == po and rk == len(po) == 290
    assert numpy.max(numpy.abs(R - Ro)) < 1e-9 * max(1.0, numpy.max(numpy.abs(Ro)))


def test_blocked_equals_plain_on_rank_deficient_sweep(gpu):
    """Both methods give the same pivots and agree to rounding on a matrix
    whose rank deficiency spans several panels."""
    rs = numpy.random.RandomState(8)
    A = rs.standard_normal((640, 37)) @ rs.standard_normal((37, 700))
    A = numpy.concatenate([A, rs.standard_normal((640, 5))], axis=1)
    Rp, pp, rp = gpu.rref_with_pivots(A, 0, method="plain")
    Rb, pb, rb = gpu.rref_with_pivots(A, 0, method="blocked")
    assert pp == pb and rp == rb == 42
    assert numpy.max(numpy.abs(Rp - Rb)) < 1e-9 * max(1.0, numpy.max(numpy.abs(Rp)))


@pytest.mark.parametrize("shape, rank, cplx, per_step", [
    ((700, 900), 400, False, 1), ((333, 517), 333, True, 3), ((1500, 1501), 1500, False, 0)])
def test_float_job_in_steps_equals_single_call(gpu, shape, rank, cplx, per_step):
    """A float job advanced a few panels per call (per_step 0: finish at once)
    gives exactly the single-call blocked result: the same kernels on the same
    data in the same order."""
    import torch
    from linac_amd.hip_float import FloatJob
    from linac_amd.hip_row_reduce import rref_device
    rs = numpy.random.RandomState(rank)
    A = rs.standard_normal((shape[0], rank)) @ rs.standard_normal((rank, shape[1]))
    if cplx:
        A = A + 1j * (rs.standard_normal((shape[0], rank)) @ rs.standard_normal((rank, shape[1])))
    dt = numpy.complex128 if cplx else numpy.float64
    one = torch.from_numpy(A.astype(dt)).cuda()
    steps = torch.from_numpy(A.astype(dt)).cuda()
    torch.cuda.synchronize()
    p1, r1 = rref_device(one.data_ptr(), shape[0], shape[1], shape[1], 0, dt, method="blocked")
    job = FloatJob(steps.data_ptr(), shape[0], shape[1], shape[1], dt)
    calls = 0
    if per_step:
        while job.step(per_step) > 0:
            calls += 1
        assert calls >= (shape[1] + 63) // 64 // per_step - 1
    p2, r2 = job.finish()
    torch.cuda.synchronize()
    assert (p1, r1) == (p2, r2) and r1 == min(rank, *shape)
    assert torch.equal(one, steps)


def test_float_job_one_per_device(gpu):
    """A second job on the same device is refused while the first is open;
    after destroy (or finish) a new one opens."""
    import torch
    from linac_amd.hip_float import FloatJob
    a = torch.rand((200, 201), dtype=torch.float64, device="cuda")
    b = a.clone()
    torch.cuda.synchronize()
    j1 = FloatJob(a.data_ptr(), 200, 201, 201, numpy.float64)
    with pytest.raises(ValueError, match="already open"):
        FloatJob(b.data_ptr(), 200, 201, 201, numpy.float64)
    j1.step(1)
    j1.close()
    j2 = FloatJob(b.data_ptr(), 200, 201, 201, numpy.float64)
    piv, rank = j2.finish()
    assert rank == 200 and piv == list(range(200))


def test_float_jobs_with_own_workspace_interleave(gpu):
    """Jobs with their own workspaces can be open together and advanced in any
    interleaving; each ends with exactly its single-call blocked result.  A
    workspace below the stated size is refused."""
    import torch
    from linac_amd.hip_float import FloatJob, float_job_workspace
    from linac_amd.hip_row_reduce import rref_device
    rs = numpy.random.RandomState(12)
    shapes = [(500, 700, numpy.float64), (900, 400, numpy.complex128), (256, 1300, numpy.float64)]
    mats, refs, jobs = [], [], []
    for rows, cols, dt in shapes:
        A = rs.standard_normal((rows, cols))
        if dt == numpy.complex128:
            A = A + 1j * rs.standard_normal((rows, cols))
        one = torch.from_numpy(A.astype(dt)).cuda()
        mats.append(torch.from_numpy(A.astype(dt)).cuda())
        torch.cuda.synchronize()
        refs.append((one, rref_device(one.data_ptr(), rows, cols, cols, 0, dt, method="blocked")))
    with pytest.raises(ValueError, match="needed"):
        small = torch.empty(float_job_workspace(500, 700) - 256, dtype=torch.uint8, device="cuda")
        FloatJob(mats[0].data_ptr(), 500, 700, 700, numpy.float64, workspace=small)
    for (rows, cols, dt), m in zip(shapes, mats):
        ws = torch.empty(float_job_workspace(rows, cols, dt), dtype=torch.uint8, device="cuda")
        jobs.append(FloatJob(m.data_ptr(), rows, cols, cols, dt, workspace=ws))
    order = rs.randint(0, 3, size=60)
    for k in order:
        if jobs[k].remaining != 0:
            jobs[k].step(int(rs.randint(1, 3)))
    for job, m, (one, ref) in zip(jobs, mats, refs):
        assert job.finish() == ref
        torch.cuda.synchronize()
        assert torch.equal(m, one)


@pytest.mark.parametrize("n, cols", [(1000, 1001), (2500, 2503), (777, 800)])
def test_mixed_method_against_lapack(gpu, n, cols):
    """FP32 trailing updates plus refinement: the pivots of the FP64 path, the
    identity block exact, and the solution columns within 1e-11 of LAPACK;
    these sizes converge without the FP64 rerun."""
    from linac_amd import _lib
    rs = numpy.random.RandomState(n)
    A = rs.random_sample((n, cols))
    R, piv, rk = gpu.rref_with_pivots(A, 0, method="mixed")
    st = _lib.refine_last()
    assert piv == list(range(n)) and rk == n and 1 <= st["passes"] <= 10
    assert st["fallback"] == 0 and st["last_correction"] <= 2.0 ** -51, st  # converged without an FP64 rerun
    assert float(numpy.max(numpy.abs(R[:, :n] - numpy.eye(n)))) <= 1e-12
    x = numpy.linalg.solve(A[:, :n], A[:, n:])
    rel = float(numpy.max(numpy.abs(R[:, n:] - x)) / numpy.max(numpy.abs(x)))
    assert rel < 1e-11, (rel, st)


def test_mixed_method_rejected_outside_float64(gpu):
    with pytest.raises(ValueError, match="float64"):
        gpu.rref_with_pivots(numpy.ones((4, 5)) + 1j, 0, method="mixed")
    with pytest.raises(ValueError, match="float64"):
        gpu.rref_with_pivots(numpy.ones((4, 5), dtype=numpy.uint32), 7, method="mixed")


def test_mixed_method_reruns_in_fp64_when_refinement_stalls(gpu):
    """cond(A) = 1e9 (singular values log-spaced 1 .. 1e-9): FP32 updates are
    far too coarse, the refinement stalls, and the call redoes the
    elimination with FP64 updates -- the result equals the refined FP64 one."""
    from linac_amd import _lib
    rs = numpy.random.RandomState(31)
    n = 600
    q1, _ = numpy.linalg.qr(rs.standard_normal((n, n)))
    q2, _ = numpy.linalg.qr(rs.standard_normal((n, n)))
    A = (q1 * numpy.logspace(0, -9, n)) @ q2.T
    Ab = numpy.concatenate([A, rs.standard_normal((n, 1))], axis=1)
    Rm, pm, rm = gpu.rref_with_pivots(Ab, 0, method="mixed")
    st = _lib.refine_last()
    Rr, pr, rr = gpu.rref_with_pivots(Ab, 0, method="blocked", refine=True)
    assert st["fallback"] == 1 and st["passes"] == 2
    assert pm == pr and rm == rr
    assert numpy.array_equal(Rm, Rr)
