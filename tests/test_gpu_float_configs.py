"""The float configurations at their stated sizes against the FP64 oracle.

Inputs are [A | b] systems (tests/certificates.float_system, seed 0):
f64 8000 uniform [0, 1), complex128 6000 with N(0, 1) parts, and an
ill-conditioned f64 4000 (singular values log-spaced from 1 to 1e-8).  The
fixture (tests/golden/make_config_fixtures.py) holds, per system, the C
oracle's FP64 RREF solution column (linac's float algorithm) and the exact
solution to extended accuracy (the oracle's answer refined with long-double
residuals); the ill-conditioned input depends on the host's LAPACK in its last
bits, so its oracle runs in the test.  The FP64 oracle is itself cond(A) * 2^-53-accurate at best, so:

* unrefined results (blocked and plain) must agree with the FP64 oracle to
  within 16 cond(A) 2^-53, the size of the error both carry;
* refined results (LINAC_REFINE) must be within 1e-10 of the exact solution
  and, for the well-conditioned systems, within 1e-10 of the FP64 oracle too;
* the mixed method (FP32 trailing updates + refinement, redone with FP64
  updates when the refinement does not converge) on f64 8000 within 1e-10 of
  the exact solution.
Errors are max |x - x_ref| / max |x_ref| over the solution column."""
import hashlib
import os

import numpy
import pytest

import certificates as cert
from conftest import GOLDEN

pytestmark = pytest.mark.gpu

U = 2.0 ** -53
CASES = [("f64", 8000), ("c128", 6000), ("illcond", 4000)]


@pytest.fixture(scope="module")
def fixture():
    return numpy.load(os.path.join(GOLDEN, "float_configs.npz"), allow_pickle=False)


def _rel(x, ref):
    return float(numpy.max(numpy.abs(x - ref)) / numpy.max(numpy.abs(ref)))


def _solve(Ab, kind, method, refine):
    import torch
    import linac_amd
    n = Ab.shape[0]
    dt = numpy.complex128 if kind == "c128" else numpy.float64
    d = torch.from_numpy(numpy.ascontiguousarray(Ab)).to("cuda")
    torch.cuda.synchronize()
    piv, rank = linac_amd.rref_device(d.data_ptr(), n, n + 1, n + 1, 0, dt, method=method, refine=refine)
    torch.cuda.synchronize()
    R = d.cpu().numpy()
    return R, piv, rank


@pytest.mark.parametrize("kind, n", CASES)
def test_float_config_against_fp64_oracle(gpu, fixture, kind, n):
    from linac_amd import _lib
    Ab = cert.float_system(kind, n, seed=0)
    if kind == "illcond":
        # built with LAPACK QR and BLAS products, whose last bits depend on the
        # host CPU: the oracle runs here (4000^2 takes seconds), cond = 1e8 by
        # construction
        from oracle import crref
        Ro, po = crref.rref_f64(Ab)
        assert po == list(range(n))
        x_or = Ro[:, n].copy()
        del Ro
        A, b = Ab[:, :n], Ab[:, n]
        x_star = cert.refine_solution(A, b, numpy.linalg.solve(A, b))
        cond = 1e8
    else:  # pure RandomState draws: the same bits everywhere
        sha = hashlib.sha256(numpy.ascontiguousarray(Ab).tobytes()).hexdigest()
        assert sha == str(fixture[f"{kind}_input_sha256"])
        x_or, x_star = fixture[f"{kind}_x_oracle"], fixture[f"{kind}_x_star"]
        cond = float(fixture[f"{kind}_cond"])
    bound = max(1e-12, 16 * cond * U)
    report = {}
    for method in ("blocked", "plain"):
        R, piv, rank = _solve(Ab, kind, method, False)
        assert rank == n and piv == list(range(n))
        assert float(numpy.max(numpy.abs(R[:, :n] - numpy.eye(n)))) <= 1e-12
        report[method] = (_rel(R[:, n], x_or), _rel(R[:, n], x_star))
        assert report[method][0] <= bound, (kind, method, report[method], bound)
    R, piv, rank = _solve(Ab, kind, "blocked", True)
    stats = _lib.refine_last()
    assert rank == n and piv == list(range(n)) and stats["passes"] == 2 and stats["solution_columns"] == 1
    report["refined"] = (_rel(R[:, n], x_or), _rel(R[:, n], x_star))
    print(kind, n, "cond", f"{cond:.3g}", {k: tuple(f"{e:.3g}" for e in v) for k, v in report.items()},
          "correction", f"{stats['first_correction']:.3g}", f"{stats['last_correction']:.3g}")
    assert report["refined"][1] <= 1e-10, report
    if 16 * cond * U <= 1e-10:
        assert report["refined"][0] <= 1e-10, report
    if kind == "f64":
        # the opt-in mixed method: FP32 trailing updates, then refinement;
        # cond of this system (4.8e8) is beyond what FP32 updates can refine
        # from, so the call redoes the elimination with FP64 updates
        R, piv, rank = _solve(Ab, kind, "mixed", False)
        ms = _lib.refine_last()
        report["mixed"] = (_rel(R[:, n], x_or), _rel(R[:, n], x_star))
        print(kind, n, "mixed", tuple(f"{e:.3g}" for e in report["mixed"]), "passes", ms["passes"],
              "correction", f"{ms['first_correction']:.3g}", f"{ms['last_correction']:.3g}",
              "fp64 rerun", ms["fallback"])
        assert rank == n and piv == list(range(n)) and 1 <= ms["passes"] <= 10
        assert float(numpy.max(numpy.abs(R[:, :n] - numpy.eye(n)))) <= 1e-12
        assert report["mixed"][1] <= 1e-10, report


@pytest.mark.parametrize("complex_", [False, True])
def test_refinement_many_solution_columns(gpu, complex_):
    """A wide system (1500 x 1700, full row rank): 200 non-pivot columns go
    through the tiled double-double residual and the replay; the refined
    result matches the extended-accuracy solution to 1e-12."""
    from linac_amd import _lib
    import linac_amd
    rs = numpy.random.RandomState(17)
    n, m = 1500, 1700
    A = rs.standard_normal((n, m))
    if complex_:
        A = A + 1j * rs.standard_normal((n, m))
    ref = cert.refine_solution(A[:, :n], A[:, n:], numpy.linalg.solve(A[:, :n], A[:, n:]))
    R0, p0, r0 = linac_amd.rref_with_pivots(A, 0, method="blocked")
    R1, p1, r1 = linac_amd.rref_with_pivots(A, 0, method="blocked", refine=True)
    st = _lib.refine_last()
    assert p0 == p1 == list(range(n)) and r0 == r1 == n and st["solution_columns"] == m - n
    e0, e1 = _rel(R0[:, n:], ref), _rel(R1[:, n:], ref)
    assert e1 <= 1e-12 and e1 <= e0, (e0, e1)
    assert float(numpy.max(numpy.abs(R1[:, :n] - numpy.eye(n)))) <= 1e-12


def test_refinement_rank_deficient_keeps_pivots(gpu):
    """Rank 300 in 500 x 640: refinement changes only the pivot rows'
    non-pivot columns; the zero rows stay zero and the pivots stay put."""
    import linac_amd
    rs = numpy.random.RandomState(23)
    A = rs.standard_normal((500, 300)) @ rs.standard_normal((300, 640))
    R0, p0, r0 = linac_amd.rref_with_pivots(A, 0, method="blocked")
    R1, p1, r1 = linac_amd.rref_with_pivots(A, 0, method="blocked", refine=True)
    assert p0 == p1 and r0 == r1 == 300
    assert not numpy.any(R1[300:])
    # R[:r] is A[Q, P]^-1 A[Q, :]; its exact value from the pivot block
    ref = cert.refine_solution(A[:300, p1], A[:300], numpy.linalg.solve(A[:300, p1], A[:300]))
    assert _rel(R1[:300], ref) <= max(1e-9, _rel(R0[:300], ref))
    assert _rel(R1[:300], ref) <= 1e-9
