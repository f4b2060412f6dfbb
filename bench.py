"""Benchmark: RREF of a 20000 x 20000 matrix over GF(2^31-1) on one MI355X.

One step = one in-place RREF (``linac_rref_u32`` through the C ABI) of a
uniform random full-rank uint32 matrix already resident in HBM, preceded by a
device-to-device restore of the pristine input (0.4 ms, inside the timed
region).  Prints ONE JSON line.

``value`` is seconds per RREF (lower is better).  ``roofline`` describes the
dominant kernel of the headline method, measured live with HIP events recorded
by the library on the launch stream around sampled launches.  ``cpu_baseline``
is the numpy restatement of linac's CPU ``row_reduce`` (the oracle) timed on
this host on a bounded sample and scaled by N^3.
"""
import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "RREF seconds at N=20000 GF(2^31-1), bit-exact; batched solves/sec + p99 latency"
A100_SECONDS_N20000 = 23.40  # reference README.md:133


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=20000)
    ap.add_argument("--method", default="auto", choices=["auto", "plain", "blocked"])
    ap.add_argument("--no-plain", action="store_true", help="skip the plain-kernel comparator")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-sizes", default="1000,2000", help="uniform sizes of the CPU baseline protocol")
    ap.add_argument("--no-host", action="store_true", help="skip the host-inclusive drop-in call")
    ap.add_argument("--no-sweep", action="store_true", help="skip the batched loader sweep (64 x 3000^2)")
    ap.add_argument("--sweep-systems", type=int, default=64)
    ap.add_argument("--no-serve", action="store_true", help="skip the serving run (2000 mixed requests)")
    ap.add_argument("--serve-requests", type=int, default=2000)
    ap.add_argument("--serve-rate", type=float, default=200.0, help="Poisson arrival rate (requests/s)")
    ap.add_argument("--no-tune", action="store_true", help="skip the autotuner (library defaults)")
    ap.add_argument("--configs", action="store_true",
                    help="instead of the headline line: one JSON line measuring every GF(p) and float configuration")
    ap.add_argument("--with-100k", action="store_true", help="with --configs: include the 100000^2 u32 run (~40 GB)")
    ap.add_argument("--only", default="", help="with --configs: comma list of cases, e.g. u32-20000-blocked,f64-8000")
    ap.add_argument("--cpu-once", type=int, default=0, metavar="N",
                    help="instead of the headline line: one timed run of the CPU baseline (numpy restatement of "
                         "linac.row_reduce, 1 core) on the uniform N x N matrix, with progress lines")
    ap.add_argument("--gemm", type=int, default=0, metavar="N",
                    help="instead of the headline line: the trailing-update kernel alone, C (N x N) -= A (N x 512) "
                         "B (512 x N) mod 2^31-1 (linac_gemm_sub_u32), --steps timed calls")
    return ap.parse_args()


def cpu_once_main(args):
    """One run of the CPU baseline at a size the default protocol skips
    (minutes per run): RandomState(N).randint(2^31-1) as in the protocol, no
    warm-up, a progress line every 512 columns."""
    import numpy
    from oracle import rref
    p, n = 2 ** 31 - 1, args.cpu_once
    A = numpy.random.RandomState(n).randint(p, size=(n, n)).astype(numpy.int64)
    t0 = time.perf_counter()

    def progress(j):
        print(f"cpu baseline {n}x{n}: column {j} at {time.perf_counter() - t0:.1f} s", flush=True)
    R = rref.rref_modp(A, p, progress=progress)
    dt = time.perf_counter() - t0
    ok = bool(numpy.array_equal(numpy.diagonal(R), numpy.ones(n, dtype=numpy.int64)))
    print(json.dumps({"cpu_once": {"n": n, "seconds": round(dt, 2), "runs": 1, "cores": 1, "identity": ok,
                                   "kind": "port", "host_cpu": _cpu_model(), "os_cpu_count": os.cpu_count(),
                                   "numpy": numpy.__version__}}), flush=True)


def gemm_main(args):
    """The modular GEMM of the trailing update in isolation: one launch of
    the int8-limb kernel over N x N outputs at depth 512 (the depth of a
    two-panel update), after the operands' limb planes are formed."""
    import torch
    from linac_amd.incremental import gemm_sub
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import certificates
    p, n, k = 2 ** 31 - 1, args.gemm, 512
    C = certificates.uniform_u32(n, n, p, 1)
    A = certificates.uniform_u32(n, k, p, 2)
    B = certificates.uniform_u32(k, n, p, 3)
    for _ in range(args.warmup):
        gemm_sub(C, A, B, p, k=k)
    torch.cuda.synchronize()
    times = []
    for _ in range(args.steps):
        t = time.perf_counter()
        gemm_sub(C, A, B, p, k=k)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t)
    ops = 2.0 * n * n * k * 16  # int8 operations: 16 limb products per modular MAC
    best = min(times)
    print(json.dumps({"gemm": {"n": n, "k": k, "p": p, "calls": args.steps, "best_ms": round(best * 1e3, 3),
                               "median_ms": round(sorted(times)[len(times) // 2] * 1e3, 3),
                               "int8_TOPs_incl_planes": round(ops / best / 1e12, 1),
                               "note": "host clock around the whole call (limb planes + kernel); the kernel's own "
                                       "time comes from rocprofv3"}}), flush=True)


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _stats(times):
    import numpy
    t = numpy.array(times)
    return {"runs": len(times), "median_s": round(float(numpy.median(t)), 4), "mean_s": round(float(t.mean()), 4),
            "std_s": round(float(t.std()), 4)}


def cpu_baseline(n_target, sizes=(1000, 2000)):
    """linac's CPU row_reduce as the numpy restatement (prime path,
    pivoting=1, scaling=False, int64 mod p), on this host, one core.
    Protocol: config C0 (500x500 of rank 400, seed 0) and uniform N x N
    matrices mod 2^31-1 (RandomState(N).randint); one warm-up, then three
    timed runs each (the largest size without its own warm-up).  `value` is
    the largest size's median scaled by (n_target / N)^3: an extrapolation,
    labelled as such; N = 5000 was measured once separately."""
    import numpy
    from oracle import modp, rref
    p = 2 ** 31 - 1
    rs = numpy.random.RandomState(0)
    L = rs.randint(p, size=(500, 400), dtype=numpy.uint32)
    U = rs.randint(p, size=(400, 500), dtype=numpy.uint32)
    C0 = modp.matmul_modp(L, U, p)
    cases = [("C0 500x500 rank 400", C0, True)]
    for n in sizes:
        cases.append((f"{n}x{n}", numpy.random.RandomState(n).randint(p, size=(n, n)).astype(numpy.int64),
                      n != sizes[-1]))
    out = {}
    for name, A, warm in cases:
        if warm:
            rref.rref_modp(A, p)
        times = []
        for _ in range(3):
            t0 = time.perf_counter()
            R = rref.rref_modp(A, p)
            times.append(time.perf_counter() - t0)
        assert R[0, 0] == 1
        out[name] = _stats(times)
    n = sizes[-1]
    med = out[f"{n}x{n}"]["median_s"]
    return {"value": round(med * (n_target / n) ** 3, 1), "unit": "s", "cores": 1, "kind": "port",
            "sample": f"numpy restatement of linac.row_reduce (prime path, pivoting=1, scaling=False), 1 core; value = "
                      f"median of the {n}x{n} runs ({med:.2f} s) x ({n_target}/{n})^3, an N^3 extrapolation to "
                      f"{n_target} (not a measurement at that size)",
            "protocol": out, "host_cpu": _cpu_model(), "os_cpu_count": os.cpu_count(),
            "numpy": numpy.__version__}


def host_inclusive(n, p=2 ** 31 - 1, seed=0):
    """The drop-in call as linac's users make it: cuda_row_reduce(numpy array,
    p) at N x N, host array in and out (conversion, H2D, RREF, D2H); the
    reference's published N = 20000 time (README.md:133) is of this kind."""
    import numpy
    import linac_amd
    A = numpy.random.RandomState(seed).randint(p, size=(n, n), dtype=numpy.uint32)
    linac_amd.cuda_row_reduce(A[:2048, :2048], p)  # warm-up (library, small workspace)
    times = []
    for _ in range(2):
        t0 = time.perf_counter()
        R = linac_amd.cuda_row_reduce(A, p)
        times.append(time.perf_counter() - t0)
    ok = bool(numpy.array_equal(R.diagonal(), numpy.ones(n, dtype=numpy.uint32))) and int(numpy.count_nonzero(R)) == n
    del R, A
    return {"n": n, "seconds": round(min(times), 4), "runs": [round(t, 4) for t in times], "correct": ok}


def sweep_bench(n_systems, p=2 ** 31 - 1):
    """64 systems of 3000 x 3000 (3000 monomials of degree <= 6 in 8 variables
    times a prefactor, 3000 sample points each) built by one loader launch in
    HBM and row-reduced in place by one batch; beside it, the same systems
    through single plain and single blocked calls (a sample of them)."""
    import numpy
    import torch
    import linac_amd
    from linac_amd import sweep
    bases, lind, _ = sweep.sweep_problem(n_systems, 8, 6, 3000, 3000, p, 0)
    sweep.reduce_sweep(bases, lind, p)  # an untimed sweep of the same size: the batch arena grows to it
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res = sweep.reduce_sweep(bases, lind, p)
    torch.cuda.synchronize()
    total = time.perf_counter() - t0
    m = res.matrices
    load_s, reduce_s = res.load_seconds, res.reduce_seconds
    ok = all(r == 3000 for r in res.ranks) and bool(m.diagonal(dim1=1, dim2=2).eq(1).all().item()) \
        and int(m.count_nonzero().item()) == n_systems * 3000
    del res, m
    torch.cuda.empty_cache()
    # sequential single-matrix calls on a sample of the systems (same loader output)
    seq = {}
    sample = 2
    for method in ("plain", "blocked"):
        mats = torch.empty((sample, 3000, 3000), dtype=torch.int32, device="cuda")
        for s in range(sample):
            A = linac_amd.cuda_load_matrix(bases[s], lind, field_characteristic=p)
            mats[s].copy_(torch.from_numpy(A.view(numpy.int32)))
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for s in range(sample):
            linac_amd.rref_device(mats[s].data_ptr(), 3000, 3000, 3000, p, numpy.uint32, method=method)
        torch.cuda.synchronize()
        seq[method] = (time.perf_counter() - t1) / sample
        del mats
    macs = n_systems * 3000 * 3000 * 3001 / 2
    return {
        "workload": f"{n_systems} systems 3000x3000 u32 GF(2^31-1): loader (8 vars, 3000 monomials of degree <= 6, "
                    "prefactor) into HBM + batched in-place RREF, after one untimed sweep of the same size",
        "correct": ok, "seconds": round(total, 4), "loader_ms": round(1e3 * load_s, 3),
        "reduce_s": round(reduce_s, 4),
        "solves_per_s": round(n_systems / total, 2),
        "modmac_per_s": float(f"{macs / total:.4g}"),
        "sequential_plain_s_per_solve": round(seq["plain"], 4),
        "sequential_blocked_s_per_solve": round(seq["blocked"], 4),
        "speedup_vs_sequential_plain": round(seq["plain"] * n_systems / total, 2),
    }


def serve_bench(n, rate):
    """Serving: n requests with Poisson arrivals at `rate`/s, N log-uniform in
    [200, 8000], cols N * U[0.8, 1.2], p from 8 seeded 31-bit primes, 10%
    float64 uniform [0, 1); each matrix generated in HBM at its arrival;
    admission by HBM footprint under 200 GB.  Saturation throughput (all
    arrivals at t = 0) beside it, and sequential single plain calls on an
    evenly spaced sample of the same requests as the comparator."""
    import numpy
    from linac_amd import serving
    serving.serve(serving.make_requests(8, 0, seed=99, n_min=200, n_max=2000), seed=99)  # warm-up
    reqs = serving.make_requests(n, rate, seed=0)
    res, wall, st = serving.serve(reqs, seed=0)
    timed = serving.latency_summary(res, wall, reqs)
    sat_reqs = serving.make_requests(n, 0, seed=0)
    res_s, wall_s, _ = serving.serve(sat_reqs, seed=0)
    sat = serving.latency_summary(res_s, wall_s, sat_reqs)
    step = max(1, n // 50)
    sample = [serving.Request(r.rid, r.rows, r.cols, r.p) for r in sat_reqs[::step]]
    t_plain = serving.sequential_plain(sample, seed=0, method="plain")
    t_auto = serving.sequential_plain(sample, seed=0, method="auto")
    plain_rate = 1.0 / float(numpy.mean(t_plain))
    return {
        "workload": f"{n} requests, Poisson {rate}/s, N log-uniform [200, 8000], cols N*U[0.8,1.2], "
                    "8 seeded 31-bit primes, 10% float64, matrices generated in HBM at arrival, 200 GB cap",
        "at_rate": timed,
        "saturated": sat,
        "max_admitted_bytes": st["max_in_use_bytes"],
        "sequential_plain_solves_per_s": round(plain_rate, 2),
        "sequential_single_auto_solves_per_s": round(1.0 / float(numpy.mean(t_auto)), 2),
        "sequential_sample": f"every {step}th request ({len(sample)}), same matrices",
        "speedup_vs_sequential_plain": round(sat["solves_per_s"] / plain_rate, 2),
    }


def _limb_matmul_mod(A, B, p, chunk=4096):
    """(A @ B) mod p for int64 device tensors in [0, p), p < 2^32: float64
    products of 16-bit limbs (exact: partial sums below 2^53), reduced in int64."""
    import torch
    b0, b1 = (B & 0xFFFF).double(), (B >> 16).double()
    out = torch.empty((A.shape[0], B.shape[1]), dtype=torch.int64, device=A.device)
    for r0 in range(0, A.shape[0], chunk):
        a = A[r0:r0 + chunk]
        a0, a1 = (a & 0xFFFF).double(), (a >> 16).double()
        q = [(x @ y).long() % p for x, y in ((a0, b0), (a0, b1), (a1, b0), (a1, b1))]
        hi = (((q[3] << 16) % p) << 16) % p
        out[r0:r0 + chunk] = (q[0] + (((q[1] + q[2]) % p) << 16) % p + hi) % p
    return out


def _gf_device_case(n, p, rank_frac, method, seed=0, dtype="u32"):
    """Seconds of an in-place device RREF of an n x n matrix: uniform in
    [0, p) (rank n w.h.p.), or of rank r = round(rank_frac * n) as the seeded
    product L U mod p (L n x r, U r x n uniform).  Checked: rank and pivots,
    the [I | X; 0] shape, and for L U Freivalds' U v == U[:, :r] (R v) mod p
    (which pins R, since RREF(L U) = RREF(U))."""
    import numpy
    import torch
    import linac_amd
    tdt = torch.int32 if dtype == "u32" else torch.int64
    r = n if rank_frac >= 1 else int(round(rank_frac * n))
    npdt = numpy.uint32 if dtype == "u32" else numpy.uint64

    def rand(rows, cols, g):
        out = torch.empty((rows, cols), dtype=tdt, device="cuda")
        chunk = max(1, (1 << 28) // cols)
        for r0 in range(0, rows, chunk):  # int64 draws in slices
            out[r0:r0 + chunk] = torch.randint(0, p, (min(chunk, rows - r0), cols), generator=g, device="cuda",
                                               dtype=torch.int64).to(tdt)
        return out

    def make():
        g = torch.Generator(device="cuda")
        g.manual_seed(seed)
        if r == n:
            return rand(n, n, g), None
        L = rand(n, r, g).long() & 0xFFFFFFFF
        U = rand(r, n, g).long() & 0xFFFFFFFF
        A = _limb_matmul_mod(L, U, p).to(tdt)
        del L
        return A, U

    A, U = make()
    if method != "plain":  # an untimed solve first: the blocked workspace grows to this size
        linac_amd.rref_device(A.data_ptr(), n, n, n, p, npdt, method=method)
        A, U = make()
    torch.cuda.synchronize()
    t = time.perf_counter()
    piv, rank = linac_amd.rref_device(A.data_ptr(), n, n, n, p, npdt, method=method)
    torch.cuda.synchronize()
    sec = time.perf_counter() - t
    ok = rank == r and piv == list(range(r))
    if ok:
        top = A[:r, :r]
        ok = bool(top.diagonal().eq(1).all().item()) and int(top.count_nonzero().item()) == r and \
            int(A[r:].count_nonzero().item()) == 0
    if ok and U is not None:
        g = torch.Generator(device="cuda")
        g.manual_seed(seed + 99)
        v = torch.randint(0, p, (n, 1), generator=g, device="cuda", dtype=torch.int64)
        w = _limb_matmul_mod(A[:r].long() & 0xFFFFFFFF, v, p)
        ok = bool(torch.equal(_limb_matmul_mod(U, v, p), _limb_matmul_mod(U[:, :r], w, p)))
    del A, U
    torch.cuda.empty_cache()
    return {"n": n, "p": p, "rank": rank, "method": method, "seconds": round(sec, 4), "correct": ok,
            "input": "uniform" if r == n else f"L U mod p, L {n}x{r}, U {r}x{n} uniform (seed {seed})"}


def _float_case(kind, n, seed=0):
    """[A | b] on the device through the blocked float path (unrefined and
    refined) and the plain one; error of the solution column against the
    exact solution (numpy's LAPACK solve refined with long-double residuals),
    plus max |R[:, :n] - I|."""
    import numpy
    import scipy.linalg
    import torch
    import linac_amd
    from linac_amd import _lib
    rs = numpy.random.RandomState(seed)
    if kind == "f64":
        A = rs.random_sample((n, n + 1))
    elif kind == "c128":
        A = rs.standard_normal((n, n + 1)) + 1j * rs.standard_normal((n, n + 1))
    else:  # ill-conditioned: U diag(logspace(0, -8)) V^T, U and V Haar (QR of Gaussians, sign-fixed)
        def haar(m):
            q, r = numpy.linalg.qr(rs.standard_normal((m, m)))
            return q * numpy.sign(numpy.diag(r))
        A = numpy.empty((n, n + 1))
        A[:, :n] = (haar(n) * numpy.logspace(0, -8, n)) @ haar(n).T
        A[:, n] = rs.random_sample(n)
    lu = scipy.linalg.lu_factor(A[:, :n])
    x = scipy.linalg.lu_solve(lu, A[:, n])
    ext = numpy.clongdouble if kind == "c128" else numpy.longdouble
    Al, bl = A[:, :n].astype(ext), A[:, n].astype(ext)
    for _ in range(3):
        x = x + scipy.linalg.lu_solve(lu, (bl - Al @ x.astype(ext)).astype(A.dtype))
    del Al, bl
    dtn = numpy.complex128 if kind == "c128" else numpy.float64
    d = torch.from_numpy(A).to("cuda", dtype=torch.complex128 if kind == "c128" else torch.float64)
    torch.cuda.synchronize()
    out = {}
    for method, refine in (("blocked", False), ("blocked", True), ("plain", False)):
        w = d.clone()
        if method == "blocked":  # an untimed solve first: the blocked workspace grows to this size
            linac_amd.rref_device(w.data_ptr(), n, n + 1, n + 1, 0, dtn, method=method, refine=refine)
            w.copy_(d)
        torch.cuda.synchronize()
        t = time.perf_counter()
        piv, rank = linac_amd.rref_device(w.data_ptr(), n, n + 1, n + 1, 0, dtn, method=method, refine=refine)
        torch.cuda.synchronize()
        sec = time.perf_counter() - t
        R = w.cpu().numpy()
        err = float(numpy.max(numpy.abs(R[:, n] - x)) / numpy.max(numpy.abs(x)))
        key = method + ("_refined" if refine else "")
        out[key] = {"seconds": round(sec, 4), "rank": rank, "rel_err_vs_exact": float(f"{err:.3g}"),
                    "max_abs_R_minus_I": float(f"{float(numpy.max(numpy.abs(R[:, :n] - numpy.eye(n)))):.3g}")}
        if refine:
            st = _lib.refine_last()
            out[key]["correction_first_last"] = [float(f"{st['first_correction']:.3g}"),
                                                 float(f"{st['last_correction']:.3g}")]
        del w
    del d
    torch.cuda.empty_cache()
    return {"kind": kind, "n": n, "cond_2": 1e8 if kind == "illcond" else None, **out}


def _coo_case(n=4000, density=0.05, seed=0):
    """A sparse system from disk: a plain-text COO file of an n x n matrix
    (density 5%, 10% of the values fractions a/b) loaded into HBM over
    GF(2^31-1) by device_matrix_from_coo (host parse and reduction, device
    scatter), against the host route of the reference's format (dense object
    matrix from the file, reduced entrywise, copied to the device)."""
    import tempfile
    import numpy
    import torch
    from linac_amd import sparse_matrix_tools as smt
    p = 2 ** 31 - 1
    rs = numpy.random.RandomState(seed)
    nnz = int(n * n * density)
    flat = rs.choice(n * n, size=nnz, replace=False)
    flat.sort()
    ri, ci = flat // n, flat % n
    vals = rs.randint(-10 ** 6, 10 ** 6, size=nnz)
    vals[vals == 0] = 1
    frac = rs.random_sample(nnz) < 0.1
    den = rs.randint(2, 1000, size=nnz)
    with tempfile.TemporaryDirectory() as d:
        path = f"{d}/system.txt"
        with open(path, "w") as f:
            f.write("".join(f"{r} {c} {v}/{q}\n" if fr else f"{r} {c} {v}\n"
                            for r, c, v, q, fr in zip(ri.tolist(), ci.tolist(), vals.tolist(), den.tolist(),
                                                      frac.tolist())))
        size_mb = os.path.getsize(path) / 1e6
        torch.cuda.synchronize()
        t = time.perf_counter()
        dev = smt.device_matrix_from_coo(path, p, shape=(n, n))
        torch.cuda.synchronize()
        t_dev = time.perf_counter() - t
        # host route: dense object matrix, entrywise residues, upload
        t = time.perf_counter()
        M = smt.matrix_from_plain_txt_coo(path)
        nzr, nzc = numpy.nonzero(M != 0)
        H = numpy.zeros(M.shape, dtype=numpy.uint32)
        H[nzr, nzc] = smt.residues_mod_p(list(M[nzr, nzc]), p).astype(numpy.uint32)
        hd = torch.from_numpy(H.view(numpy.int32)).cuda()
        torch.cuda.synchronize()
        t_host = time.perf_counter() - t
        same = bool(torch.equal(hd, dev))
    del dev, hd
    torch.cuda.empty_cache()
    return {"workload": f"{n}x{n} GF(2^31-1), {nnz} entries ({size_mb:.1f} MB plain-text COO, 10% fractions)",
            "device_route_s": round(t_dev, 3), "entries_per_s": float(f"{nnz / t_dev:.3g}"),
            "host_dense_route_s": round(t_host, 3), "speedup": round(t_host / t_dev, 2), "identical": same}


def _growing_case(dim=1000, cols=2000, start=20, step=20, max_iteration=50, seed=0):
    """The row-growing loop of a function-space basis search
    (VectorSpaceOfFunctions.__get_pivots__): sample rows of a hidden
    dim-dimensional space over GF(2^31-1) are added `step` at a time until
    one is dependent.  The reference's pattern reduces the whole stacked
    matrix again after every addition (here with this library's device RREF);
    the resident path (linac_amd.incremental) keeps the RREF in HBM and
    eliminates only the new rows.  Sample rows are generated on the device."""
    import numpy
    import torch
    from linac_amd import incremental, rref_device
    p = 2 ** 31 - 1
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    n_rows = start + step * max_iteration
    V = torch.randint(0, p, (dim, cols), generator=g, device="cuda", dtype=torch.int64).to(torch.int32)
    Cf = torch.randint(0, p, (n_rows, dim), generator=g, device="cuda", dtype=torch.int64).to(torch.int32)
    S = torch.zeros((n_rows, cols), dtype=torch.int32, device="cuda")
    incremental.gemm_sub(S, Cf, V, p, k=dim)  # S = -Cf V: rows spanning the hidden space
    torch.cuda.synchronize()
    ld = (cols + 3) // 4 * 4
    work = torch.zeros((n_rows, ld), dtype=torch.int32, device="cuda")

    def reference_pattern():
        n, it = start, 0
        while True:
            work[:n, :cols] = S[:n]
            torch.cuda.synchronize()
            piv, rank = rref_device(work.data_ptr(), n, cols, ld, p, numpy.uint32)
            if rank < n:
                return piv, work[:rank, :cols].clone(), it
            if it >= max_iteration:
                raise RuntimeError("dimension exceeded")
            n += step
            it += 1

    def resident():
        return incremental.pivots_of_growing_matrix(lambda lo, hi: S[lo:hi], cols, p, start, step, max_iteration)

    reference_pattern()  # warm-up (kernels, scratch)
    torch.cuda.synchronize()
    t = time.perf_counter()
    piv_r, R_r, iters = reference_pattern()
    torch.cuda.synchronize()
    t_ref = time.perf_counter() - t
    resident()
    t = time.perf_counter()
    R_i, piv_i = resident()
    torch.cuda.synchronize()
    t_inc = time.perf_counter() - t
    same = piv_i == piv_r and numpy.array_equal(R_i, R_r.cpu().numpy().view(numpy.uint32))
    del V, Cf, S, work
    torch.cuda.empty_cache()
    return {"workload": f"hidden space of dimension {dim} in {cols} columns over GF(2^31-1), rows added {step} at a "
                        f"time from {start} ({iters + 1} additions) until one is dependent",
            "reference_pattern_s": round(t_ref, 4), "resident_s": round(t_inc, 4),
            "speedup": round(t_ref / t_inc, 2), "rank": len(piv_i), "identical": bool(same)}


def configs_main(args):
    """Every configuration of the GF(p) and float forms, device-resident."""
    import torch
    from linac_amd import _lib
    torch.cuda.set_device(0)
    _lib.check(_lib.load().linac_init(0))
    p31, p61 = 2 ** 31 - 1, 2 ** 61 - 1
    only = set(x for x in args.only.split(",") if x)
    gf_cases = [(f"u32-{n}-{m}", (n, p31, 1, m, 0, "u32")) for n in (5000, 10000, 20000) for m in ("blocked", "plain")]
    gf_cases += [("u32-20000-rank0.9-blocked", (20000, p31, 0.9, "blocked", 0, "u32")),
                 ("u32-20000-rank0.9-plain", (20000, p31, 0.9, "plain", 0, "u32")),
                 ("u64-10000-blocked", (10000, p61, 1, "blocked", 0, "u64")),
                 ("u64-10000-plain", (10000, p61, 1, "plain", 0, "u64"))]
    if args.with_100k:
        gf_cases.append(("u32-100000-blocked", (100000, p31, 1, "blocked", 0, "u32")))
    fl_cases = [("f64-8000", ("f64", 8000)), ("c128-6000", ("c128", 6000)), ("illcond-4000", ("illcond", 4000))]
    unknown = only - {k for k, _ in gf_cases + fl_cases} - {"coo", "growing"}
    if unknown:
        print(f"unknown cases: {sorted(unknown)}", file=sys.stderr)
        sys.exit(2)
    _gf_device_case(2000, p31, 1, "auto")  # warm-up
    gf = [_gf_device_case(*a) for k, a in gf_cases if not only or k in only]
    fl = [_float_case(*a) for k, a in fl_cases if not only or k in only]
    out = {"gf": gf, "float": fl}
    if not only or "coo" in only:
        out["coo"] = _coo_case()
    if not only or "growing" in only:
        out["growing"] = _growing_case()
    print(json.dumps({"configs": out, "host": platform.node()}), flush=True)


def main():
    args = parse()
    if args.cpu_once:
        cpu_once_main(args)
        return
    if args.gemm:
        if args.gpus != 1:
            print("bench.py measures one GPU only", file=sys.stderr)
            sys.exit(2)
        gemm_main(args)
        return
    if args.configs:
        if args.gpus != 1:
            print("bench.py measures one GPU only", file=sys.stderr)
            sys.exit(2)
        configs_main(args)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(independent solves shard across GPUs without exchange; one solve's row stripes would "
              "need a broadcast of each panel's pivot rows)", file=sys.stderr)
        sys.exit(2)
    import numpy
    import torch
    import linac_amd
    from linac_amd import _lib

    N = args.n
    p = 2 ** 31 - 1
    torch.cuda.set_device(0)
    _lib.check(_lib.load().linac_init(0))
    tuning = None
    if not args.no_tune:
        from linac_amd import autotune
        t0 = time.perf_counter()
        tuned = autotune.tune(save_result=False)
        tuning = {"knobs": tuned["knobs"], "seconds": round(time.perf_counter() - t0, 2),
                  "gemm_candidates": [{k: (round(v, 5) if isinstance(v, float) else v) for k, v in m.items()}
                                      for m in tuned["measurements"]["gemm"]]}
    g = torch.Generator(device="cuda")
    g.manual_seed(0)
    src = torch.randint(0, p, (N, N), generator=g, device="cuda", dtype=torch.int64).to(torch.int32)
    work = torch.empty_like(src)
    stream = torch.cuda.current_stream().cuda_stream
    torch.cuda.synchronize()

    def step(method, profile=False):
        work.copy_(src)
        return linac_amd.rref_device(work.data_ptr(), N, N, N, p, numpy.uint32, method=method, stream=stream,
                                     profile=profile)

    method = args.method
    for _ in range(args.warmup):
        step(method)
    # every timed step reduces its own copy of the input, made before the timed
    # region (the RREF overwrites its matrix; the inputs are resident in HBM)
    inputs = [src.clone() for _ in range(max(1, args.steps))]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        piv, rank = linac_amd.rref_device(inputs[k].data_ptr(), N, N, N, p, numpy.uint32, method=method,
                                          stream=stream)
    torch.cuda.synchronize()
    sec = (time.perf_counter() - t0) / max(1, args.steps)
    work.copy_(inputs[max(0, args.steps - 1)])
    del inputs
    ok = rank == N and bool(work.diagonal().eq(1).all().item()) and int(work.count_nonzero().item()) == N
    if not ok:
        print(json.dumps({"error": "RREF of the full-rank input is not the identity", "rank": rank}))
        sys.exit(1)

    # one profiled run of the headline method (events around sampled launches)
    step(method, profile=True)
    prof = _lib.profile_last()
    used = {1: "plain", 2: "blocked"}.get(int(prof.get("method", 0)), "?")
    secs_s = prof["sampled_ms"] / 1e3
    if used == "plain":
        achieved = prof["sampled_bytes"] / secs_s / 1e9
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": 8000.0, "unit": "GB/s",
                    "frac": round(achieved / 8000.0, 4), "traffic": None,
                    "kernel": "k_elim (row update of one pivot step)",
                    "per_launch": "2*4*(N-1)*(N-j) bytes at column j", "samples": int(prof["samples"])}
    else:
        limb_products = 16
        achieved = 2.0 * limb_products * prof["sampled_macs"] / secs_s / 1e12
        roofline = {"bound": "mfma", "achieved": round(achieved, 1), "peak": 5000.0, "unit": "TFLOP/s",
                    "frac": round(achieved / 5000.0, 4), "traffic": None,
                    "kernel": "k_gemm_outer_lds<0>: mod-p GEMM (int8 limb MFMA, 16 limb products per modular MAC)",
                    "per_launch": "(rows - k)*k*cols modular MACs, rows = the rows the launch updates (from the "
                                  "panel's row offset on, forward elimination), k = depth (pivots of one panel or of "
                                  "two composed), cols = its columns; every launch of the kernel in the solve: the "
                                  "trailing updates and the full-rank fast path's Z = -P M^-1",
                    "samples": int(prof["samples"])}

    roofline["avg_launch_us"] = round(1e3 * prof["sampled_ms"] / max(1, prof["samples"]), 2)
    if used != "plain":  # C read and written plus both operands' limb planes once, per launch, from the schedule
        roofline["traffic_algorithmic"] = round(prof["sampled_bytes"] / max(1, prof["samples"]))
    # HBM traffic per launch of the same kernel from the committed rocprofv3 PMC summary
    # (FETCH_SIZE / WRITE_SIZE, calibrated against a device copy of known size)
    tfile = os.path.join(ROOT, "profiles", f"rref_{used}_u32_N{N}_traffic.json")
    if os.path.exists(tfile):
        t = json.load(open(tfile))
        roofline["traffic"] = round(t["traffic_bytes_per_launch"])
        roofline.setdefault("traffic_algorithmic", round(t["algorithmic_bytes_per_launch"]))
        roofline["traffic_source"] = os.path.relpath(tfile, ROOT)
    comparators = {"device_ms_profiled_run": round(prof["device_ms"], 2)}
    host = None
    if not args.no_host:
        host = host_inclusive(N)
        comparators["host_inclusive_drop_in"] = host
    if method != "plain" and not args.no_plain and used != "plain":
        t1 = time.perf_counter()
        step("plain")
        torch.cuda.synchronize()
        plain_s = time.perf_counter() - t1
        comparators["plain_seconds"] = round(plain_s, 4)
        comparators["speedup_vs_plain"] = round(plain_s / sec, 2)
    del src, work
    torch.cuda.empty_cache()

    out = {
        "metric": METRIC,
        "value": round(sec, 5),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(sec * 1e3, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": round(host["seconds"] / A100_SECONDS_N20000, 5) if (N == 20000 and host) else None,
        "vs_baseline_note": "host-inclusive cuda_row_reduce(numpy, p) seconds at N=20000 (comparators) / 23.40 s, "
                            "the reference's A100 time of the same call (README.md:133); below 1 is faster. "
                            "`value` itself is device-resident.",
        "dtype": "u32",
        "data": "synthetic",
        "config": {"workload": f"rref_gf_u32_N{N}", "rows": N, "cols": N, "p": p, "method": used,
                   "input": "uniform [0,p) from torch.randint on cuda:0, seed 0; full rank, result checked == I"},
        "roofline": roofline,
        "comparators": comparators,
        "tuning": tuning,
        "host": platform.node(),
    }
    if not args.no_sweep:
        out["batched"] = sweep_bench(args.sweep_systems)
    if not args.no_serve:
        out["serving"] = serve_bench(args.serve_requests, args.serve_rate)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(N, tuple(int(x) for x in args.cpu_sizes.split(",")))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
