"""Continuous-batching request loop for row reductions (no counterpart in the
reference, whose GPU path solves one matrix per call,
``linac/pycuda_row_reduce.py:18-151``).

A request is one matrix already in HBM (a device tensor, as produced by the
matrix loader) with its field: a prime p for GF(p) or 0 for float64.

* GF(p) requests join a ``Batch`` between steps and leave it as soon as they
  finish: every step advances every resident matrix by one 256-column panel in
  shared launches, so requests that arrive while others are on the GPU start
  at the next step and their work overlaps on the device.  They are reduced
  in place (the request's buffer receives the RREF).
* float64 requests run through the blocked float path as resumable jobs
  (``FloatJob``, each with its own workspace): between batch steps, panel by
  panel for up to a time budget per round, the job with the least remaining
  work first, so that a long float solve holds back neither the GF(p)
  requests nor shorter float requests.  In place.
* Admission is by HBM footprint: a request is admitted only while the bytes
  of every admitted request (its matrix plus the solver's workspace for it)
  stay under the cap; otherwise it waits, first come first served.

``Server.pump`` performs one round (admit, float panels for up to
``float_budget_s``, one batch step, collect); ``serve`` drives a list of timed requests through it
against the wall clock and reports per-request latency.
"""
import collections
import time
from dataclasses import dataclass, field

import numpy

from .batch import Batch, footprint
from .hip_float import FloatJob, float_job_workspace
from .hip_row_reduce import rref_device

# Eight 31-bit primes: numpy.random.RandomState(2024), x = randint(2**30, 2**31),
# next prime after x (all below 2**31).
SERVING_PRIMES = [1104465403, 1261898177, 1451761303, 1482257969, 1688125709, 1881848357, 1928907403, 2099153543]
# Primes that need 64-bit storage (2^32 <= p < 2^63): the Mersenne prime
# 2^61 - 1 and the largest primes below 2^62 and 2^63.
SERVING_PRIMES_64 = [2 ** 61 - 1, 2 ** 62 - 57, 2 ** 63 - 25]
WIDE_P_MIN = 2 ** 32


@dataclass
class Request:
    rid: int
    rows: int
    cols: int
    p: int                 # prime, or 0 for float64 (p >= 2^32: uint64 storage)
    arrival: float = 0.0   # seconds after the start of the run
    tensor: object = None  # device tensor (int32 bits of uint32 / int64 for GF, float64), set when materialised


@dataclass
class Result:
    rid: int
    pivots: list
    rank: int
    arrival: float
    finished: float
    tensor: object = field(default=None, repr=False)

    @property
    def latency(self):
        return self.finished - self.arrival


def float_footprint(rows, cols):
    """Device bytes of a float64 request beyond its matrix: its job's
    workspace and the finishing copy (the RREF is gathered into a scratch
    matrix before it is written back)."""
    return float_job_workspace(rows, cols) + rows * cols * 8


class Server:
    """GF(p) continuous batch plus interleaved float64 solves under one HBM
    cap.  ``float_budget_s``: wall time per round given to the float jobs (at
    least one panel per round).  ``float_age_s``: once a float request has
    waited this long its jobs get four times the budget (off by default: at
    200 requests/s it moved the overall p99 from 0.11 to 0.20-0.27 s, the GF(p)
    requests paying more than the float ones gained)."""

    def __init__(self, mem_cap_bytes=200 * 10 ** 9, float_budget_s=0.0015, float_age_s=None):
        self.cap = int(mem_cap_bytes)
        self.float_budget_s = float(float_budget_s)
        self.float_age_s = float_age_s       # a float request older than this gets 4 x the budget per round
        self.last_step_s = 0.0
        self.fjobs = []                      # [request, bytes, FloatJob] of the open float solves
        self.batch = Batch(self.cap)
        self.in_use = 0
        self.waiting = collections.deque()   # arrived, not admitted (FIFO)
        self.active = {}                     # batch id -> (request, bytes)
        self.max_in_use = 0

    def close(self):
        for _, _, job in self.fjobs:
            job.close()
        self.fjobs = []
        self.batch.close()

    def submit(self, req):
        """Queue a materialised request; MemoryError if it could never fit the cap."""
        if self._bytes(req) > self.cap:
            raise MemoryError(f"request {req.rid} needs {self._bytes(req)} bytes, above the cap of {self.cap}")
        self.waiting.append(req)

    def idle(self):
        return not (self.waiting or self.active or self.fjobs)

    def _bytes(self, req):
        if req.p == 0:
            return req.rows * req.cols * 8 + float_footprint(req.rows, req.cols)
        w = 8 if req.p >= WIDE_P_MIN else 4
        return req.rows * req.cols * w + footprint(req.rows, req.cols, req.p) - \
            (req.rows * ((req.cols + 63) // 64 * 64) * w + 255) // 256 * 256

    def _admit(self):
        while self.waiting:
            req = self.waiting[0]
            need = self._bytes(req)
            if self.in_use + need > self.cap:
                break  # wait for memory to free up (first come, first served)
            if req.p == 0:
                import torch
                ws = torch.empty(float_job_workspace(req.rows, req.cols), dtype=torch.uint8, device=req.tensor.device)
                job = FloatJob(req.tensor.data_ptr(), req.rows, req.cols, req.cols, numpy.float64, workspace=ws)
                self.fjobs.append([req, need, job])
            else:
                bid = self.batch.add_device(req.tensor.data_ptr(), req.rows, req.cols, req.cols, req.p,
                                            in_place=True)
                self.active[bid] = (req, need)
            self.in_use += need
            self.max_in_use = max(self.max_in_use, self.in_use)
            self.waiting.popleft()

    def pump(self, clock):
        """One round; returns the requests finished in it."""
        done = []
        self._admit()
        if self.fjobs:
            budget = self.float_budget_s
            if self.float_age_s is not None:
                now = clock()
                if any(now - r.arrival > self.float_age_s for r, _, _ in self.fjobs):
                    budget *= 4
            t_end = time.perf_counter() + budget
            while self.fjobs:
                ent = min(self.fjobs, key=_remaining_work)
                req, need, job = ent
                if job.step(1) == 0:
                    piv, rank = job.finish()
                    done.append(Result(req.rid, piv, rank, req.arrival, clock(), req.tensor))
                    self.in_use -= need
                    self.fjobs.remove(ent)
                if time.perf_counter() >= t_end:
                    break
        if self.active:
            t = time.perf_counter()
            self.batch.step()
            self.last_step_s = time.perf_counter() - t
            fin = [b for b in self.active if self.batch.done(b)]
            taken = self.batch.take_many(fin)
            now = clock()
            for bid, (_, piv, rank) in zip(fin, taken):
                req, need = self.active.pop(bid)
                done.append(Result(req.rid, piv, rank, req.arrival, now, req.tensor))
                self.in_use -= need
        return done


def _remaining_work(ent):
    """Work left in a float job, in units of rows x columns^2 (the trailing
    updates of its remaining panels)."""
    req, _, job = ent
    panels = (req.cols + 63) // 64 if job.remaining is None else job.remaining
    return req.rows * min(req.cols, 64 * panels) ** 2


def make_requests(n, rate, seed=0, n_min=200, n_max=8000, float_share=0.1, primes=SERVING_PRIMES):
    """Poisson arrivals at ``rate`` per second; N log-uniform in [n_min, n_max],
    cols = round(N * U[0.8, 1.2]); GF(p) with p drawn from ``primes``, or
    float64 with probability ``float_share``.  Matrices are materialised later
    (``materialise``)."""
    rs = numpy.random.RandomState(seed)
    t = numpy.cumsum(rs.exponential(1.0 / rate, size=n)) if rate else numpy.zeros(n)
    reqs = []
    for i in range(n):
        rows = int(round(numpy.exp(rs.uniform(numpy.log(n_min), numpy.log(n_max)))))
        cols = max(1, int(round(rows * rs.uniform(0.8, 1.2))))
        p = 0 if rs.uniform() < float_share else int(primes[rs.randint(len(primes))])
        reqs.append(Request(i, rows, cols, p, float(t[i])))
    return reqs


def materialise(req, seed=0):
    """The request's matrix in HBM: uniform [0, p) (GF; int32 bits of uint32,
    or int64 for p >= 2^32) or [0, 1) (float64), generated on the device from
    (seed, rid)."""
    import torch
    g = torch.Generator(device="cuda")
    g.manual_seed(seed * 1000003 + req.rid)
    if req.p == 0:
        req.tensor = torch.rand((req.rows, req.cols), generator=g, device="cuda", dtype=torch.float64)
    else:
        t = torch.randint(0, req.p, (req.rows, req.cols), generator=g, device="cuda", dtype=torch.int64)
        req.tensor = t if req.p >= WIDE_P_MIN else t.to(torch.int32)
    return req


def serve(requests, mem_cap_bytes=200 * 10 ** 9, seed=0, keep=False, float_budget_s=0.0015, float_age_s=None):
    """Run timed requests through a Server against the wall clock.

    Each request is materialised in HBM when it arrives (not before), then
    submitted.  Returns (results sorted by rid, wall seconds, server stats)."""
    import torch
    reqs = sorted(requests, key=lambda r: r.arrival)
    srv = Server(mem_cap_bytes, float_budget_s, float_age_s)
    results = []
    t0 = time.perf_counter()

    def clock():
        return time.perf_counter() - t0

    i = 0
    try:
        while len(results) < len(reqs):
            now = clock()
            while i < len(reqs) and reqs[i].arrival <= now:
                materialise(reqs[i], seed)
                torch.cuda.current_stream().synchronize()  # the library works on its own stream
                srv.submit(reqs[i])
                i += 1
            if srv.idle():
                if i < len(reqs):
                    time.sleep(max(0.0, min(reqs[i].arrival - clock(), 0.01)))
                continue
            for r in srv.pump(clock):
                if not keep:
                    r.tensor = None  # results do not keep the buffers
                results.append(r)
        wall = clock()
    finally:
        stats = {"max_in_use_bytes": srv.max_in_use, "cap_bytes": srv.cap}
        srv.close()
    if not keep:
        for r in reqs:
            r.tensor = None
    results.sort(key=lambda r: r.rid)
    return results, wall, stats


def sequential_plain(requests, seed=0, method="plain"):
    """The same requests solved one after another by single calls (each matrix
    materialised in HBM first, outside the timing); returns seconds per
    request."""
    import torch
    out = []
    for req in requests:
        materialise(req, seed)
        torch.cuda.synchronize()
        t = time.perf_counter()
        if req.p == 0:
            rref_device(req.tensor.data_ptr(), req.rows, req.cols, req.cols, 0, numpy.float64, method=method)
        else:
            _rref_gf_unaligned(req, method)
        torch.cuda.synchronize()
        out.append(time.perf_counter() - t)
        req.tensor = None
    return out


def _rref_gf_unaligned(req, method):
    """Single-call GF(p) solve of a device matrix whose row stride need not be
    16-byte aligned: copied into an aligned buffer (the device entry point's
    requirement), solved there, copied back."""
    import torch
    wide = req.p >= WIDE_P_MIN
    per = 2 if wide else 4  # elements per 16 bytes
    dt = numpy.uint64 if wide else numpy.uint32
    ld = (req.cols + per - 1) // per * per
    if ld == req.cols:
        rref_device(req.tensor.data_ptr(), req.rows, req.cols, req.cols, req.p, dt, method=method)
        return
    buf = torch.zeros((req.rows, ld), dtype=req.tensor.dtype, device="cuda")
    buf[:, :req.cols].copy_(req.tensor)
    torch.cuda.synchronize()
    rref_device(buf.data_ptr(), req.rows, req.cols, ld, req.p, dt, method=method)
    req.tensor.copy_(buf[:, :req.cols])


SIZE_CLASSES = [(200, 500), (500, 1000), (1000, 2000), (2000, 4000), (4000, 8001)]


def latency_summary(results, wall, requests=None):
    """Throughput and p50/p99 latency; with the requests, also per size class
    (rows) for GF(p) and for float64 requests."""
    lat = numpy.array([r.latency for r in results])
    out = {"requests": len(results), "wall_s": round(wall, 3), "solves_per_s": round(len(results) / wall, 2),
           "p50_latency_s": round(float(numpy.percentile(lat, 50)), 4),
           "p99_latency_s": round(float(numpy.percentile(lat, 99)), 4)}
    if requests is not None:
        by_id = {r.rid: r for r in requests}
        classes = {}
        for lo, hi in SIZE_CLASSES:
            for kind in ("gf", "f64"):
                sel = [res.latency for res in results
                       if lo <= by_id[res.rid].rows < hi and (by_id[res.rid].p == 0) == (kind == "f64")]
                if sel:
                    classes[f"{kind} N in [{lo}, {hi})"] = {
                        "n": len(sel), "p50_s": round(float(numpy.percentile(sel, 50)), 4),
                        "p99_s": round(float(numpy.percentile(sel, 99)), 4)}
        out["by_class"] = classes
    return out
