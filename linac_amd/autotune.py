"""Autotuner for the library's run-time knobs (no counterpart in the
reference, whose kernels are specialised per shape by nvcc JIT,
``linac/pycuda_tools.py:17-49``).

Knobs (``include/linac_amd.h``, ``linac_tune_set``): the sizes from which
``method="auto"`` switches to the blocked GF(p) and float methods, whether a
panel's trailing update overlaps the next panel's inner steps
(``lookahead``), and, per size class of the matrix (largest side below 3072,
below 12288, larger: keys ``<knob>.s`` / ``.m`` / ``.l``), the GEMM shapes: the
number of tile rows the GEMMs walk together (``gemm_group``: the tile order's
L2 reuse), the inner GEMM's workgroups per 64-row stripe (``inner_split``) and
the 64-bit trailing GEMM's tile, 64 x 32 or 64 x 64 (``tile64``), and the
float64 panel kernel's rows per wave, 16 or 32 (``fpanel_rows``: half the
workgroups in the per-column candidate exchange).  ``tune()``
times the candidates on this device with matrices already in HBM, sets the
winners and saves them to a per-device JSON file; the package applies that
file when it loads the library.  GF(p) results
do not depend on the knobs (the RREF is unique); float results may change in
the last bits when a crossover moves a size to the other method.
"""
import ctypes
import json
import os
import re
import time

import numpy

from . import _lib

SIZED = ("gemm_group", "inner_split", "tile64", "fpanel_rows")  # knobs that also take per-size-class values
CLASSES = ("s", "m", "l")                        # largest side < 3072, < 12288, larger
KNOBS = ("gf_blocked_min_rows", "gf_blocked_min_cols", "float_blocked_min_rows", "float_blocked_min_cols",
         "gemm_group", "lookahead", "timeslice", "gemm_lds", "inner_split", "tile64", "fpanel_rows") + \
    tuple(f"{k}.{c}" for k in SIZED for c in CLASSES)


def get(key):
    v = ctypes.c_int64(0)
    _lib.check(_lib.load().linac_tune_get(key.encode(), ctypes.byref(v)))
    return v.value


def set_knob(key, value):
    _lib.check(_lib.load().linac_tune_set(key.encode(), int(value)))


def current():
    return {k: get(k) for k in KNOBS}


def device_name():
    buf = ctypes.create_string_buffer(256)
    _lib.check(_lib.load().linac_device_name(buf, 256))
    return buf.value.decode(errors="replace")


def tuning_file():
    """LINAC_TUNING_FILE, else ~/.cache/linac_amd/tuning-<device>.json."""
    if os.environ.get("LINAC_TUNING_FILE"):
        return os.environ["LINAC_TUNING_FILE"]
    tag = re.sub(r"[^A-Za-z0-9_.-]+", "_", device_name())
    return os.path.join(os.path.expanduser("~"), ".cache", "linac_amd", f"tuning-{tag}.json")


# Knobs that move float matrices between the plain and blocked methods, whose
# results agree only within rounding; every other knob leaves every result
# unchanged (GF(p) results are exact and unique).
FLOAT_CROSSOVERS = ("float_blocked_min_rows", "float_blocked_min_cols")


def apply(values, float_crossovers=True):
    for k, v in values.items():
        if k in KNOBS and (float_crossovers or k not in FLOAT_CROSSOVERS):
            set_knob(k, v)


def load(path=None, float_crossovers=True):
    """Apply a saved tuning (returns it), or None when there is none.
    ``float_crossovers=False`` applies only the knobs that cannot change a
    result (the package does this by itself when it loads the library, so
    that float results do not depend on a file outside the repository unless
    the caller asks for it)."""
    path = path or tuning_file()
    if not os.path.exists(path):
        return None
    with open(path) as f:
        data = json.load(f)
    apply(data.get("knobs", {}), float_crossovers)
    return data


def save(data, path=None):
    path = path or tuning_file()
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    with open(path, "w") as f:
        json.dump(data, f, indent=1)
    return path


def _timed(fn, repeat):
    import torch
    best = float("inf")
    for _ in range(repeat):
        torch.cuda.synchronize()
        t = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        best = min(best, time.perf_counter() - t)
    return best


def _gf_case(n, method, seed=0, wide=False):
    """Closure timing one in-place RREF of an n x n matrix in HBM over
    GF(2^31-1), or GF(2^61-1) in 64-bit storage (wide)."""
    import torch
    from .hip_row_reduce import rref_device
    p = 2 ** 61 - 1 if wide else 2 ** 31 - 1
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    src = torch.randint(0, p, (n, n), generator=g, device="cuda", dtype=torch.int64)
    if not wide:
        src = src.to(torch.int32)
    work = torch.empty_like(src)

    def run():
        work.copy_(src)
        rref_device(work.data_ptr(), n, n, n, p, numpy.uint64 if wide else numpy.uint32, method=method)
    return run


def _float_case(n, method, seed=0):
    import torch
    from .hip_row_reduce import rref_device
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    src = torch.rand((n, n), generator=g, device="cuda", dtype=torch.float64)
    work = torch.empty_like(src)

    def run():
        work.copy_(src)
        rref_device(work.data_ptr(), n, n, n, 0, numpy.float64, method=method)
    return run


def _crossover(case, sizes, repeat):
    """Smallest size from which blocked beats plain at it and every larger
    candidate; the timings beside it."""
    rows = []
    for n in sizes:
        tp = _timed(case(n, "plain"), repeat)
        tb = _timed(case(n, "blocked"), repeat)
        rows.append({"n": n, "plain_s": tp, "blocked_s": tb})
    choice = sizes[-1] * 2
    for i in range(len(rows) - 1, -1, -1):
        if rows[i]["blocked_s"] < rows[i]["plain_s"]:
            choice = rows[i]["n"]
        else:
            break
    return choice, rows


def tune(quick=False, save_result=True, path=None):
    """Measure the knobs on this device, set the winners and (optionally)
    save them.  Returns {"device", "knobs", "measurements"}."""
    reps = 1 if quick else 3
    meas = {}
    # tile order and look-ahead on a matrix large enough for both to matter
    n_big = 2048 if quick else 8000
    best = None
    meas["gemm"] = []
    for grp in ((8,) if quick else (4, 8, 16)):
        for la in (0, 1):
            set_knob("gemm_group", grp)
            set_knob("lookahead", la)
            t = _timed(_gf_case(n_big, "blocked"), reps)
            meas["gemm"].append({"gemm_group": grp, "lookahead": la, "n": n_big, "seconds": t})
            if best is None or t < best[0]:
                best = (t, grp, la)
    set_knob("gemm_group", best[1])
    set_knob("lookahead", best[2])
    # per size class: the GEMM tile order, the inner GEMM's workgroups per row
    # stripe and the 64-bit trailing tile width, one knob after the other
    # (coordinate descent) on a matrix of the class
    sizes = {"s": 1024, "m": 4096} if quick else {"s": 2048, "m": 8000, "l": 16000}
    sizes64 = {"s": 1024, "m": 4096} if quick else {"s": 2048, "m": 8000, "l": 12000}
    meas["by_class"] = {}
    sizes_f = {"s": 1024, "m": 4096} if quick else {"s": 2048, "m": 8000}
    for c in CLASSES:
        for k in SIZED:
            set_knob(f"{k}.{c}", 0)
    by_class = {}
    for c, n in sizes.items():
        rows = []
        run = _gf_case(n, "blocked")
        run()  # (the first call at a size grows the workspace)
        for key, cands in (("gemm_group", (4, 8, 16)), ("inner_split", (4, 8, 16))):
            times = {}
            for v in cands:
                set_knob(f"{key}.{c}", v)
                times[v] = _timed(run, reps)
                rows.append({"knob": f"{key}.{c}", "value": v, "n": n, "seconds": times[v]})
            best_v = min(times, key=times.get)
            set_knob(f"{key}.{c}", best_v)
            by_class[f"{key}.{c}"] = best_v
        run = _gf_case(sizes64[c], "blocked", wide=True)
        run()
        times = {}
        for v in (1, 2):
            set_knob(f"tile64.{c}", v)
            times[v] = _timed(run, reps)
            rows.append({"knob": f"tile64.{c}", "value": v, "n": sizes64[c], "seconds": times[v], "p": "2^61-1"})
        best_v = min(times, key=times.get)
        set_knob(f"tile64.{c}", best_v)
        by_class[f"tile64.{c}"] = best_v
        if c in sizes_f:  # float64 panel kernel: rows per wave
            run = _float_case(sizes_f[c], "blocked")
            run()
            times = {}
            for v in (16, 32):
                set_knob(f"fpanel_rows.{c}", v)
                times[v] = _timed(run, reps)
                rows.append({"knob": f"fpanel_rows.{c}", "value": v, "n": sizes_f[c], "seconds": times[v],
                             "dtype": "float64"})
            best_v = min(times, key=times.get)
            set_knob(f"fpanel_rows.{c}", best_v)
            by_class[f"fpanel_rows.{c}"] = best_v
        meas["by_class"][c] = rows
    gf_sizes = (256, 512, 1024) if quick else (64, 96, 128, 192, 256, 384, 512, 768, 1024, 1536)
    gf_n, meas["gf_crossover"] = _crossover(_gf_case, gf_sizes, reps)
    fl_sizes = (128, 256) if quick else (32, 48, 64, 96, 128, 192, 256, 384, 512)
    fl_n, meas["float_crossover"] = _crossover(_float_case, fl_sizes, reps)
    knobs = {"gemm_group": best[1], "lookahead": best[2], "gf_blocked_min_rows": gf_n, "gf_blocked_min_cols": gf_n,
             "float_blocked_min_rows": fl_n, "float_blocked_min_cols": fl_n, **by_class}
    apply(knobs)
    data = {"device": device_name(), "knobs": knobs, "measurements": meas}
    if save_result:
        data["file"] = save(data, path)
    return data
