"""ctypes binding of liblinac_amd.so, the C ABI declared in include/linac_amd.h.

The shared library is built in-tree by ``linac_amd.build`` (hipcc, gfx950).
There is no JIT and no CPU fallback: if the library is missing, every compute
call raises ``RuntimeError`` naming the file it looked for.
"""
import atexit
import ctypes
import importlib.util
import os
import threading

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblinac_amd.so")

OK = 0
ERR_INVALID = -1
ERR_NOMEM = -2
ERR_DEVICE = -3
ERR_NOT_INVERTIBLE = -4
ERR_FORKED = -5

DEVICE_PTR = 0x1
IN_PLACE = 0x400
METHOD_AUTO = 0x00
METHOD_PLAIN = 0x10
METHOD_BLOCKED = 0x20
NO_FINALIZE = 0x100
PROFILE = 0x200
REFINE = 0x800

METHOD_MIXED = 0x30  # float64: FP32 trailing updates + refinement
METHODS = {"auto": METHOD_AUTO, "plain": METHOD_PLAIN, "blocked": METHOD_BLOCKED, "mixed": METHOD_MIXED}

_lock = threading.Lock()
_lib = None
_load_error = None
_load_pid = os.getpid()

_i64 = ctypes.c_int64
_u32 = ctypes.c_uint32
_u64 = ctypes.c_uint64
_vp = ctypes.c_void_p
_i64p = ctypes.POINTER(ctypes.c_int64)

# name -> (restype, argtypes); every entry point declared in include/linac_amd.h
SIGNATURES = {
    "linac_tune_set": (ctypes.c_int, [ctypes.c_char_p, _i64]),
    "linac_tune_get": (ctypes.c_int, [ctypes.c_char_p, _i64p]),
    "linac_device_name": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int]),
    "linac_last_error": (ctypes.c_char_p, []),
    "linac_version": (ctypes.c_int, []),
    "linac_device_count": (ctypes.c_int, [ctypes.POINTER(ctypes.c_int)]),
    "linac_available": (ctypes.c_int, []),
    "linac_init": (ctypes.c_int, [ctypes.c_int]),
    "linac_release_cache": (ctypes.c_int, []),
    "linac_shutdown": (ctypes.c_int, []),
    "linac_is_prime": (ctypes.c_int, [_u64]),
    "linac_profile_last": (ctypes.c_int, [ctypes.POINTER(ctypes.c_double), ctypes.c_int]),
    "linac_refine_last": (ctypes.c_int, [ctypes.POINTER(ctypes.c_double), ctypes.c_int]),
    "linac_rref_u32": (ctypes.c_int, [_vp, _i64, _i64, _i64, _u32, _u32, _vp, _i64p, _i64p]),
    "linac_rref_u64": (ctypes.c_int, [_vp, _i64, _i64, _i64, _u64, _u32, _vp, _i64p, _i64p]),
    "linac_rref_u32_noalloc": (ctypes.c_int, [_vp, _i64, _i64, _i64, _u32, _u32, _vp, _i64p, _i64p]),
    "linac_rref_f64": (ctypes.c_int, [_vp, _i64, _i64, _i64, ctypes.c_double, _u32, _vp, _i64p, _i64p]),
    "linac_rref_c128": (ctypes.c_int, [_vp, _i64, _i64, _i64, ctypes.c_double, _u32, _vp, _i64p, _i64p]),
    "linac_batch_footprint_u32": (ctypes.c_int, [_i64, _i64, ctypes.POINTER(_u64)]),
    "linac_batch_create": (ctypes.c_int, [ctypes.POINTER(_vp), _u64]),
    "linac_batch_destroy": (ctypes.c_int, [_vp]),
    "linac_batch_memory": (ctypes.c_int, [_vp, ctypes.POINTER(_u64), ctypes.POINTER(_u64)]),
    "linac_batch_add_u32": (ctypes.c_int, [_vp, _vp, _i64, _i64, _i64, _u32, _u32, _i64p]),
    "linac_batch_footprint_u64": (ctypes.c_int, [_i64, _i64, ctypes.POINTER(_u64)]),
    "linac_batch_add_u64": (ctypes.c_int, [_vp, _vp, _i64, _i64, _i64, _u64, _u32, _i64p]),
    "linac_batch_take_u64": (ctypes.c_int, [_vp, _i64, _vp, _i64, _u32, _i64p, _i64p]),
    "linac_batch_step": (ctypes.c_int, [_vp, _i64p]),
    "linac_batch_run": (ctypes.c_int, [_vp]),
    "linac_batch_status": (ctypes.c_int, [_vp, _i64]),
    "linac_batch_take_u32": (ctypes.c_int, [_vp, _i64, _vp, _i64, _u32, _i64p, _i64p]),
    "linac_batch_take_many_u32": (ctypes.c_int, [_vp, ctypes.POINTER(_i64), _i64, ctypes.POINTER(_vp), ctypes.POINTER(_i64),
                                                  _u32, ctypes.POINTER(_vp), ctypes.POINTER(_i64)]),
    "linac_batch_take_many_u64": (ctypes.c_int, [_vp, ctypes.POINTER(_i64), _i64, ctypes.POINTER(_vp), ctypes.POINTER(_i64),
                                                  _u32, ctypes.POINTER(_vp), ctypes.POINTER(_i64)]),
    "linac_rref_batch_u32": (ctypes.c_int, [_vp, _i64p, _i64p, _i64p, ctypes.POINTER(_u32), _i64, _u32, _vp, _i64p]),
    "linac_float_job_workspace": (ctypes.c_int, [_i64, _i64, ctypes.c_int, _i64p]),
    "linac_float_job_create": (ctypes.c_int, [ctypes.POINTER(_vp), _vp, _i64, _i64, _i64, ctypes.c_int, ctypes.c_double,
                                              _u32, _vp, _i64]),
    "linac_float_job_step": (ctypes.c_int, [_vp, _i64, _i64p]),
    "linac_float_job_finish": (ctypes.c_int, [_vp, _i64p, _i64p]),
    "linac_float_job_destroy": (ctypes.c_int, [_vp]),
    "linac_load_matrix_u32": (ctypes.c_int, [_vp, _vp, _vp, _i64, _i64, _i64, _i64, _u32, _u32, _vp]),
    "linac_load_matrix_u64": (ctypes.c_int, [_vp, _vp, _vp, _i64, _i64, _i64, _i64, _u64, _u32, _vp]),
    "linac_load_matrix_f64": (ctypes.c_int, [_vp, _vp, _vp, _i64, _i64, _i64, _i64, _u32, _vp]),
    "linac_load_matrix_c128": (ctypes.c_int, [_vp, _vp, _vp, _i64, _i64, _i64, _i64, _u32, _vp]),
    "linac_gemm_sub_u32": (ctypes.c_int, [_vp, _i64, _vp, _i64, _i64, _vp, _vp, _i64, _i64, _vp, _i64, _i64, _i64,
                                          _u32, _u32]),
    "linac_coo_scatter_u32": (ctypes.c_int, [_vp, _i64, _i64, _i64, _vp, _vp, _vp, _i64, _u32]),
    "linac_coo_scatter_u64": (ctypes.c_int, [_vp, _i64, _i64, _i64, _vp, _vp, _vp, _i64, _u32]),
}


class LibraryMissing(RuntimeError):
    pass


def _preload_torch_hip_runtime():
    """Make torch's bundled HIP runtime the process's HIP runtime.

    torch ships its own libamdhip64/libhsa-runtime64 with the same sonames as
    /opt/rocm's, so whichever is loaded first serves the whole process.  torch
    does not work on a runtime other than its own (it reports no GPU), while
    this library runs on either; loading torch's first (without importing
    torch) keeps both usable in one process whatever the import order."""
    spec = importlib.util.find_spec("torch")
    if spec is None or not spec.origin:
        return
    path = os.path.join(os.path.dirname(spec.origin), "lib", "libamdhip64.so")
    if os.path.exists(path):
        ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)


def load():
    """Load (once) and return the ctypes library; raise if it is not built."""
    global _lib, _load_error
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise LibraryMissing(f"{LIB_PATH} is not built; run linac_amd.build.build() (hipcc, gfx950)")
        _preload_torch_hip_runtime()
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name, None)
            if fn is None:
                continue
            fn.restype = res
            fn.argtypes = args
        _lib = lib
        # release streams, pinned memory and device buffers while torch and
        # the HIP runtime are still up (the library's own exit hook would run
        # after the interpreter is gone)
        atexit.register(_shutdown_at_exit)
    # a saved tuning for this device (linac_amd.autotune) applies from the
    # start, except the float crossovers (they can change float results in the
    # last bits; autotune.load() applies them on request)
    try:
        from . import autotune
        if lib.linac_available():
            autotune.load(float_crossovers=False)
    except (OSError, ValueError, RuntimeError):
        pass
    return lib


def _shutdown_at_exit():
    lib = _lib
    if lib is not None and os.getpid() == _load_pid:
        lib.linac_shutdown()


def last_error():
    msg = load().linac_last_error()
    return msg.decode(errors="replace") if msg else ""


def check(rc):
    """Map a C status to the exception the Python API promises."""
    if rc == OK:
        return
    msg = last_error()
    if rc == ERR_INVALID:
        raise ValueError(msg)
    if rc == ERR_NOMEM:
        raise MemoryError(msg)
    if rc == ERR_NOT_INVERTIBLE:
        raise ZeroDivisionError(msg)
    raise RuntimeError(msg or f"linac_amd error {rc}")


def profile_last():
    """Timings of this thread's last LINAC_PROFILE call (see linac_profile_last)."""
    buf = (ctypes.c_double * 6)()
    n = load().linac_profile_last(buf, 6)
    keys = ["device_ms", "samples", "sampled_ms", "sampled_bytes", "sampled_macs", "method"]
    return dict(zip(keys[:n], list(buf)[:n]))


def refine_last():
    """Statistics of this thread's last LINAC_REFINE call (see linac_refine_last)."""
    buf = (ctypes.c_double * 5)()
    n = load().linac_refine_last(buf, 5)
    keys = ["passes", "first_correction", "last_correction", "solution_columns", "fallback"]
    return dict(zip(keys[:n], list(buf)[:n]))


def device_count():
    n = ctypes.c_int(0)
    check(load().linac_device_count(ctypes.byref(n)))
    return n.value


def available():
    try:
        return bool(load().linac_available())
    except (OSError, LibraryMissing):
        return False
