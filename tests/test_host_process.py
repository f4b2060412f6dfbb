"""Host-side behaviour of the drop-in package: linac's timing print rule and
the process model of the C ABI (CPU only; no compute reaches a GPU)."""
import ctypes
import multiprocessing
import re

import linac_amd
from linac_amd import _lib
from linac_amd.timeit_decorator import timeit


@timeit
def _work(x, verbose=False):
    return x + 1


def test_timeit_prints_only_for_keyword_verbose_true(capsys):
    """linac/timeit_decorator.py:19-20: the report appears only when the call
    passes verbose=True as a keyword; a positional True stays silent."""
    assert _work(1) == 2
    assert _work(1, verbose=False) == 2
    assert _work(1, True) == 2
    assert capsys.readouterr().out == ""
    assert _work(1, verbose=True) == 2
    out = capsys.readouterr().out
    assert re.fullmatch(r"\rTime elapsed in _work: 0:00:00:\d{3}\. {40}\n", out), repr(out)
    assert _work.__name__ == "_work"


def test_timeit_report_fields():
    """h:mm:ss:mmm from the elapsed seconds (the reference's divmod chain)."""
    from linac_amd.timeit_decorator import _report
    assert _report("f", 3723.0456).startswith("\rTime elapsed in f: 1:02:03:045.")
    assert _report("f", 59.9999).startswith("\rTime elapsed in f: 0:00:59:999.")
    assert _report("g", 0.0).startswith("\rTime elapsed in g: 0:00:00:000.")


def _child_device_count(q):
    try:
        q.put(("ok", _lib.device_count()))
    except Exception as e:  # noqa: BLE001
        q.put(("error", f"{type(e).__name__}: {e}"))


def test_spawn_after_runtime_init_works():
    """Reference tests/test_cuda_concurrency.py:9-15: a child started with
    'spawn' after the parent touched the runtime can use the library (a fork
    cannot; see test_capi)."""
    n = _lib.device_count()
    ctx = multiprocessing.get_context("spawn")
    q = ctx.Queue()
    pr = ctx.Process(target=_child_device_count, args=(q,))
    pr.start()
    pr.join(120)
    assert pr.exitcode == 0
    assert q.get(timeout=5) == ("ok", n)


def test_shutdown_is_idempotent_and_reinitialises():
    lib = _lib.load()
    assert lib.linac_shutdown() == 0
    assert lib.linac_shutdown() == 0
    a = ctypes.c_int(-1)
    assert lib.linac_device_count(ctypes.byref(a)) == 0 and a.value >= 0
    assert isinstance(linac_amd.GPU_FOUND, bool)
