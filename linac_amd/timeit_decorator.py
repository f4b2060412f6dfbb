"""Wall-clock reporting for the drop-in calls (``linac/timeit_decorator.py:8-22``).

Callers of linac see a report only when they pass ``verbose=True`` as a
keyword argument (a positional ``True`` stays silent, as in linac), printed
as ``Time elapsed in <name>: h:mm:ss:mmm.`` after a carriage return and
padded with blanks so that it overwrites a progress line.
"""
import functools
import time

_PAD = " " * 40


def _report(name, seconds):
    whole = int(seconds)
    millis = int((seconds * 1000.0) % 1000.0)
    return "\rTime elapsed in %s: %d:%02d:%02d:%03d.%s" % (
        name, whole // 3600, (whole // 60) % 60, whole % 60, millis, _PAD)


def timeit(func):
    """Time ``func``; report the wall time when called with ``verbose=True``."""
    @functools.wraps(func)
    def timed(*args, **kwargs):
        t0 = time.time()
        result = func(*args, **kwargs)
        if kwargs.get("verbose") is True:
            print(_report(func.__name__, time.time() - t0))
        return result
    return timed
