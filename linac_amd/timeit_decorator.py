"""Wall-clock wrapper with linac's printing rule (``linac/timeit_decorator.py:8-22``).

The elapsed time is printed only when the call passes ``verbose=True`` as a
keyword argument, in the reference's h:mm:ss:mmm format.
"""
import functools
import time


def timeit(func):
    """Print the runtime of the decorated function as hours:minutes:seconds:milliseconds"""
    @functools.wraps(func)
    def wrapper_timeit(*args, **kwargs):
        start = time.time()
        value = func(*args, **kwargs)
        seconds = time.time() - start
        _, ms = divmod(seconds * 1000, 1000)
        m, s = divmod(seconds, 60)
        h, m = divmod(m, 60)
        if kwargs.get("verbose") is True:
            print("\rTime elapsed in %s: %d:%02d:%02d:%03d.                                        "
                  % (func.__name__, h, m, s, ms))
        return value
    return wrapper_timeit
