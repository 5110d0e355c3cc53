"""Wall-clock timing decorator (stencil_benchmarks/tools/timing.py:38-46)."""

import functools
import time


def return_time(func):
    """Wrap `func` so that it returns ``{'time': seconds}`` of its run."""

    @functools.wraps(func)
    def timed(*args, **kwargs):
        start = time.perf_counter()
        func(*args, **kwargs)
        return dict(time=time.perf_counter() - start)

    return timed
