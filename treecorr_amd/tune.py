"""The autotuner of the persistent traversal.

Results never depend on these settings; only the time does.  autotune() runs a
caller's correlation under each candidate (job granularity: the number of jobs the
breadth-first split makes before the persistent launch; queue depth: the share of the
jobs a workgroup takes from the shared counter at a time) and installs the fastest.
The reference has no counterpart (OpenMP's schedule(dynamic) hands out one top-level
pair at a time, src/Corr2.cpp:158-203).
"""
import time

from . import _lib

MIN_JOBS = (1 << 22, 1 << 24, 1 << 26)
GRAB_DIV = (4, 16, 64)


def set_params(min_jobs=0, grab_div=16, levels=0):
    L = _lib.lib()
    _lib.check(L.tc_set_persist_params(levels, min_jobs))
    _lib.check(L.tc_set_persist_grab(grab_div))


def autotune(run, min_jobs=MIN_JOBS, grab_div=GRAB_DIV, sync=None, repeat=1):
    """Time run() (one correlation, or any callable that ends with its results on the
    host) for every (min_jobs, grab_div), keep the fastest and return
    (best, {(min_jobs, grab_div): seconds})."""
    times = {}
    for mj in min_jobs:
        for gd in grab_div:
            set_params(mj, gd)
            run()                       # warm: allocations and captured graphs
            if sync:
                sync()
            t0 = time.perf_counter()
            for _ in range(repeat):
                run()
            if sync:
                sync()
            times[(mj, gd)] = (time.perf_counter() - t0) / repeat
    best = min(times, key=times.get)
    set_params(*best)
    return best, times
