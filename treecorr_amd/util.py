"""Thread-count helpers (treecorr/util.py:43-107 set_omp_threads/get_omp_threads).

The GPU engine does not use host threads for the hot path; the functions are
kept so TreeCorr code that calls them keeps working, and because the default
min_top of a field depends on the thread count (treecorr/field.py:94-104).
"""
import ctypes
import os

from . import _lib

_NUM_THREADS = None


def set_omp_threads(num_threads, logger=None):
    global _NUM_THREADS
    if num_threads is None or num_threads <= 0:
        num_threads = os.cpu_count() or 1
    _NUM_THREADS = int(num_threads)
    try:
        _lib.lib().tc_set_num_threads(_NUM_THREADS)
    except _lib.TreeCorrAMDError:
        pass
    return _NUM_THREADS


def get_omp_threads():
    if _NUM_THREADS is None:
        return os.cpu_count() or 1
    return _NUM_THREADS
