"""tc_kmeans_run_noalloc (static device scratch, no allocation) against tc_kmeans_run.

The allocation-free entry point takes device-resident columns and centres; its results
must equal those of tc_kmeans_run on the same data (the same kernels, the centre update
done on the device with the host's arithmetic and order), on two calls in a row, the
second with fewer objects and centres.  The sums reach the centres through FP64
atomics in both, so the comparison allows rounding at the 1e-12 level of the field
size; the iteration counts are equal."""
import ctypes

import numpy as np
import pytest


class _Hip:
    """Device buffers for the test (the entry point itself allocates nothing)."""

    def __init__(self):
        self.L = ctypes.CDLL('libamdhip64.so')
        self.L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                     ctypes.c_int]
        self.L.hipFree.argtypes = [ctypes.c_void_p]
        self.bufs = []

    def put(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        p = ctypes.c_void_p()
        assert self.L.hipMalloc(ctypes.byref(p), a.nbytes) == 0
        assert self.L.hipMemcpy(p, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, 1) == 0
        self.bufs.append(p)
        return p

    def get(self, p, n):
        out = np.empty(n)
        assert self.L.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), p, out.nbytes, 2) == 0
        return out

    def free(self):
        for p in self.bufs:
            self.L.hipFree(p)
        self.bufs = []


def _case(rng, n, npatch, dim):
    pts = rng.uniform(0., 100., (n, dim))
    # clumps, so that the patches are uneven
    k = rng.integers(0, 40, n // 3)
    pts[:n // 3] = rng.uniform(0., 100., (40, dim))[k] + rng.normal(0., 2., (n // 3, dim))
    w = rng.uniform(0.5, 1.5, n)
    cen = pts[rng.choice(n, npatch, replace=False)].copy()
    return pts, w, cen


@pytest.mark.gpu
def test_kmeans_noalloc_matches_kmeans_run(gpu):
    from treecorr_amd import _lib
    L = _lib.lib()
    hip = _Hip()
    rng = np.random.default_rng(2024)
    try:
        for n, npatch, dim in ((200_000, 300, 3), (30_000, 37, 2)):
            print('no-alloc k-means case n', n, 'npatch', npatch, 'dim', dim, flush=True)
            pts, w, cen = _case(rng, n, npatch, dim)
            cols = [np.ascontiguousarray(pts[:, q]) for q in range(dim)]
            size = 100.
            ref = cen.copy().ravel()
            it_ref = ctypes.c_int()
            _lib.check(L.tc_kmeans_run(_lib.dptr(ref), npatch, dim, _lib.dptr(cols[0]),
                                       _lib.dptr(cols[1]),
                                       _lib.dptr(cols[2]) if dim == 3 else None, _lib.dptr(w),
                                       n, 12, 1e-9, size, 0, ctypes.byref(it_ref)))
            dcen = hip.put(cen.ravel())
            dcols = [hip.put(c) for c in cols]
            dw = hip.put(w)
            it = ctypes.c_int()
            _lib.check(L.tc_kmeans_run_noalloc(dcen, npatch, dim, dcols[0], dcols[1],
                                               dcols[2] if dim == 3 else None, dw, n, 12, 1e-9,
                                               size, 0, ctypes.byref(it)))
            got = hip.get(dcen, npatch * dim)
            assert it.value == it_ref.value
            assert np.max(np.abs(got - ref)) <= 1e-12 * size
        # beyond the static scratch: a clear error, nothing run
        big = 8193
        dcen = hip.put(np.zeros(big * 2))
        dx = hip.put(np.zeros(10))
        rc = L.tc_kmeans_run_noalloc(dcen, big, 2, dx, dx, None, dx, 10, 1, 0., 1., 0, None)
        assert rc != 0
        assert b'8192' in L.tc_last_error()
    finally:
        hip.free()
