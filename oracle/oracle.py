"""TEST INFRASTRUCTURE ONLY -- Python access to the CPU oracle (oracle/libtc_oracle.so).

The oracle restates the TreeCorr 5.1.3 C++ engine (tree build, 2-point
traversal, 3-point multipole) on the CPU.  tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg are the only users; the product package never
imports this module.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

NDATA, KDATA, GDATA = 0, 1, 4
FLAT, THREED, SPHERE = 1, 2, 3
MIDDLE, MEDIAN, MEAN, RANDOM = 0, 1, 2, 3

_dp = ctypes.POINTER(ctypes.c_double)


def _ptr(a):
    if a is None:
        return None
    return a.ctypes.data_as(_dp)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, 'libtc_oracle.so')
        if not os.path.exists(path):
            raise RuntimeError('oracle library missing: build it with make -f oracle/Makefile')
        L = ctypes.CDLL(path)
        vp = ctypes.c_void_p
        L.tco_field_new.restype = vp
        L.tco_field_new.argtypes = [ctypes.c_int, ctypes.c_int] + [_dp] * 7 + [
            ctypes.c_long, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int,
            ctypes.c_int, ctypes.c_int]
        L.tco_field_free.argtypes = [vp]
        L.tco_set_periodic.argtypes = [ctypes.c_double] * 3
        L.tco_set_arc.argtypes = [ctypes.c_int]
        L.tco_set_metric3.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                      ctypes.c_double]
        L.tco_field_nnodes.argtypes = [vp]
        L.tco_field_nnodes.restype = ctypes.c_long
        L.tco_field_ntop.argtypes = [vp]
        L.tco_field_ntop.restype = ctypes.c_long
        L.tco_field_sizesq.argtypes = [vp]
        L.tco_field_sizesq.restype = ctypes.c_double
        L.tco_field_center.argtypes = [vp, _dp]
        L.tco_field_dump.argtypes = [vp] + [ctypes.c_void_p] * 9
        L.tco_corr2.argtypes = [vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int] + [
            ctypes.c_double] * 5 + [ctypes.c_int] * 3 + [_dp] * 8 + [
            ctypes.POINTER(ctypes.c_longlong)]
        L.tco_corr2_sampled.argtypes = L.tco_corr2.argtypes + [ctypes.c_int]
        L.tco_corr2_brute.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_long] + [_dp] * 6 + [
            ctypes.c_long] + [_dp] * 6 + [ctypes.c_int] + [ctypes.c_double] * 3 + [
            ctypes.c_int, ctypes.c_int] + [_dp] * 8
        if hasattr(L, 'tco_corr3_multipole'):
            L.tco_corr3_multipole.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                              ctypes.c_int] + [ctypes.c_double] * 5 + [
                ctypes.c_int] * 3 + [_dp] * 17 + [ctypes.POINTER(ctypes.c_longlong)]
        if hasattr(L, 'tco_corr3_multipole_x'):
            L.tco_corr3_multipole_x.argtypes = [vp, vp, vp] + [ctypes.c_int] * 5 + [
                ctypes.c_double] * 5 + [ctypes.c_int] + [_dp] * 15 + [
                ctypes.POINTER(ctypes.c_longlong)]
        _LIB = L
    return _LIB


def _c(a):
    return None if a is None else np.ascontiguousarray(a, dtype=float)


class Field:
    """A tree built by the oracle (the reference's Field + BuildCells)."""

    def __init__(self, d, coords, x, y, z=None, k=None, g1=None, g2=None, w=None, wpos=None,
                 min_size=0., max_size=np.inf, split_method=MEAN, brute=False, min_top=3,
                 max_top=10):
        x, y, z = _c(x), _c(y), _c(z)
        n = len(x)
        w = np.ones(n) if w is None else _c(w)
        wpos = _c(wpos)
        d1 = _c(k) if d == KDATA else _c(g1)
        d2 = _c(g2)
        self._keep = (x, y, z, d1, d2, w, wpos)
        self.d, self.coords, self.n = d, coords, n
        self.h = lib().tco_field_new(d, coords, _ptr(x), _ptr(y), _ptr(z), _ptr(d1), _ptr(d2),
                                     _ptr(w), _ptr(wpos), n, float(min_size), float(max_size),
                                     int(split_method), int(bool(brute)), int(min_top),
                                     int(max_top))

    def __del__(self):
        if getattr(self, 'h', None):
            lib().tco_field_free(self.h)
            self.h = None

    @property
    def ntop(self):
        return lib().tco_field_ntop(self.h)

    @property
    def sizesq(self):
        return lib().tco_field_sizesq(self.h)

    def dump(self):
        L = lib()
        nn = L.tco_field_nnodes(self.h)
        nt = L.tco_field_ntop(self.h)
        out = dict(pos=np.zeros((nn, 3)), size=np.zeros(nn, np.float32),
                   w=np.zeros(nn, np.float32), wk=np.zeros(nn, np.float32),
                   g=np.zeros((nn, 2), np.float32), n=np.zeros(nn, np.int64),
                   left=np.zeros(nn, np.int32), right=np.zeros(nn, np.int32),
                   top=np.zeros(nt, np.int32))
        L.tco_field_dump(self.h, *[out[k].ctypes.data for k in
                                   ('pos', 'size', 'w', 'wk', 'g', 'n', 'left', 'right', 'top')])
        return out


# codes of the mixed cross correlations in tco_corr2_sampled (tc_oracle.h Corr2Kind)
MIXED = {(0, 1): 8, (0, 4): 9, (1, 4): 10, (0, 2): 11, (1, 2): 12, (0, 3): 13, (0, 5): 14,
         (0, 6): 15, (1, 3): 16, (1, 5): 17, (1, 6): 18}


def corr2(d, coords, f1, f2, nbins, min_sep, max_sep, bin_size, b, a, quick=False, exact=True,
          nthreads=0, row_stride=1, d2=None, period=None, arc=False, metric3=0, rpar=None):
    """Run the oracle's 2-point traversal; returns dict of arrays and counters.

    d2 (the second field's type) differs from d for NK, NG and KG (cross only).
    row_stride > 1 processes only every row_stride-th top-level row (timing samples).
    period (xp, yp, zp) selects the Periodic metric, arc=True the Arc metric, metric3 the
    Rperp (2), Rlens (3) or OldRperp (5) metric; rpar (min, max) the r_par range."""
    out = {k: np.zeros(nbins) for k in
           ('xi0', 'xi1', 'xi2', 'xi3', 'meanr', 'meanlogr', 'weight', 'npairs')}
    cnt = (ctypes.c_longlong * 4)()
    if d2 is not None and d2 != d:
        if f2 is None:
            raise ValueError('mixed correlations are cross correlations')
        d = MIXED[(d, d2)]
    lib().tco_set_periodic(*(period if period is not None else (0., 0., 0.)))
    lib().tco_set_arc(int(bool(arc)))
    lib().tco_set_metric3(int(metric3), int(rpar is not None),
                          *(rpar if rpar is not None else (0., 0.)))
    rc = lib().tco_corr2_sampled(f1.h, f2.h if f2 is not None else None, d, coords, nbins,
                                 float(min_sep), float(max_sep), float(bin_size), float(b),
                                 float(a), int(quick), int(exact), int(nthreads),
                                 *[_ptr(out[k]) for k in ('xi0', 'xi1', 'xi2', 'xi3', 'meanr',
                                                          'meanlogr', 'weight', 'npairs')], cnt,
                                 int(row_stride))
    lib().tco_set_periodic(0., 0., 0.)
    lib().tco_set_arc(0)
    lib().tco_set_metric3(0, 0, 0., 0.)
    if rc != 0:
        raise RuntimeError('oracle corr2 failed')
    out['counters'] = dict(visits=cnt[0], accepted=cnt[1], leafpairs=cnt[2], refined=cnt[3])
    return out


def brute2(d, coords, cat1, cat2, nbins, min_sep, max_sep, bin_size, nthreads=0):
    """O(N^2) brute-force pair binning.  cat = dict(x,y,z,k|g1,g2,w)."""
    def cols(c):
        n = len(c['x'])
        w = np.ones(n) if c.get('w') is None else _c(c['w'])
        a = _c(c.get('k')) if d == KDATA else _c(c.get('g1'))
        return (n, _c(c['x']), _c(c['y']), _c(c.get('z')), a, _c(c.get('g2')), w)
    A = cols(cat1)
    if cat2 is None:
        B = (0, None, None, None, None, None, None)
    else:
        B = cols(cat2)
    out = {k: np.zeros(nbins) for k in
           ('xi0', 'xi1', 'xi2', 'xi3', 'meanr', 'meanlogr', 'weight', 'npairs')}
    rc = lib().tco_corr2_brute(d, coords, A[0], *[_ptr(v) for v in A[1:]], B[0],
                               *[_ptr(v) for v in B[1:]], nbins, float(min_sep), float(max_sep),
                               float(bin_size), 1, int(nthreads),
                               *[_ptr(out[k]) for k in ('xi0', 'xi1', 'xi2', 'xi3', 'meanr',
                                                        'meanlogr', 'weight', 'npairs')])
    if rc != 0:
        raise RuntimeError('oracle brute failed')
    return out


Z3_KEYS = ('zeta0', 'zeta1', 'zeta2', 'zeta3', 'zeta4', 'zeta5', 'zeta6', 'zeta7', 'meand2',
           'meanlogd2', 'meand3', 'meanlogd3', 'weight', 'weight_im', 'ntri')


def corr3_multipole(d, coords, f, nbins, min_sep, max_sep, bin_size, b, a, max_n, nthreads=0):
    """Run the oracle's multipole 3-point auto-correlation (NNN d=0, GGG d=4).

    Returns raw sums (before finalize), each of shape (nbins, nbins, 2*max_n+1),
    plus counters (Gn calls, zeta calls, zeta rows, finish stages)."""
    L = lib()
    ntot = nbins * nbins * (2 * max_n + 1)
    out = {k: np.zeros(ntot) for k in Z3_KEYS}
    cnt = (ctypes.c_longlong * 4)()
    rc = L.tco_corr3_multipole(f.h, d, coords, nbins, int(max_n), float(min_sep), float(max_sep),
                               float(bin_size), float(b), float(a), 0, 0, int(nthreads),
                               *[_ptr(out[k]) for k in Z3_KEYS], None, None, cnt)
    if rc != 0:
        raise RuntimeError('oracle multipole failed')
    shape = (nbins, nbins, 2 * max_n + 1)
    res = {k: v.reshape(shape) for k, v in out.items()}
    res['counters'] = dict(gn_calls=cnt[0], zeta_calls=cnt[1], zeta_rows=cnt[2], stages=cnt[3])
    return res


# The multipole passes of the cross-correlation drivers (src/Corr3.cpp: ProcessCross12c
# :3731-3768, ProcessCross21c :3831-3870, ProcessCrossc :3938-3985): (c1 field, list
# fields, ordered) per pass, fields named 1, 2, 3; one list field = the one-list form.
MULTIPOLE_PASSES = {
    ('cross12', 0): [((2, 1, 2), 1), ((1, 2), None)],
    ('cross12', 1): [((1, 2), None)],
    ('cross21', 0): [((2, 1, 1), 1), ((1, 2, 1), 1), ((1, 1, 2), 1)],
    ('cross21', 3): [((1, 1, 2), 4)],
    ('cross', 0): [((2, 1, 3), 1), ((3, 1, 2), 1), ((1, 2, 3), 1)],
    ('cross', 1): [((1, 2, 3), 1)],
    ('cross', 2): [((1, 2, 3), 4), ((3, 2, 1), 4)],
    ('cross', 3): [((1, 2, 3), 4), ((2, 1, 3), 4)],
    ('cross', 4): [((1, 2, 3), 4)],
}


def corr3_multipole_cross(d, coords, kind, ordered, fields, nbins, min_sep, max_sep, bin_size,
                          b, a, max_n, nthreads=0):
    """The oracle's multipole cross-correlation: kind 'cross12' (fields f1, f2; ordered 0
    or 1), 'cross21' (f1, f2; 0 or 3) or 'cross' (f1, f2, f3; 0 to 4), run as the
    reference's drivers run it, pass by pass.  Raw sums as corr3_multipole returns them."""
    L = lib()
    ntot = nbins * nbins * (2 * max_n + 1)
    out = {k: np.zeros(ntot) for k in Z3_KEYS}
    cnt = (ctypes.c_longlong * 4)()
    for names, order in MULTIPOLE_PASSES[(kind, ordered)]:
        fs = [fields[i - 1].h for i in names]
        f3 = fs[2] if len(fs) == 3 else None
        rc = L.tco_corr3_multipole_x(fs[0], fs[1], f3, 0 if order is None else order, d, coords,
                                     nbins, int(max_n), float(min_sep), float(max_sep),
                                     float(bin_size), float(b), float(a), int(nthreads),
                                     *[_ptr(out[k]) for k in Z3_KEYS], cnt)
        if rc != 0:
            raise RuntimeError('oracle multipole failed')
    shape = (nbins, nbins, 2 * max_n + 1)
    res = {k: v.reshape(shape) for k, v in out.items()}
    res['counters'] = dict(gn_calls=cnt[0], zeta_calls=cnt[1], zeta_rows=cnt[2], stages=cnt[3])
    return res


def select_patch(patch, centers, x, y, z=None, chunk=1 << 16):
    """SelectPatch (src/KMeans.cpp:944-1011): 1 for each object unless a centre other than
    `patch` is strictly nearer.  Squared distances: dx^2 + dy^2 in 2-d; in 3-d the
    compiled reference (its -ffast-math build) adds the selected centre's terms as
    (dx^2 + dy^2) + dz^2 and the other centres' as dx^2 + (dy^2 + dz^2), which decides
    objects equidistant from two equal centres (found by matching its outputs; pinned by
    tests/test_oracle_patches.py).  numpy restatement, evaluated chunk by chunk."""
    centers = np.asarray(centers, dtype=float)
    dim = 2 if z is None else 3
    centers = centers.reshape(-1, dim)
    out = np.empty(len(x), dtype=np.int64)
    others = np.arange(len(centers)) != patch
    for s in range(0, len(x), chunk):
        e = min(len(x), s + chunk)
        dx = x[s:e, None] - centers[None, :, 0]
        dy = y[s:e, None] - centers[None, :, 1]
        if z is None:
            rsq = dx * dx + dy * dy
            psq = rsq[:, patch:patch + 1]
        else:
            dz = z[s:e, None] - centers[None, :, 2]
            rsq = dx * dx + (dy * dy + dz * dz)
            psq = (dx[:, patch:patch + 1] * dx[:, patch:patch + 1] +
                   dy[:, patch:patch + 1] * dy[:, patch:patch + 1]) + \
                dz[:, patch:patch + 1] * dz[:, patch:patch + 1]
        out[s:e] = ~np.any(rsq[:, others] < psq, axis=1)
    return out


def quick_assign(centers, x, y, z=None, chunk=1 << 16):
    """QuickAssign (src/KMeans.cpp:890-942): for each object the index of the nearest
    patch centre, squared distance (dx^2 + dy^2) + dz^2 in that order, the first
    centre winning a tie (the reference keeps kmin unless rsq < min_rsq).  numpy
    restatement, evaluated chunk by chunk."""
    centers = np.asarray(centers, dtype=float)
    dim = 2 if z is None else 3
    centers = centers.reshape(-1, dim)
    out = np.empty(len(x), dtype=np.int64)
    for s in range(0, len(x), chunk):
        e = min(len(x), s + chunk)
        dx = x[s:e, None] - centers[None, :, 0]
        dy = y[s:e, None] - centers[None, :, 1]
        rsq = dx * dx + dy * dy
        if z is not None:
            dz = z[s:e, None] - centers[None, :, 2]
            rsq = rsq + dz * dz
        out[s:e] = np.argmin(rsq, axis=1)      # the first minimum
    return out
