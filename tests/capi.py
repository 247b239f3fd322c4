"""Direct C-ABI calls for the GPU parity tests (include/treecorr_amd.h).

The parity tests go through the boundary itself: fields are built with
tc_field_create and correlations run with tc_corr2_* / tc_corr3_*, exactly as
a TreeCorr binding would call them.
"""
import ctypes
import math

import numpy as np

from treecorr_amd import _lib

KEYS2 = ('xi0', 'xi1', 'xi2', 'xi3', 'meanr', 'meanlogr', 'weight', 'npairs')
KEYS3 = ('zeta0', 'zeta1', 'zeta2', 'zeta3', 'zeta4', 'zeta5', 'zeta6', 'zeta7', 'meand1',
         'meanlogd1', 'meand2', 'meanlogd2', 'meand3', 'meanlogd3', 'meanu', 'meanv', 'weight',
         'weight_im', 'ntri')


def _c(a):
    return None if a is None else np.ascontiguousarray(a, dtype=float)


class Field:
    def __init__(self, d, coords, cat, min_size=0., max_size=np.inf, split_method=2, brute=False,
                 min_top=3, max_top=10):
        self.cols = [_c(cat['x']), _c(cat['y']), _c(cat.get('z')) if coords != 1 else None,
                     _c(cat.get('k')) if d == 1 else _c(cat.get('g1')) if d >= 2 else None,
                     _c(cat.get('g2')) if d >= 2 else None,
                     _c(cat['w']) if cat.get('w') is not None else np.ones(len(cat['x'])),
                     _c(cat.get('wpos'))]
        self.h = ctypes.c_void_p()
        _lib.check(_lib.lib().tc_field_create(d, coords, *[_lib.dptr(a) for a in self.cols],
                                              len(self.cols[0]), float(min_size), float(max_size),
                                              int(split_method), 0, int(bool(brute)),
                                              int(min_top), int(max_top), ctypes.byref(self.h)))

    def __del__(self):
        if getattr(self, 'h', None) is not None and self.h.value:
            _lib.lib().tc_field_destroy(self.h)
            self.h = None

    @property
    def ntop(self):
        n = ctypes.c_longlong()
        _lib.check(_lib.lib().tc_field_ntop(self.h, ctypes.byref(n)))
        return n.value


def corr2(d, f1, f2, nbins, min_sep, max_sep, bin_size, b, a, quick=False, d2=None,
          period=None, arc=False, metric3=0, rpar=None):
    """d2 != d: the mixed cross correlations (f1 the d side); period (xp, yp, zp): the
    Periodic metric; arc: the Arc metric; metric3: Rperp (2), Rlens (3) or OldRperp (5);
    rpar: the (min, max) r_par range."""
    out = {k: np.zeros(nbins) for k in KEYS2}
    d2 = d if d2 is None else d2
    nxi = (0 if d == 0 else 1 if d == 1 else 4) if d2 == d else (2 if d2 >= 2 else 1)
    xi = [_lib.dptr(out['xi%d' % i]) if i < nxi else None for i in range(4)]
    h = ctypes.c_void_p()
    L = _lib.lib()
    mnr, mxr = rpar if rpar is not None else (-1.7976931348623157e308, 1.7976931348623157e308)
    _lib.check(L.tc_corr2_create(d, d2, 0, float(min_sep), float(max_sep), int(nbins),
                                 float(bin_size), float(b), float(a), float(mnr), float(mxr),
                                 *(period if period is not None else (0., 0., 0.)), *xi,
                                 *[_lib.dptr(out[k]) for k in KEYS2[4:]], ctypes.byref(h)))
    metric = 6 if period is not None else 4 if arc else metric3 if metric3 else 1
    try:
        if f2 is None:
            _lib.check(L.tc_corr2_process_auto(h, f1.h, 0, int(quick), metric))
        else:
            _lib.check(L.tc_corr2_process_cross(h, f1.h, f2.h, 0, int(quick), metric))
        c = (ctypes.c_longlong * 4)()
        _lib.check(L.tc_corr2_counters(h, c))
        out['counters'] = dict(visits=c[0], accepted=c[1], leafpairs=c[2], refined=c[3])
    finally:
        L.tc_corr2_destroy(h)
    for i in range(nxi, 4):
        out['xi%d' % i] = np.array([])
    return out


def corr3(d, f, nbins, min_sep, max_sep, bin_size, b, a, max_n, quick=False):
    ntot = nbins * nbins * (2 * max_n + 1)
    out = {k: np.zeros(ntot) for k in KEYS3}
    ptrs = [_lib.dptr(out[k]) if (d == 4 or not k.startswith('zeta')) else None for k in KEYS3]
    h = ctypes.c_void_p()
    L = _lib.lib()
    _lib.check(L.tc_corr3_create(d, 5, float(min_sep), float(max_sep), int(nbins),
                                 float(bin_size), float(b), float(a), int(max_n), *ptrs,
                                 ctypes.byref(h)))
    try:
        _lib.check(L.tc_corr3_process_auto(h, f.h, 0, int(quick), 1))
        c = (ctypes.c_longlong * 4)()
        _lib.check(L.tc_corr3_counters(h, c))
    finally:
        L.tc_corr3_destroy(h)
    shape = (nbins, nbins, 2 * max_n + 1)
    res = {k: v.reshape(shape) for k, v in out.items()}
    res['counters'] = dict(gn_calls=c[0], zeta_calls=c[1], zeta_rows=c[2], stages=c[3])
    return res


def corr3_cross(d, kind, ordered, fields, nbins, min_sep, max_sep, bin_size, b, a, max_n,
                quick=False):
    """tc_corr3_process_cross12 / _cross21 (fields f1, f2) or _cross (f1, f2, f3)."""
    ntot = nbins * nbins * (2 * max_n + 1)
    out = {k: np.zeros(ntot) for k in KEYS3}
    ptrs = [_lib.dptr(out[k]) if (d == 4 or not k.startswith('zeta')) else None for k in KEYS3]
    h = ctypes.c_void_p()
    L = _lib.lib()
    _lib.check(L.tc_corr3_create(d, 5, float(min_sep), float(max_sep), int(nbins),
                                 float(bin_size), float(b), float(a), int(max_n), *ptrs,
                                 ctypes.byref(h)))
    try:
        fn = getattr(L, 'tc_corr3_process_' + kind)
        _lib.check(fn(h, *[f.h for f in fields], int(ordered), 0, int(quick), 1))
        c = (ctypes.c_longlong * 4)()
        _lib.check(L.tc_corr3_counters(h, c))
    finally:
        L.tc_corr3_destroy(h)
    shape = (nbins, nbins, 2 * max_n + 1)
    res = {k: v.reshape(shape) for k, v in out.items()}
    res['counters'] = dict(gn_calls=c[0], zeta_calls=c[1], zeta_rows=c[2], stages=c[3])
    return res


def multipole_binning(scale, nbins, lo, hi, max_n, bin_slop=None):
    """LogMultipole binning and field sizes (corr3base.py:714-745, 2955-2960)."""
    min_sep, max_sep = scale * lo, scale * hi
    bin_size = math.log(max_sep / min_sep) / nbins
    a = math.pi / (2 * max_n + 1)
    if bin_slop is None:
        b = bin_size if bin_size <= 0.1 else 0.1
    else:
        b = bin_size * bin_slop
    b1 = min(a, b)
    return dict(nbins=nbins, min_sep=min_sep, max_sep=max_sep, bin_size=bin_size, b=b, a=a,
                max_n=max_n, min_size=2 * min_sep * b1 / (2 + 3 * b1), max_size=max_sep * b1)
