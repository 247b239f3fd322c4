"""TEST INFRASTRUCTURE ONLY -- access to the literal reference engine built in oracle/_ref.

oracle/Makefile.ref compiles the TreeCorr 5.1.3 C++ sources (src/*.cpp) into
  oracle/_ref/_treecorr<ext>    the reference's own pybind11 module
  oracle/_ref/libtcref_dump.so  the reference tree build + a pre-order cell dump
Both are optional: helpers return None when the build is absent, and tests that
need them skip.  The reference's Python package is not importable here (it needs
LSSTDESC.Coord), so this module drives the compiled engine directly, restating
only the glue the Python layer adds (treecorr/corr2base.py:686-695, 1070-1160).
"""
import ctypes
import glob
import importlib.util
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REFDIR = os.path.join(HERE, '_ref')
_MOD = None
_DUMP = None


def module():
    """The compiled reference pybind module, or None."""
    global _MOD
    if _MOD is None:
        hits = glob.glob(os.path.join(REFDIR, '_treecorr*.so'))
        if not hits:
            return None
        spec = importlib.util.spec_from_file_location('_treecorr', hits[0])
        m = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(m)
        _MOD = m
    return _MOD


def dumplib():
    global _DUMP
    if _DUMP is None:
        path = os.path.join(REFDIR, 'libtcref_dump.so')
        if not os.path.exists(path):
            return None
        L = ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)
        dp = ctypes.POINTER(ctypes.c_double)
        L.refdump_build.restype = ctypes.c_void_p
        L.refdump_build.argtypes = [ctypes.c_int, ctypes.c_int] + [dp] * 7 + [
            ctypes.c_long, ctypes.c_double, ctypes.c_double] + [ctypes.c_int] * 4
        L.refdump_nnodes.argtypes = [ctypes.c_void_p]
        L.refdump_nnodes.restype = ctypes.c_long
        L.refdump_ntop.argtypes = [ctypes.c_void_p]
        L.refdump_ntop.restype = ctypes.c_long
        L.refdump_get.argtypes = [ctypes.c_void_p] * 10
        L.refdump_free.argtypes = [ctypes.c_void_p]
        _DUMP = L
    return _DUMP


def tree_dump(d, coords, x, y, z=None, k=None, g1=None, g2=None, w=None, wpos=None,
              min_size=0., max_size=np.inf, split_method=2, brute=False, min_top=3, max_top=10):
    """Build a reference Field and return its cells in pre-order (None if not built)."""
    L = dumplib()
    if L is None:
        return None
    dp = ctypes.POINTER(ctypes.c_double)

    def c(a):
        return None if a is None else np.ascontiguousarray(a, dtype=float)
    n = len(x)
    arrs = [c(x), c(y), c(z), c(k) if d == 1 else c(g1), c(g2),
            c(w) if w is not None else np.ones(n), c(wpos)]
    ptrs = [None if a is None else a.ctypes.data_as(dp) for a in arrs]
    h = L.refdump_build(d, coords, *ptrs, n, float(min_size), float(max_size), int(split_method),
                        int(bool(brute)), int(min_top), int(max_top))
    nn = L.refdump_nnodes(h)
    nt = L.refdump_ntop(h)
    out = dict(pos=np.zeros((nn, 3)), size=np.zeros(nn, np.float32), w=np.zeros(nn, np.float32),
               wk=np.zeros(nn, np.float32), g=np.zeros((nn, 2), np.float32),
               n=np.zeros(nn, np.int64), left=np.zeros(nn, np.int32),
               right=np.zeros(nn, np.int32), top=np.zeros(nt, np.int32))
    L.refdump_get(h, *[ctypes.c_void_p(out[key].ctypes.data) for key in
                       ('pos', 'size', 'w', 'wk', 'g', 'n', 'left', 'right', 'top')])
    L.refdump_free(h)
    return out


_FIELD = {(0, 1): 'NFieldFlat', (0, 2): 'NFieldThreeD', (0, 3): 'NFieldSphere',
          (1, 1): 'KFieldFlat', (1, 2): 'KFieldThreeD', (1, 3): 'KFieldSphere',
          (4, 1): 'GFieldFlat', (4, 2): 'GFieldThreeD', (4, 3): 'GFieldSphere',
          (2, 1): 'ZFieldFlat', (2, 2): 'ZFieldThreeD', (2, 3): 'ZFieldSphere',
          (3, 1): 'VFieldFlat', (3, 2): 'VFieldThreeD', (3, 3): 'VFieldSphere',
          (5, 1): 'TFieldFlat', (5, 2): 'TFieldThreeD', (5, 3): 'TFieldSphere',
          (6, 1): 'QFieldFlat', (6, 2): 'QFieldThreeD', (6, 3): 'QFieldSphere'}
_CORR = {0: 'NNCorr', 1: 'KKCorr', 2: 'ZZCorr', 3: 'VVCorr', 4: 'GGCorr', 5: 'TTCorr',
         6: 'QQCorr'}
_CORR_MIXED = {(0, 1): 'NKCorr', (0, 4): 'NGCorr', (1, 4): 'KGCorr', (0, 2): 'NZCorr',
               (1, 2): 'KZCorr', (0, 3): 'NVCorr', (0, 5): 'NTCorr', (0, 6): 'NQCorr',
               (1, 3): 'KVCorr', (1, 5): 'KTCorr', (1, 6): 'KQCorr'}


def make_field(d, coords, x, y, z=None, k=None, g1=None, g2=None, w=None, wpos=None,
               min_size=0., max_size=np.inf, split_method=2, brute=False, min_top=3, max_top=10):
    """A reference _treecorr Field (pybind signature src/Field.cpp:644-691)."""
    m = module()
    e = np.array([])

    def c(a):
        return e if a is None else np.ascontiguousarray(a, dtype=float)
    w = np.ones(len(x)) if w is None else w
    sm = [m.Middle, m.Median, m.Mean, m.Random][split_method]
    cls = getattr(m, _FIELD[(d, coords)])
    if d == 0:
        return cls(c(x), c(y), c(z), c(w), c(wpos), float(min_size), float(max_size), sm, 0,
                   bool(brute), int(min_top), int(max_top))
    if d == 1:
        return cls(c(x), c(y), c(z), c(k), c(w), c(wpos), float(min_size), float(max_size), sm,
                   0, bool(brute), int(min_top), int(max_top))
    return cls(c(x), c(y), c(z), c(g1), c(g2), c(w), c(wpos), float(min_size), float(max_size),
               sm, 0, bool(brute), int(min_top), int(max_top))


def corr2(d, f1, f2, nbins, min_sep, max_sep, bin_size, b, a, quick=False, nthreads=None,
          d2=None, period=None, arc=False, metric3=0, rpar=None):
    """Run the reference 2-point engine (auto when f2 is None).  Returns raw sums.
    d2 != d selects NK, NG or KG (treecorr/nkcorrelation.py, ngcorrelation.py,
    kgcorrelation.py: xi, and xi_im for a G second field)."""
    m = module()
    if nthreads is not None:
        m.SetOMPThreads(int(nthreads))
    out = {k: np.zeros(nbins) for k in
           ('xi0', 'xi1', 'xi2', 'xi3', 'meanr', 'meanlogr', 'weight', 'npairs')}
    e = np.array([])
    xi = [out['xi0'], out['xi1'], out['xi2'], out['xi3']]
    mixed = d2 is not None and d2 != d
    nxi = (2 if d2 >= 2 else 1) if mixed else (0 if d == 0 else 1 if d == 1 else 4)
    xi = [xi[i] if i < nxi else e for i in range(4)]
    name = _CORR_MIXED[(d, d2)] if mixed else _CORR[d]
    xp, yp, zp = period if period is not None else (0., 0., 0.)
    metric = m.Periodic if period is not None else m.Arc if arc else \
        {0: m.Euclidean, 2: m.Rperp, 3: m.Rlens, 5: m.OldRperp}[metric3]
    minrpar, maxrpar = rpar if rpar is not None else (-1.7976931348623157e308,
                                                      1.7976931348623157e308)
    corr = getattr(m, name)(m.Log, float(min_sep), float(max_sep), int(nbins),
                            float(bin_size), float(b), float(a), float(minrpar),
                            float(maxrpar), float(xp), float(yp), float(zp), *xi,
                            out['meanr'], out['meanlogr'], out['weight'], out['npairs'])
    if f2 is None:
        corr.processAuto(f1, False, bool(quick), metric)
    else:
        corr.processCross(f1, f2, False, bool(quick), metric)
    return out


_CORR3 = {0: 'NNNCorr', 4: 'GGGCorr'}


def _corr3_mp(d, nbins, min_sep, max_sep, bin_size, b, a, max_n, nthreads):
    m = module()
    if nthreads is not None:
        m.SetOMPThreads(int(nthreads))
    ntot = nbins * nbins * (2 * max_n + 1)
    keys = ('zeta0', 'zeta1', 'zeta2', 'zeta3', 'zeta4', 'zeta5', 'zeta6', 'zeta7', 'meand1',
            'meanlogd1', 'meand2', 'meanlogd2', 'meand3', 'meanlogd3', 'meanu', 'meanv',
            'weight', 'weight_im', 'ntri')
    out = {k: np.zeros(ntot) for k in keys}
    e = np.array([])
    z = [out['zeta%d' % i] if d == 4 else e for i in range(8)]
    corr = getattr(m, _CORR3[d])(m.LogMultipole, float(min_sep), float(max_sep), int(nbins),
                                 float(bin_size), float(b), float(a), 0., 0., int(max_n), 0., 0.,
                                 0., 0., 0, 0., 0., -1.7976931348623157e308,
                                 1.7976931348623157e308, 0., 0., 0., *z, out['meand1'],
                                 out['meanlogd1'], out['meand2'], out['meanlogd2'], out['meand3'],
                                 out['meanlogd3'], out['meanu'], out['meanv'], out['weight'],
                                 out['weight_im'], out['ntri'])
    return m, corr, out


def corr3_multipole(d, f, nbins, min_sep, max_sep, bin_size, b, a, max_n, nthreads=None):
    """Run the reference LogMultipole auto-correlation (src/Corr3.cpp:1124-1175).

    Constructor arguments follow treecorr/corr3base.py:1116-1131 with the
    LogMultipole settings of :691-705 (nubins = max_n, u/v ranges zero)."""
    m, corr, out = _corr3_mp(d, nbins, min_sep, max_sep, bin_size, b, a, max_n, nthreads)
    corr.processAuto(f, False, False, m.Euclidean)
    shape = (nbins, nbins, 2 * max_n + 1)
    return {k: v.reshape(shape) for k, v in out.items()}


def corr3_multipole_cross(d, kind, ordered, fields, nbins, min_sep, max_sep, bin_size, b, a,
                          max_n, nthreads=None):
    """The reference's LogMultipole cross-correlations through its own drivers:
    processCross12 / processCross21 (fields f1, f2) and processCross (f1, f2, f3) with the
    ordered code the Python layer passes (corr3base.py:1968, 2027, 2086)."""
    m, corr, out = _corr3_mp(d, nbins, min_sep, max_sep, bin_size, b, a, max_n, nthreads)
    fn = {'cross12': corr.processCross12, 'cross21': corr.processCross21,
          'cross': corr.processCross}[kind]
    fn(*fields, int(ordered), False, False, m.Euclidean)
    shape = (nbins, nbins, 2 * max_n + 1)
    return {k: v.reshape(shape) for k, v in out.items()}
