"""2-point correlations on the GPU: NNCorrelation, KKCorrelation, GGCorrelation.

Restates the Python layer of TreeCorr that drives the engine -- binning setup
(treecorr/corr2base.py:349-571), field sizes (:1595-1625), process/process_auto/
process_cross with patch loops (:836-1272), finalize (:1274-1294, unit handling
:1586-1593) and the per-class finalizers (nncorrelation.py, kkcorrelation.py,
zzcorrelation.py/ggcorrelation.py) -- over the tc_corr2 C ABI.
"""
import ctypes
import math
import sys

import numpy as np

from . import _lib
from .catalog import Catalog, calculateVarG, calculateVarK
from .config import parse_unit
from .field import coord_enum
from .util import set_omp_threads

LOG = 0
EUCLIDEAN = 1


class Corr2:
    """Base class of the 2-point correlation objects (treecorr/corr2base.py:Corr2)."""
    _default_angle_slop = 0.1
    _letters = None
    _d = None
    _nxi = 0

    def __init__(self, config=None, *, logger=None, rng=None, **kwargs):
        cfg = dict(config or {})
        cfg.update(kwargs)
        self.config = cfg
        bin_type = cfg.get('bin_type', 'Log')
        if bin_type != 'Log':
            raise NotImplementedError("treecorr_amd implements Log binning for 2-point correlations")
        self.bin_type = bin_type
        self.sep_units = cfg.get('sep_units', '')
        self._sep_units = parse_unit(cfg.get('sep_units') or 'radians')
        self._log_sep_units = math.log(self._sep_units)
        nbins = cfg.get('nbins')
        bin_size = cfg.get('bin_size')
        min_sep = cfg.get('min_sep')
        max_sep = cfg.get('max_sep')
        # corr2base.py:382-449
        if nbins is None:
            if max_sep is None:
                raise TypeError("Missing required parameter max_sep")
            if min_sep is None:
                raise TypeError("Missing required parameter min_sep")
            if bin_size is None:
                raise TypeError("Missing required parameter bin_size")
            min_sep, max_sep, bin_size = float(min_sep), float(max_sep), float(bin_size)
            if min_sep >= max_sep:
                raise ValueError("max_sep must be larger than min_sep")
            nbins = int(math.ceil(math.log(max_sep / min_sep) / bin_size))
            bin_size = math.log(max_sep / min_sep) / nbins
        elif bin_size is None:
            if max_sep is None:
                raise TypeError("Missing required parameter max_sep")
            if min_sep is None:
                raise TypeError("Missing required parameter min_sep")
            min_sep, max_sep, nbins = float(min_sep), float(max_sep), int(nbins)
            if min_sep >= max_sep:
                raise ValueError("max_sep must be larger than min_sep")
            bin_size = math.log(max_sep / min_sep) / nbins
        elif max_sep is None:
            if min_sep is None:
                raise TypeError("Missing required parameter min_sep")
            min_sep, nbins, bin_size = float(min_sep), int(nbins), float(bin_size)
            max_sep = math.exp(nbins * bin_size) * min_sep
        else:
            if min_sep is not None:
                raise TypeError("Only 3 of min_sep, max_sep, bin_size, nbins are allowed.")
            max_sep, nbins, bin_size = float(max_sep), int(nbins), float(bin_size)
            min_sep = max_sep * math.exp(-nbins * bin_size)
        self.nbins, self.bin_size, self.min_sep, self.max_sep = nbins, bin_size, min_sep, max_sep
        self.logr = np.linspace(0, nbins * bin_size, nbins, endpoint=False, dtype=float)
        self.logr += math.log(min_sep) + 0.5 * bin_size
        self.rnom = np.exp(self.logr)
        half_bin = np.exp(0.5 * bin_size)
        self.left_edges = self.rnom / half_bin
        self.right_edges = self.rnom * half_bin
        max_good_slop = 0.1 / bin_size
        self._min_sep = min_sep * self._sep_units
        self._max_sep = max_sep * self._sep_units
        self._bin_size = bin_size
        self.split_method = cfg.get('split_method', 'mean')
        self.min_top = cfg.get('min_top')
        self.max_top = int(cfg.get('max_top', 10))
        bin_slop = cfg.get('bin_slop')
        self.angle_slop = float(cfg.get('angle_slop', self._default_angle_slop))
        if bin_slop is None or bin_slop < 0:
            bin_slop = min(max_good_slop, 1.0)
        self.bin_slop = float(bin_slop)
        self.b = self._bin_size * self.bin_slop
        self.brute = cfg.get('brute', False)
        self.min_rpar = -sys.float_info.max
        self.max_rpar = sys.float_info.max
        if cfg.get('min_rpar') is not None or cfg.get('max_rpar') is not None:
            raise NotImplementedError("min_rpar/max_rpar are not supported by treecorr_amd")
        if cfg.get('period') or cfg.get('xperiod') or cfg.get('yperiod') or cfg.get('zperiod'):
            raise NotImplementedError("periodic metrics are not supported by treecorr_amd")
        self.var_method = cfg.get('var_method', 'shot')
        self.coords = None
        self.metric = None
        self.results = {}
        self.npatch1 = self.npatch2 = 1
        self.meanr = np.zeros(nbins)
        self.meanlogr = np.zeros(nbins)
        self.weight = np.zeros(nbins)
        self.npairs = np.zeros(nbins)
        self._xi = [np.zeros(nbins) if i < self._nxi else None for i in range(4)]
        self._var_num = 0.
        self._varxi = None
        self._cov = None
        self._corr = None
        self._corr_only = False
        self._processed_cats1 = []
        self._processed_cats2 = []
        self.counters = None

    # ---- the C-ABI correlation object (corr2base.py:686-695) ----
    @property
    def corr(self):
        if self._corr is None:
            h = ctypes.c_void_p()
            xi = [_lib.dptr(a) if a is not None else None for a in self._xi]
            _lib.check(_lib.lib().tc_corr2_create(
                self._d, self._d, LOG, self._min_sep, self._max_sep, self.nbins, self._bin_size,
                self.b, self.angle_slop, self.min_rpar, self.max_rpar, 0., 0., 0., *xi,
                _lib.dptr(self.meanr), _lib.dptr(self.meanlogr), _lib.dptr(self.weight),
                _lib.dptr(self.npairs), ctypes.byref(h)))
            self._corr = h
        return self._corr

    def __del__(self):
        h = getattr(self, '_corr', None)
        if h is not None and h.value:
            try:
                _lib.lib().tc_corr2_destroy(h)
            except Exception:
                pass

    def _set_metric(self, metric, coords1, coords2=None):
        metric = metric or self.config.get('metric', 'Euclidean')
        if metric != 'Euclidean':
            raise NotImplementedError("treecorr_amd implements the Euclidean metric")
        if coords2 is not None and coords1 != coords2:
            raise ValueError("Cannot correlate catalogs with different coordinate systems.")
        if self.sep_units != '' and coords1 == '3d':
            raise ValueError("sep_units is invalid with 3d coordinates. ")
        self.coords = coords1
        self.metric = metric

    def _get_minmax_size(self):
        # corr2base.py:1595-1625 (Log, Euclidean)
        b1 = min(self.angle_slop, self.b)
        min_size = self._min_sep * b1 / (2. + 3. * b1)
        max_size = self._max_sep * b1
        return min_size, max_size

    def clear(self):
        for a in self._xi:
            if a is not None:
                a[:] = 0
        self.meanr[:] = 0
        self.meanlogr[:] = 0
        self.weight[:] = 0
        self.npairs[:] = 0
        self.results = {}
        self._varxi = None
        self._cov = None

    def _field(self, cat, **kw):
        getter = getattr(cat, 'get%sField' % self._letters[0])
        return getter(**kw)

    def process_auto(self, cat, *, metric=None, num_threads=None, corr_only=False):
        """Accumulate the auto-correlation of one catalog (corr2base.py:1070-1111)."""
        self._set_metric(metric, cat.coords)
        if num_threads is not None:
            set_omp_threads(num_threads)
        self._corr_only = corr_only
        self._meanww = cat.meanw ** 2
        min_size, max_size = self._get_minmax_size()
        field = self._field(cat, min_size=min_size, max_size=max_size,
                            split_method=self.split_method, brute=bool(self.brute),
                            min_top=self.min_top, max_top=self.max_top, coords=self.coords)
        _lib.check(_lib.lib().tc_corr2_process_auto(self.corr, field.data, 0, int(bool(corr_only)),
                                                     EUCLIDEAN))
        self._read_counters()

    def process_cross(self, cat1, cat2, *, metric=None, num_threads=None, corr_only=False):
        """Accumulate the cross-correlation of two catalogs (corr2base.py:1113-1160)."""
        self._set_metric(metric, cat1.coords, cat2.coords)
        if num_threads is not None:
            set_omp_threads(num_threads)
        self._corr_only = corr_only
        self._meanww = cat1.meanw * cat2.meanw
        min_size, max_size = self._get_minmax_size()
        f1 = self._field(cat1, min_size=min_size, max_size=max_size,
                         split_method=self.split_method,
                         brute=self.brute is True or self.brute == 1,
                         min_top=self.min_top, max_top=self.max_top, coords=self.coords)
        f2 = self._field(cat2, min_size=min_size, max_size=max_size,
                         split_method=self.split_method,
                         brute=self.brute is True or self.brute == 2,
                         min_top=self.min_top, max_top=self.max_top, coords=self.coords)
        _lib.check(_lib.lib().tc_corr2_process_cross(self.corr, f1.data, f2.data, 0,
                                                      int(bool(corr_only)), EUCLIDEAN))
        self._read_counters()

    def _read_counters(self):
        c = (ctypes.c_longlong * 4)()
        _lib.check(_lib.lib().tc_corr2_counters(self.corr, c))
        self.counters = dict(visits=c[0], accepted=c[1], leafpairs=c[2], refined=c[3])

    def _trivially_zero(self, c1, c2):
        x1, y1, z1, s1 = c1._get_center_size()
        x2, y2, z2, s2 = c2._get_center_size()
        r = ctypes.c_int()
        _lib.check(_lib.lib().tc_corr2_trivially_zero(self.corr, EUCLIDEAN, coord_enum(self.coords),
                                                       x1, y1, z1, s1, x2, y2, z2, s2,
                                                       ctypes.byref(r)))
        return bool(r.value)

    def _snapshot(self):
        return dict(xi=[None if a is None else a.copy() for a in self._xi], meanr=self.meanr.copy(),
                    meanlogr=self.meanlogr.copy(), weight=self.weight.copy(),
                    npairs=self.npairs.copy())

    def _delta_since(self, snap):
        d = self._snapshot()
        for key in ('meanr', 'meanlogr', 'weight', 'npairs'):
            d[key] = d[key] - snap[key]
        d['xi'] = [None if a is None else a - b for a, b in zip(d['xi'], snap['xi'])]
        return d

    def process(self, cat1, cat2=None, metric=None, num_threads=None, comm=None, low_mem=False,
                initialize=True, finalize=True, patch_method='global', corr_only=False):
        """Compute the correlation function (corr2base.py:1162-1272)."""
        if comm is not None:
            raise NotImplementedError("MPI is not part of the single-GPU engine")
        if patch_method not in ('local', 'global'):
            raise ValueError("Invalid patch_method %s" % patch_method)
        if patch_method == 'local':
            raise NotImplementedError("patch_method='local' is not implemented")
        if initialize:
            self.clear()
            self._processed_cats1 = []
            self._processed_cats2 = []
            self._corr_only = corr_only
        if not isinstance(cat1, list):
            cat1 = cat1.get_patches()
        if cat2 is not None and not isinstance(cat2, list):
            cat2 = cat2.get_patches()
        if cat2 is None:
            self._process_all_auto(cat1, metric, num_threads, corr_only)
        else:
            self._process_all_cross(cat1, cat2, metric, num_threads, corr_only)
        self._processed_cats1.extend(cat1)
        if cat2 is not None:
            self._processed_cats2.extend(cat2)
        if finalize:
            self._do_finalize(cat2 is None, corr_only)
            self._processed_cats1 = []
            self._processed_cats2 = []

    def _do_finalize(self, auto, corr_only):
        if corr_only:
            w1 = np.sum([c.sumw for c in self._processed_cats1]) / \
                np.sum([c.nobj for c in self._processed_cats1])
            cats2 = self._processed_cats1 if auto else self._processed_cats2
            w2 = np.sum([c.sumw for c in cats2]) / np.sum([c.nobj for c in cats2])
            self._meanww = w1 * w2
        v1 = self._calculate_var(self._processed_cats1)
        v2 = v1 if auto else self._calculate_var(self._processed_cats2)
        if v1 is None:
            self.finalize()
        else:
            self.finalize(v1, v2)

    def _calculate_var(self, cats):
        return None

    def _single_pair(self, c1, c2, ij, metric, num_threads, corr_only, force=False):
        # corr2base.py:836-853: skip trivially-zero patch pairs, keep per-pair results
        snap = self._snapshot()
        if c2 is not None and not self._trivially_zero(c1, c2):
            self.process_cross(c1, c2, metric=metric, num_threads=num_threads,
                               corr_only=corr_only)
        delta = self._delta_since(snap)
        if np.any(delta['weight']) or force:
            self.results[ij] = delta
        self._add_tot(ij, c1, c2)

    def _add_tot(self, ij, c1, c2):
        pass

    def _process_all_auto(self, cat1, metric, num_threads, corr_only):
        # corr2base.py:855-964 (global patch method, single process)
        if len(cat1) == 1 and cat1[0].npatch == 1:
            self.process_auto(cat1[0], metric=metric, num_threads=num_threads,
                              corr_only=corr_only)
            return
        self.npatch1 = self.npatch2 = len(cat1)
        self._set_metric(metric, cat1[0].coords)
        for ii, c1 in enumerate(cat1):
            i = c1._single_patch if c1._single_patch is not None else ii
            snap = self._snapshot()
            self.process_auto(c1, metric=metric, num_threads=num_threads, corr_only=corr_only)
            self.results[(i, i)] = self._delta_since(snap)
            for jj in range(len(cat1) - 1, -1, -1):
                c2 = cat1[jj]
                j = c2._single_patch if c2._single_patch is not None else jj
                if i < j:
                    self._single_pair(c1, c2, (i, j), metric, num_threads, corr_only)

    def _process_all_cross(self, cat1, cat2, metric, num_threads, corr_only):
        # corr2base.py:966-1050 (global patch method, single process)
        if len(cat1) == 1 and len(cat2) == 1 and cat1[0].npatch == 1 and cat2[0].npatch == 1:
            self.process_cross(cat1[0], cat2[0], metric=metric, num_threads=num_threads,
                               corr_only=corr_only)
            return
        self.npatch1, self.npatch2 = len(cat1), len(cat2)
        self._set_metric(metric, cat1[0].coords, cat2[0].coords)
        n1, n2 = len(cat1), len(cat2)
        for ii, c1 in enumerate(cat1):
            i = c1._single_patch if c1._single_patch is not None else ii
            for jj, c2 in enumerate(cat2):
                j = c2._single_patch if c2._single_patch is not None else jj
                self._single_pair(c1, c2, (i, j), metric, num_threads, corr_only,
                                  force=(i == j or n1 == 1 or n2 == 1))

    def _finalize(self):
        # corr2base.py:1274-1294
        mask1 = self.weight != 0
        for a in self._xi:
            if a is not None:
                a[mask1] /= self.weight[mask1]
        if self._corr_only:
            self.meanr[:] = self.rnom
            self.meanlogr[:] = self.logr
            self.npairs[:] = self.weight / self._meanww
        else:
            self.meanr[mask1] /= self.weight[mask1]
            self.meanlogr[mask1] /= self.weight[mask1]
            self._apply_units(mask1)
            mask2 = self.weight == 0
            self.meanr[mask2] = self.rnom[mask2]
            self.meanlogr[mask2] = self.logr[mask2]

    def _apply_units(self, mask):
        # corr2base.py:1586-1593
        if self.coords == 'spherical' and self.metric == 'Euclidean':
            self.meanr[mask] = 2. * np.arcsin(self.meanr[mask] / 2.)
            self.meanlogr[mask] = np.log(2. * np.arcsin(np.exp(self.meanlogr[mask]) / 2.))
        self.meanr[mask] /= self._sep_units
        self.meanlogr[mask] -= self._log_sep_units

    def getWeight(self):
        return self.weight.ravel()

    def _calculate_varxi(self, n):
        # corr2base.py:1354-1358 with _cov_shot :2393-2408
        self._varxi = [np.zeros(self.nbins) for _ in range(n)]
        if self._var_num != 0:
            v = self.getWeight().copy()
            mask = v != 0
            v[mask] = self._var_num / v[mask]
            for i in range(n):
                self._varxi[i][:] = v[i * self.nbins:(i + 1) * self.nbins]


class NNCorrelation(Corr2):
    """Count-count correlation (treecorr/nncorrelation.py)."""
    _letters = 'NN'
    _d = 0
    _nxi = 0
    _default_angle_slop = 1

    def __init__(self, config=None, *, logger=None, **kwargs):
        super().__init__(config, logger=logger, **kwargs)
        self.tot = 0.
        self.xi = None
        self.varxi = None

    def clear(self):
        super().clear()
        self.tot = 0.

    def process_auto(self, cat, **kw):
        super().process_auto(cat, **kw)
        self.tot += 0.5 * cat.sumw ** 2

    def process_cross(self, cat1, cat2, **kw):
        super().process_cross(cat1, cat2, **kw)
        self.tot += cat1.sumw * cat2.sumw

    def _add_tot(self, ij, c1, c2):
        # nncorrelation.py:194-204: skipped pairs still count toward tot
        pass

    def _single_pair(self, c1, c2, ij, metric, num_threads, corr_only, force=False):
        tz = c2 is not None and self._trivially_zero(c1, c2)
        super()._single_pair(c1, c2, ij, metric, num_threads, corr_only, force)
        if tz or c2 is None:
            self.tot += c1.sumw * (c2.sumw if c2 is not None else 0.)

    def finalize(self):
        self._finalize()

    def _mean_weight(self):
        mean_np = np.mean(self.npairs)
        return 1 if mean_np == 0 else np.mean(self.weight) / mean_np

    def calculateXi(self, *, rr, dr=None, rd=None):
        """xi from DD (self) and RR [, DR, RD] (nncorrelation.py:230-340), shot variance."""
        if rr.tot == 0:
            raise ValueError("rr has tot=0.")
        rrf = self.tot / rr.tot
        denom = rr.weight * rrf
        if dr is None and rd is None:
            xi = self.weight - denom
        elif rd is not None and dr is None:
            xi = self.weight - 2. * rd.weight * (self.tot / rd.tot) + denom
        elif dr is not None and rd is None:
            xi = self.weight - 2. * dr.weight * (self.tot / dr.tot) + denom
        else:
            xi = self.weight - dr.weight * (self.tot / dr.tot) - rd.weight * (self.tot / rd.tot) + denom
        denom = denom.copy()
        denom[rr.weight == 0] = 1.
        xi = xi / denom
        ddw = self._mean_weight()
        rrw = rr._mean_weight()
        if dr is None and rd is None:
            varxi_factor = 1 + rrf * rrw / ddw
        elif rd is not None and dr is None:
            varxi_factor = 1 + 2 * (self.tot / rd.tot) * rd._mean_weight() / ddw + rrf * rrw / ddw
        elif dr is not None and rd is None:
            varxi_factor = 1 + 2 * (self.tot / dr.tot) * dr._mean_weight() / ddw + rrf * rrw / ddw
        else:
            varxi_factor = (1 + (self.tot / rd.tot) * rd._mean_weight() / ddw +
                            (self.tot / dr.tot) * dr._mean_weight() / ddw + rrf * rrw / ddw)
        self._var_num = ddw * varxi_factor ** 2
        self._rr_weight = rr.weight * rrf
        w = self._rr_weight.copy()
        mask = w != 0
        varxi = np.zeros(self.nbins)
        varxi[mask] = self._var_num / w[mask]
        self.xi, self.varxi = xi, varxi
        return xi, varxi


class KKCorrelation(Corr2):
    """Scalar-scalar correlation (treecorr/kkcorrelation.py)."""
    _letters = 'KK'
    _d = 1
    _nxi = 1
    _default_angle_slop = 1

    @property
    def xi(self):
        return self._xi[0]

    def _calculate_var(self, cats):
        return calculateVarK(cats)

    def finalize(self, vark1, vark2):
        # kkcorrelation.py:100-112
        self._finalize()
        self._var_num = vark1 * vark2

    @property
    def varxi(self):
        if self._varxi is None:
            self._calculate_varxi(1)
        return self._varxi[0]


class GGCorrelation(Corr2):
    """Shear-shear correlation (treecorr/ggcorrelation.py, zzcorrelation.py)."""
    _letters = 'GG'
    _d = 4
    _nxi = 4

    xip = property(lambda self: self._xi[0])
    xip_im = property(lambda self: self._xi[1])
    xim = property(lambda self: self._xi[2])
    xim_im = property(lambda self: self._xi[3])

    def _calculate_var(self, cats):
        return calculateVarG(cats)

    def finalize(self, varg1, varg2):
        # zzcorrelation.py:76-88
        self._finalize()
        self._var_num = 2. * varg1 * varg2

    def getWeight(self):
        return np.concatenate([self.weight.ravel(), self.weight.ravel()])

    @property
    def varxip(self):
        if self._varxi is None:
            self._calculate_varxi(2)
        return self._varxi[0]

    @property
    def varxim(self):
        if self._varxi is None:
            self._calculate_varxi(2)
        return self._varxi[1]
