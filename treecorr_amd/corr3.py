"""3-point multipole correlations on the GPU: NNNCorrelation, GGGCorrelation.

Restates TreeCorr's LogMultipole front-end (treecorr/corr3base.py: binning
:349-760, arrays :820-860, _get_minmax_size :2934-2964, process_auto
:1870-1911, process :2090-2310 with the local-patch rule for LogMultipole,
_finalize :2312-2396, _apply_units :2890-2932; gggcorrelation.py:162-171,
nnncorrelation.py:130-149) over the tc_corr3 C ABI.
"""
import ctypes
import math
import sys

import numpy as np

from . import _lib
from .catalog import calculateVarG
from .config import parse_unit
from .util import set_omp_threads

LOGMULTIPOLE = 5
EUCLIDEAN = 1


class Corr3:
    _d = None
    _letters = None

    def __init__(self, config=None, *, logger=None, rng=None, **kwargs):
        cfg = dict(config or {})
        cfg.update(kwargs)
        self.config = cfg
        self.bin_type = cfg.get('bin_type', 'LogSAS')
        if self.bin_type != 'LogMultipole':
            raise NotImplementedError("treecorr_amd implements bin_type='LogMultipole' for "
                                      "3-point correlations")
        if cfg.get('max_n') is None:
            raise TypeError("Missing required parameter max_n")
        self.max_n = int(cfg['max_n'])
        if self.max_n < 0:
            raise ValueError("max_n must be non-negative")
        for key in ('min_u', 'max_u', 'nubins', 'ubin_size', 'min_v', 'max_v', 'nvbins',
                    'vbin_size', 'min_phi', 'max_phi', 'nphi_bins', 'phi_bin_size'):
            if key in cfg:
                raise TypeError("%s is invalid for bin_type=LogMultipole" % key)
        self.sep_units = cfg.get('sep_units', '')
        self._sep_units = parse_unit(cfg.get('sep_units') or 'radians')
        self._log_sep_units = math.log(self._sep_units)
        nbins, bin_size = cfg.get('nbins'), cfg.get('bin_size')
        min_sep, max_sep = cfg.get('min_sep'), cfg.get('max_sep')
        if min_sep is None or max_sep is None:
            raise TypeError("min_sep and max_sep are required")
        min_sep, max_sep = float(min_sep), float(max_sep)
        if min_sep >= max_sep:
            raise ValueError("max_sep must be larger than min_sep")
        if nbins is None:
            if bin_size is None:
                raise TypeError("Missing required parameter bin_size")
            nbins = int(math.ceil(math.log(max_sep / min_sep) / float(bin_size)))
        nbins = int(nbins)
        bin_size = math.log(max_sep / min_sep) / nbins
        self.nbins, self.bin_size, self.min_sep, self.max_sep = nbins, bin_size, min_sep, max_sep
        self._min_sep = min_sep * self._sep_units
        self._max_sep = max_sep * self._sep_units
        self._bin_size = bin_size
        self.split_method = cfg.get('split_method', 'mean')
        self.min_top = cfg.get('min_top')
        self.max_top = int(cfg.get('max_top', 10))
        # corr3base.py:714-745
        self.angle_slop = float(cfg.get('angle_slop', np.pi / (2 * self.max_n + 1)))
        bin_slop = cfg.get('bin_slop', -1.0)
        if bin_slop is None or bin_slop < 0:
            if self.bin_size <= 0.1:
                self.bin_slop = 1.0
                self.b = self.bin_size
            else:
                self.bin_slop = 0.1 / self.bin_size
                self.b = 0.1
        else:
            self.bin_slop = float(bin_slop)
            self.b = self._bin_size * self.bin_slop
        self.brute = bool(cfg.get('brute', False))
        self.var_method = cfg.get('var_method', 'shot')
        logr1d = np.linspace(0, nbins * bin_size, nbins, endpoint=False, dtype=float)
        logr1d += math.log(min_sep) + 0.5 * bin_size
        nn = 2 * self.max_n + 1
        self.logd2 = np.tile(logr1d[:, None, None], (1, nbins, nn))
        self.logd3 = np.tile(logr1d[None, :, None], (nbins, 1, nn))
        self.d2nom = np.exp(self.logd2)
        self.d3nom = np.exp(self.logd3)
        self.n1d = np.arange(-self.max_n, self.max_n + 1, dtype=int)
        self.n = np.tile(self.n1d[None, None, :], (nbins, nbins, 1))
        self.data_shape = self.logd2.shape
        for nm in ('meand1', 'meanlogd1', 'meand2', 'meanlogd2', 'meand3', 'meanlogd3', 'meanu',
                   'meanv', 'weightr', 'weighti', 'ntri'):
            setattr(self, nm, np.zeros(self.data_shape))
        self._z = [np.zeros(self.data_shape) if self._d == 4 else None for _ in range(8)]
        self._var_num = 0.
        self._corr = None
        self._corr_only = False
        self.coords = None
        self.metric = None
        self.counters = None

    @property
    def weight(self):
        return self.weightr + 1j * self.weighti

    @property
    def corr(self):
        if self._corr is None:
            h = ctypes.c_void_p()
            z = [_lib.dptr(a) if a is not None else None for a in self._z]
            _lib.check(_lib.lib().tc_corr3_create(
                self._d, LOGMULTIPOLE, self._min_sep, self._max_sep, self.nbins, self._bin_size,
                self.b, self.angle_slop, self.max_n, *z, _lib.dptr(self.meand1),
                _lib.dptr(self.meanlogd1), _lib.dptr(self.meand2), _lib.dptr(self.meanlogd2),
                _lib.dptr(self.meand3), _lib.dptr(self.meanlogd3), _lib.dptr(self.meanu),
                _lib.dptr(self.meanv), _lib.dptr(self.weightr), _lib.dptr(self.weighti),
                _lib.dptr(self.ntri), ctypes.byref(h)))
            self._corr = h
        return self._corr

    def __del__(self):
        h = getattr(self, '_corr', None)
        if h is not None and h.value:
            try:
                _lib.lib().tc_corr3_destroy(h)
            except Exception:
                pass

    def _set_metric(self, metric, coords):
        metric = metric or self.config.get('metric', 'Euclidean')
        if metric != 'Euclidean':
            raise NotImplementedError("treecorr_amd implements the Euclidean metric")
        if self.sep_units != '' and coords == '3d':
            raise ValueError("sep_units is invalid with 3d coordinates.")
        self.coords, self.metric = coords, metric

    def _get_minmax_size(self):
        # corr3base.py:2955-2960 (LogMultipole)
        b1 = min(self.angle_slop, self.b)
        return 2 * self._min_sep * b1 / (2. + 3 * b1), self._max_sep * b1

    def clear(self):
        for a in self._z:
            if a is not None:
                a[...] = 0
        for nm in ('meand1', 'meanlogd1', 'meand2', 'meanlogd2', 'meand3', 'meanlogd3', 'meanu',
                   'meanv', 'weightr', 'weighti', 'ntri'):
            getattr(self, nm)[...] = 0

    def process_auto(self, cat, *, metric=None, num_threads=None, corr_only=False):
        """Accumulate the auto-correlation of one catalog (corr3base.py:1870-1911)."""
        self._set_metric(metric, cat.coords)
        if num_threads is not None:
            set_omp_threads(num_threads)
        self._corr_only = corr_only
        min_size, max_size = self._get_minmax_size()
        getter = getattr(cat, 'get%sField' % self._letters[0])
        field = getter(min_size=min_size, max_size=max_size, split_method=self.split_method,
                       brute=self.brute, min_top=self.min_top, max_top=self.max_top,
                       coords=self.coords)
        _lib.check(_lib.lib().tc_corr3_process_auto(self.corr, field.data, 0, int(bool(corr_only)),
                                                     EUCLIDEAN))
        c = (ctypes.c_longlong * 4)()
        _lib.check(_lib.lib().tc_corr3_counters(self.corr, c))
        self.counters = dict(gn_calls=c[0], zeta_calls=c[1], zeta_rows=c[2], stages=c[3])

    def process(self, cat1, cat2=None, cat3=None, *, metric=None, ordered=True, num_threads=None,
                comm=None, low_mem=False, initialize=True, finalize=True, patch_method=None,
                algo=None, max_n=None, corr_only=False):
        """Compute the multipole correlation of one catalog (corr3base.py:2090-2310)."""
        if cat2 is not None or cat3 is not None:
            raise NotImplementedError("multipole cross-correlations are not implemented")
        if algo not in (None, 'multipole'):
            raise ValueError("LogMultipole binning requires algo='multipole'")
        if patch_method == 'global':
            raise ValueError("LogMultipole binning cannot use patch_method='global'")
        if comm is not None:
            raise NotImplementedError("MPI is not part of the single-GPU engine")
        if cat1.npatch != 1:
            raise NotImplementedError("patch-based multipole processing is not implemented")
        if initialize:
            self.clear()
            self._corr_only = corr_only
        self.process_auto(cat1, metric=metric, num_threads=num_threads, corr_only=corr_only)
        if finalize:
            self._do_finalize(cat1, corr_only)

    def _do_finalize(self, cat, corr_only):
        if corr_only:
            w = cat.sumw / cat.nobj
            self._meanwww = w ** 3
        self.finalize()

    def _apply_units(self, mask):
        if self.coords == 'spherical' and self.metric == 'Euclidean':
            self.meand2[mask] = 2. * np.arcsin(self.meand2[mask] / 2.)
            self.meanlogd2[mask] = np.log(2. * np.arcsin(np.exp(self.meanlogd2[mask]) / 2.))
            self.meand3[mask] = 2. * np.arcsin(self.meand3[mask] / 2.)
            self.meanlogd3[mask] = np.log(2. * np.arcsin(np.exp(self.meanlogd3[mask]) / 2.))
        self.meand2[mask] /= self._sep_units
        self.meanlogd2[mask] -= self._log_sep_units
        self.meand3[mask] /= self._sep_units
        self.meanlogd3[mask] -= self._log_sep_units

    def _finalize(self):
        # corr3base.py:2312-2396 (LogMultipole branches)
        mask1 = self.weightr != 0
        if self._corr_only:
            self.meand2[:] = self.d2nom
            self.meanlogd2[:] = self.logd2
            self.meand3[:] = self.d3nom
            self.meanlogd3[:] = self.logd3
            self.meanu[:] = 0
            self.meand1[:] = 0
            self.meanlogd1[:] = 0
            self.ntri[:] = self.weightr / self._meanwww
        else:
            for nm in ('meand2', 'meanlogd2', 'meand3', 'meanlogd3'):
                getattr(self, nm)[mask1] /= self.weightr[mask1]
            self._apply_units(mask1)
            mask2 = self.weightr == 0
            self.meand2[mask2] = self.d2nom[mask2]
            self.meanlogd2[mask2] = self.logd2[mask2]
            self.meand3[mask2] = self.d3nom[mask2]
            self.meanlogd3[mask2] = self.logd3[mask2]
            self.meanu[mask2] = 0
            self.meand1[mask2] = 0
            self.meanlogd1[mask2] = 0
            M = self.max_n
            for nm in ('ntri', 'meand2', 'meanlogd2', 'meand3', 'meanlogd3'):
                a = getattr(self, nm)
                a[:, :, :] = a[:, :, M][:, :, None]


class NNNCorrelation(Corr3):
    """Count-count-count multipole correlation (treecorr/nnncorrelation.py)."""
    _d = 0
    _letters = 'NNN'

    def __init__(self, config=None, *, logger=None, **kwargs):
        super().__init__(config, logger=logger, **kwargs)
        self.tot = 0.

    def process_auto(self, cat, **kw):
        super().process_auto(cat, **kw)
        self.tot += (1. / 6.) * cat.sumw ** 3

    def finalize(self):
        self._finalize()


class GGGCorrelation(Corr3):
    """Shear-shear-shear multipole correlation (treecorr/gggcorrelation.py)."""
    _d = 4
    _letters = 'GGG'

    gam0r = property(lambda self: self._z[0])
    gam0i = property(lambda self: self._z[1])
    gam1r = property(lambda self: self._z[2])
    gam1i = property(lambda self: self._z[3])
    gam2r = property(lambda self: self._z[4])
    gam2i = property(lambda self: self._z[5])
    gam3r = property(lambda self: self._z[6])
    gam3i = property(lambda self: self._z[7])
    gam0 = property(lambda self: self._z[0] + 1j * self._z[1])
    gam1 = property(lambda self: self._z[2] + 1j * self._z[3])
    gam2 = property(lambda self: self._z[4] + 1j * self._z[5])
    gam3 = property(lambda self: self._z[6] + 1j * self._z[7])

    def _do_finalize(self, cat, corr_only):
        if corr_only:
            w = cat.sumw / cat.nobj
            self._meanwww = w ** 3
        v = calculateVarG(cat)
        self.finalize(v, v, v)

    def finalize(self, varg1, varg2, varg3):
        # gggcorrelation.py:162-171
        self._finalize()
        self._var_num = 4 * varg1 * varg2 * varg3
