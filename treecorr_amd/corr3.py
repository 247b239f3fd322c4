"""3-point correlations on the GPU by the multipole algorithm: NNNCorrelation,
GGGCorrelation, with LogMultipole binning or LogSAS binning through toSAS.

Restates TreeCorr's 3-point front-end for the multipole algorithm (treecorr/corr3base.py:
binning :349-760 (LogSAS phi bins :652-690, :797-816), arrays :820-860,
_get_minmax_size :2934-2964, process_auto :1870-1911, process_cross12 :1913-1970,
process_cross21 :1972-2029, process_cross :2031-2088, process :2090-2310 with the
multipole default and the local-patch rule, _finalize :2312-2396, _apply_units
:2890-2932, _x_to_natural_phases :2489-2551, toSAS :2553-2677; gggcorrelation.py:162-178,
nnncorrelation.py:130-145, 220-227) over the tc_corr3 C ABI.  The toSAS Fourier sums
are FP64 products on the GPU (tc_cov_rowsums).  The triangle algorithm (LogRUV, or
algo='triangle') is not part of this engine.
"""
import ctypes
import math

import numpy as np

from . import _lib
from . import cov as _cov
from .catalog import calculateVarG
from .config import parse_unit
from .corr2 import Corr2
from .field import coord_enum
from .io import literal as _io_literal, make_minimal_config, make_reader, make_writer
from .util import set_omp_threads

LOGMULTIPOLE = 5
EUCLIDEAN = 1

_ARRAYS = ('meand1', 'meanlogd1', 'meand2', 'meanlogd2', 'meand3', 'meanlogd3', 'meanu',
           'meanv', 'weightr', 'weighti', 'ntri')


def _gemm(A, B):
    """A @ B in FP64 on the GPU (row-major host arrays)."""
    A = np.ascontiguousarray(A, dtype=float)
    B = np.ascontiguousarray(B, dtype=float)
    m, k = A.shape
    k2, n = B.shape
    assert k == k2
    C = np.zeros((m, n))
    _lib.check(_lib.lib().tc_cov_rowsums(m, k, n, _lib.dptr(A), _lib.dptr(B), _lib.dptr(C)))
    return C


# corr3base.py:480-520: the constructor parameters (and their defaults) kept in the
# params line of an output file
_VALID3 = {
    'nbins': None, 'bin_size': None, 'min_sep': None, 'max_sep': None, 'sep_units': None,
    'bin_slop': None, 'angle_slop': None, 'brute': False, 'verbose': 1, 'log_file': None,
    'output_dots': None, 'split_method': 'mean', 'min_top': None, 'max_top': 10,
    'precision': 4, 'metric': 'Euclidean', 'bin_type': 'LogSAS', 'min_phi': None,
    'max_phi': None, 'nphi_bins': None, 'phi_bin_size': None, 'phi_units': None,
    'max_n': None, 'var_method': 'shot', 'num_bootstrap': 500, 'cross_patch_weight': None,
    'num_threads': None,
}


def _spin(letter):
    return {'N': 0, 'K': 0, 'G': 2}[letter]


class Corr3:
    _d = None
    _letters = None

    def __init__(self, config=None, *, logger=None, rng=None, **kwargs):
        cfg = dict(config or {})
        cfg.update(kwargs)
        self.config = cfg
        self.bin_type = cfg.get('bin_type', 'LogSAS')
        if self.bin_type == 'LogRUV':
            raise NotImplementedError("bin_type='LogRUV' needs the triangle algorithm, which "
                                      "treecorr_amd does not implement")
        if self.bin_type not in ('LogSAS', 'LogMultipole'):
            raise ValueError("Invalid bin_type %s" % self.bin_type)
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
        cfg_bin_size = bin_size
        bin_size = math.log(max_sep / min_sep) / nbins
        self.nbins, self.bin_size, self.min_sep, self.max_sep = nbins, bin_size, min_sep, max_sep
        self._min_sep = min_sep * self._sep_units
        self._max_sep = max_sep * self._sep_units
        self._bin_size = bin_size
        for key in ('min_u', 'max_u', 'nubins', 'ubin_size', 'min_v', 'max_v', 'nvbins',
                    'vbin_size'):
            if key in cfg:
                raise TypeError("%s is invalid for bin_type=%s" % (key, self.bin_type))
        if self.bin_type == 'LogMultipole':
            for key in ('min_phi', 'max_phi', 'nphi_bins', 'phi_bin_size'):
                if key in cfg:
                    raise TypeError("%s is invalid for bin_type=LogMultipole" % key)
            if cfg.get('max_n') is None:
                raise TypeError("Missing required parameter max_n")
            self.max_n = int(cfg['max_n'])
            if self.max_n < 0:
                raise ValueError("max_n must be non-negative")
        else:
            self._init_phi(cfg, bin_size if cfg_bin_size is None else float(cfg_bin_size))
        self.split_method = cfg.get('split_method', 'mean')
        self.min_top = cfg.get('min_top')
        self.max_top = int(cfg.get('max_top', 10))
        # corr3base.py:714-745
        if self.bin_type == 'LogMultipole':
            self.angle_slop = float(cfg.get('angle_slop', np.pi / (2 * self.max_n + 1)))
        else:
            self.angle_slop = float(cfg.get('angle_slop', 0.1))
        bin_slop = cfg.get('bin_slop', -1.0)
        if bin_slop is None or bin_slop < 0:
            if self.bin_size <= 0.1:
                self.bin_slop = 1.0
                self.b = self.bin_size
            else:
                self.bin_slop = 0.1 / self.bin_size
                self.b = 0.1
            if self.bin_type == 'LogSAS':
                self.bu = self.phi_bin_size if self.phi_bin_size <= 0.1 else 0.1
            else:
                self.bu = 0.
        else:
            self.bin_slop = float(bin_slop)
            self.b = self._bin_size * self.bin_slop
            self.bu = self.phi_bin_size * self.bin_slop if self.bin_type == 'LogSAS' else 0.
        self.bv = 0.
        self.brute = bool(cfg.get('brute', False))
        self.var_method = cfg.get('var_method', 'shot')
        self.num_bootstrap = int(cfg.get('num_bootstrap', 500))
        self._rng = rng
        self.cross_patch_weight = cfg.get('cross_patch_weight')
        if self.cross_patch_weight not in (None, 'simple', 'mean', 'match', 'geom'):
            raise ValueError("Invalid cross_patch_weight %s" % self.cross_patch_weight)
        self._cov = None
        self._varzeta = None
        logr1d = np.linspace(0, nbins * bin_size, nbins, endpoint=False, dtype=float)
        logr1d += math.log(min_sep) + 0.5 * bin_size
        self.logr1d = logr1d
        self.rnom1d = np.exp(logr1d)
        if self.bin_type == 'LogMultipole':
            nn = 2 * self.max_n + 1
            self.n1d = np.arange(-self.max_n, self.max_n + 1, dtype=int)
            self.n = np.tile(self.n1d[None, None, :], (nbins, nbins, 1))
        else:
            nn = self.nphi_bins
            self.phi = np.tile(self.phi1d[None, None, :], (nbins, nbins, 1))
        self.logd2 = np.tile(logr1d[:, None, None], (1, nbins, nn))
        self.logd3 = np.tile(logr1d[None, :, None], (nbins, 1, nn))
        self.d2nom = np.exp(self.logd2)
        self.d3nom = np.exp(self.logd3)
        self.data_shape = self.logd2.shape
        for nm in _ARRAYS:
            setattr(self, nm, np.zeros(self.data_shape))
        if self.bin_type != 'LogMultipole':
            self.weighti = np.array([])       # corr3base.py:847-851: real weights for SAS
        self._z = [np.zeros(self.data_shape) if self._d == 4 else None for _ in range(8)]
        self._var_num = 0.
        self._corr = None
        self._corr_only = False
        self.coords = None
        self.metric = None
        self.counters = None
        self.results = {}
        self.npatch1 = self.npatch2 = self.npatch3 = 1

    def _init_phi(self, cfg, bin_size):
        # corr3base.py:652-690: phi in radians internally, phi_units for input and meanphi
        self.phi_units = cfg.get('phi_units', '')
        self._phi_units = parse_unit(cfg.get('phi_units') or 'radians')
        self.min_phi = float(cfg.get('min_phi', 0.)) * self._phi_units
        self.max_phi = float(cfg.get('max_phi', np.pi / self._phi_units)) * self._phi_units
        if self.min_phi < 0 or self.max_phi > np.pi:
            raise ValueError("Invalid range for phi: %f - %f" % (
                self.min_phi / self._phi_units, self.max_phi / self._phi_units))
        if self.min_phi >= self.max_phi:
            raise ValueError("max_phi must be larger than min_phi")
        self.phi_bin_size = float(cfg.get('phi_bin_size', bin_size))
        if 'nphi_bins' not in cfg:
            self.nphi_bins = int(math.ceil((self.max_phi - self.min_phi - 1.e-10) /
                                           self.phi_bin_size))
        elif 'max_phi' in cfg and 'min_phi' in cfg and 'phi_bin_size' in cfg:
            raise TypeError("Only 3 of min_phi, max_phi, phi_bin_size, and nphi_bins are "
                            "allowed.")
        else:
            self.nphi_bins = int(cfg['nphi_bins'])
            if 'phi_bin_size' in cfg:
                if 'max_phi' not in cfg:
                    self.max_phi = min(self.min_phi + self.nphi_bins * self.phi_bin_size, np.pi)
                else:
                    self.min_phi = max(self.max_phi - self.nphi_bins * self.phi_bin_size, 0.)
        self.phi_bin_size = (self.max_phi - self.min_phi) / self.nphi_bins
        self.phi1d = np.linspace(start=0, stop=self.nphi_bins * self.phi_bin_size,
                                 num=self.nphi_bins, endpoint=False)
        self.phi1d += self.min_phi + 0.5 * self.phi_bin_size

    @property
    def meanphi(self):
        assert self.bin_type == 'LogSAS'
        return self.meanu

    @property
    def weight(self):
        return self.weightr + 1j * self.weighti if self.weighti.size else self.weightr

    @property
    def _zetas(self):
        if self._d != 4:
            return []
        return [self._z[2 * i] + 1j * self._z[2 * i + 1] for i in range(4)]

    @property
    def corr(self):
        if self.bin_type != 'LogMultipole':
            raise TypeError("the C-ABI correlation object is the LogMultipole one")
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

    def _set_metric(self, metric, *coords):
        metric = metric or self.config.get('metric', 'Euclidean')
        if metric != 'Euclidean':
            raise NotImplementedError("treecorr_amd implements the Euclidean metric")
        c = coords[0]
        for other in coords[1:]:
            if other != c:
                raise ValueError("Cannot correlate catalogs with different coordinate systems.")
        if self.sep_units != '' and c == '3d':
            raise ValueError("sep_units is invalid with 3d coordinates.")
        self.coords, self.metric = c, metric

    def _get_minmax_size(self):
        # corr3base.py:2955-2960 (LogMultipole)
        b1 = min(self.angle_slop, self.b)
        return 2 * self._min_sep * b1 / (2. + 3 * b1), self._max_sep * b1

    def _clear(self):
        # corr3base.py:2398-2417: the accumulated arrays only
        self._cov = None
        self._varzeta = None
        for a in self._z:
            if a is not None:
                a[...] = 0
        for nm in _ARRAYS:
            a = getattr(self, nm)
            if a.size:
                a[...] = 0

    def clear(self):
        """Clear the sums, the per-patch results and the patch counts
        (corr3base.py:1254-1260)."""
        self._clear()
        self.results = {}
        self.npatch1 = self.npatch2 = self.npatch3 = 1

    def _field(self, cat, letter):
        min_size, max_size = self._get_minmax_size()
        getter = getattr(cat, 'get%sField' % letter)
        return getter(min_size=min_size, max_size=max_size, split_method=self.split_method,
                      brute=self.brute, min_top=self.min_top, max_top=self.max_top,
                      coords=self.coords)

    def _read_counters(self):
        c = (ctypes.c_longlong * 4)()
        _lib.check(_lib.lib().tc_corr3_counters(self.corr, c))
        self.counters = dict(gn_calls=c[0], zeta_calls=c[1], zeta_rows=c[2], stages=c[3])

    def _check_multipole(self, what):
        if self.bin_type != 'LogMultipole':
            raise TypeError("%s accumulates LogMultipole sums; LogSAS objects use process()"
                            % what)

    def process_auto(self, cat, *, metric=None, num_threads=None, corr_only=False):
        """Accumulate the auto-correlation of one catalog (corr3base.py:1870-1911)."""
        self._check_multipole('process_auto')
        self._set_metric(metric, cat.coords)
        if num_threads is not None:
            set_omp_threads(num_threads)
        self._corr_only = corr_only
        field = self._field(cat, self._letters[0])
        _lib.check(_lib.lib().tc_corr3_process_auto(self.corr, field.data, 0, int(bool(corr_only)),
                                                     EUCLIDEAN))
        self._read_counters()

    def process_cross12(self, cat1, cat2, *, metric=None, ordered=True, num_threads=None,
                        corr_only=False):
        """One point of each triangle from cat1, two from cat2 (corr3base.py:1913-1970)."""
        self._check_multipole('process_cross12')
        self._set_metric(metric, cat1.coords, cat2.coords)
        if num_threads is not None:
            set_omp_threads(num_threads)
        self._corr_only = corr_only
        f1 = self._field(cat1, self._letters[0])
        f2 = self._field(cat2, self._letters[1])
        ocode = 1 if ordered or self._letters[0] != self._letters[1] else 0
        _lib.check(_lib.lib().tc_corr3_process_cross12(self.corr, f1.data, f2.data, ocode, 0,
                                                        int(bool(corr_only)), EUCLIDEAN))
        self._read_counters()

    def process_cross21(self, cat1, cat2, *, metric=None, ordered=True, num_threads=None,
                        corr_only=False):
        """Two points of each triangle from cat1, one from cat2 (corr3base.py:1972-2029)."""
        self._check_multipole('process_cross21')
        self._set_metric(metric, cat1.coords, cat2.coords)
        if num_threads is not None:
            set_omp_threads(num_threads)
        self._corr_only = corr_only
        f1 = self._field(cat1, self._letters[0])
        f2 = self._field(cat2, self._letters[2])
        ocode = 3 if ordered or self._letters[1] != self._letters[2] else 0
        _lib.check(_lib.lib().tc_corr3_process_cross21(self.corr, f1.data, f2.data, ocode, 0,
                                                        int(bool(corr_only)), EUCLIDEAN))
        self._read_counters()

    def process_cross(self, cat1, cat2, cat3, *, metric=None, ordered=True, num_threads=None,
                      corr_only=False):
        """One point of each triangle from each catalog (corr3base.py:2031-2088)."""
        self._check_multipole('process_cross')
        self._set_metric(metric, cat1.coords, cat2.coords, cat3.coords)
        if num_threads is not None:
            set_omp_threads(num_threads)
        self._corr_only = corr_only
        f1 = self._field(cat1, self._letters[0])
        f2 = self._field(cat2, self._letters[1])
        f3 = self._field(cat3, self._letters[2])
        ocode = 4 if ordered is True else 0 if ordered is False else int(ordered)
        _lib.check(_lib.lib().tc_corr3_process_cross(self.corr, f1.data, f2.data, f3.data, ocode,
                                                      0, int(bool(corr_only)), EUCLIDEAN))
        self._read_counters()

    def process(self, cat1, cat2=None, cat3=None, *, metric=None, ordered=True, num_threads=None,
                comm=None, low_mem=False, initialize=True, finalize=True, patch_method=None,
                algo=None, max_n=None, corr_only=False):
        """Compute the correlation of one, two or three catalogs (corr3base.py:2090-2310)."""
        if algo is None:
            multipole = True
        else:
            if algo not in ('triangle', 'multipole'):
                raise ValueError("Invalid algo %s" % algo)
            multipole = algo == 'multipole'
        if not multipole:
            raise NotImplementedError("treecorr_amd computes 3-point correlations with the "
                                      "multipole algorithm only")
        if comm is not None:
            raise NotImplementedError("MPI is not part of the single-GPU engine")
        if self.bin_type != 'LogMultipole':
            # corr3base.py:2190-2206: a LogMultipole run, converted by toSAS
            config = self.config.copy()
            config['bin_type'] = 'LogMultipole'
            if max_n is None:
                max_n = 2. * np.pi / self.phi_bin_size
                max_n = config.get('max_n', max_n)
            for key in ('min_phi', 'max_phi', 'nphi_bins', 'phi_bin_size'):
                config.pop(key, None)
            config['max_n'] = max_n
            corr = self.__class__(config)
            corr.process(cat1, cat2, cat3, metric=metric, ordered=ordered,
                         num_threads=num_threads, corr_only=corr_only, low_mem=low_mem,
                         initialize=initialize, finalize=finalize, patch_method=patch_method,
                         algo='multipole')
            corr.toSAS(target=self)
            if corr_only:
                self._meanwww = corr._meanwww
                self.ntri[:] = np.abs(self.weight) / self._meanwww
            return
        if patch_method == 'global':
            raise ValueError("LogMultipole binning cannot use patch_method='global'")
        if patch_method not in (None, 'local', 'global'):
            raise ValueError("Invalid patch_method %s" % patch_method)
        if not isinstance(cat1, list):
            cat1 = cat1.get_patches()
        if cat2 is not None and not isinstance(cat2, list):
            cat2 = cat2.get_patches()
        if cat3 is not None and not isinstance(cat3, list):
            cat3 = cat3.get_patches()
        single = all(c is None or (len(c) == 1 and c[0].npatch == 1) for c in (cat1, cat2, cat3))
        if initialize:
            self.clear()
            self._corr_only = corr_only
        kw = dict(metric=metric, num_threads=num_threads, corr_only=corr_only)
        if cat2 is None:
            if cat3 is not None:
                raise ValueError("For two catalog case, use cat1,cat2, not cat1,cat3")
            if not (ordered is True or ordered is False):
                raise ValueError("The integer options for ordered are only valid with 3 catalogs")
            if single:
                self.process_auto(cat1[0], **kw)
            else:
                self._local_auto(cat1, kw)
        elif cat3 is None:
            if not (ordered is True or ordered is False):
                raise ValueError("The integer options for ordered are only valid with 3 catalogs")
            if self._letters[1] == self._letters[2]:
                if single:
                    self.process_cross12(cat1[0], cat2[0], ordered=ordered, **kw)
                else:
                    self._local_cross12(cat1, cat2, ordered, kw)
            elif single:
                self.process_cross21(cat1[0], cat2[0], ordered=ordered, **kw)
            else:
                self._local_cross21(cat1, cat2, ordered, kw)
        elif single:
            self.process_cross(cat1[0], cat2[0], cat3[0], ordered=ordered, **kw)
        else:
            self._local_cross(cat1, cat2, cat3, ordered, kw)
        if finalize:
            self._do_finalize(cat1, cat2, cat3, corr_only)

    # ---- patches: the local rule of the multipole algorithm ----
    # Each patch's triangles with their first vertex in the patch and the other two among
    # the nearby objects of the other catalogs (the expanded patch), stored as results[(i,
    # i, i)] and added to the totals (corr3base.py:1287-1351, 1388-1421, 1533-1563,
    # 1660-1680, 1760-1818).

    _make_expanded_patch = Corr2._make_expanded_patch

    def copy(self):
        """A copy with its own arrays (and its own engine object, made on use);
        corr3base.py:1219-1233."""
        ret = self.__class__.__new__(self.__class__)
        for key, item in self.__dict__.items():
            ret.__dict__[key] = item.copy() if isinstance(item, np.ndarray) else item
        ret._z = [None if a is None else a.copy() for a in self._z]
        ret.results = dict(self.results)
        ret._corr = None
        return ret

    @property
    def nonzero(self):
        """Whether any weight is accumulated (corr3base.py:1263-1266)."""
        return bool(np.any(self.weight))

    def _sums(self):
        """The accumulated arrays (the zetas of the class and the _ARRAYS in use)."""
        return [a for a in self._z if a is not None] + \
            [getattr(self, nm) for nm in _ARRAYS if getattr(self, nm).size]

    def _same_binning(self, other):
        return (self.bin_type == other.bin_type and self.nbins == other.nbins and
                self.min_sep == other.min_sep and self.max_sep == other.max_sep and
                self.data_shape == other.data_shape and self.sep_units == other.sep_units)

    def __iadd__(self, other):
        # corr3base.py:2419-2449
        if not isinstance(other, self.__class__):
            raise TypeError("Can only add another %s object" % type(self).__name__)
        if not self._same_binning(other):
            raise ValueError("%s to be added is not compatible with this one."
                             % type(self).__name__)
        if other.coords is not None:
            if self.coords is not None and self.coords != other.coords:
                raise ValueError("Cannot correlate catalogs with different coordinate systems.")
            self.coords, self.metric = other.coords, other.metric
        if not other.nonzero:
            return self
        for a, b in zip(self._sums(), other._sums()):
            a += b
        return self

    def _trivially_zero(self, c1, c2, c3, ordered):
        # corr3base.py:1277-1283
        x1, y1, z1, s1 = c1._get_center_size()
        x2, y2, z2, s2 = c2._get_center_size()
        x3, y3, z3, s3 = c3._get_center_size()
        r = ctypes.c_int()
        _lib.check(_lib.lib().tc_corr3_trivially_zero(
            self.corr, coord_enum(self.coords), x1, y1, z1, s1, x2, y2, z2, s2, x3, y3, z3, s3,
            int(ordered), int(c2 is c3 or c1 is c2), ctypes.byref(r)))
        return bool(r.value)

    def _single_process(self, c1, c2, c3, ijk, ordered, temp, kw):
        # corr3base.py:1287-1351 with force_write, as every local call has it: one vertex
        # from c1 and two from c2 (c2 is c3) or one each from c2 and c3
        temp._clear()
        if c2 is not None and c3 is not None and not self._trivially_zero(c1, c2, c3, ordered):
            if c2 is c3:
                temp.process_cross12(c1, c2, ordered=ordered, **kw)
            else:
                temp.process_cross(c1, c2, c3, ordered=ordered, **kw)
        if ijk in self.results and self.results[ijk].nonzero:
            self.results[ijk] += temp
        else:
            self.results[ijk] = temp.copy()
        self += temp

    @staticmethod
    def _npatch_of(cats):
        return cats[0].npatch if cats[0].npatch != 1 else len(cats)

    @staticmethod
    def _patch_index(c, ii):
        return c._single_patch if c._single_patch is not None else ii

    def _local_temp(self, coords, metric):
        self._set_metric(metric, *coords)
        temp = self.copy()
        temp.results = {}
        return temp

    def _local_auto(self, cat1, kw):
        # corr3base.py:1392-1421
        if self.npatch1 == 1:
            self.npatch1 = self.npatch2 = self.npatch3 = self._npatch_of(cat1)
        temp = self._local_temp([cat1[0].coords], kw['metric'])
        for ii, c1 in enumerate(cat1):
            i = self._patch_index(c1, ii)
            c1e = self._make_expanded_patch(c1, cat1)
            self._single_process(c1, c1e, c1e, (i, i, i), 1, temp, kw)

    def _local_cross12(self, cat1, cat2, ordered, kw):
        # corr3base.py:1507-1563
        if self.npatch1 == 1:
            self.npatch1 = self._npatch_of(cat1)
        if self.npatch2 == 1:
            self.npatch2 = self.npatch3 = self._npatch_of(cat2)
        if self.npatch1 != self.npatch2 and self.npatch1 != 1 and self.npatch2 != 1:
            raise RuntimeError("Cross correlation requires both catalogs use the same patches.")
        temp = self._local_temp([cat1[0].coords, cat2[0].coords], kw['metric'])
        for ii, c1 in enumerate(cat1):
            i = self._patch_index(c1, ii)
            c2e = self._make_expanded_patch(c1, cat2)
            self._single_process(c1, c2e, c2e, (i, i, i), 1, temp, kw)
        if not ordered and self._letters[0] == self._letters[1]:
            # the local rule runs with ordered=1 only: the triangles with the first vertex
            # from cat2 are a second pass
            for ii, c2 in enumerate(cat2):
                i = self._patch_index(c2, ii)
                c1e = self._make_expanded_patch(c2, cat1)
                c2e = self._make_expanded_patch(c2, cat2)
                self._single_process(c2, c1e, c2e, (i, i, i), 1, temp, kw)

    def _local_cross21(self, cat1, cat2, ordered, kw):
        # corr3base.py:1634-1680
        if self.npatch1 == 1:
            self.npatch1 = self.npatch2 = self._npatch_of(cat1)
        if self.npatch3 == 1:
            self.npatch3 = self._npatch_of(cat2)
        if self.npatch1 != self.npatch3 and self.npatch1 != 1 and self.npatch3 != 1:
            raise RuntimeError("Cross correlation requires both catalogs use the same patches.")
        temp = self._local_temp([cat1[0].coords, cat2[0].coords], kw['metric'])
        for ii, c1 in enumerate(cat1):
            i = self._patch_index(c1, ii)
            c2e = self._make_expanded_patch(c1, cat1)
            c3e = self._make_expanded_patch(c1, cat2)
            self._single_process(c1, c2e, c3e, (i, i, i), 1 if not ordered else 4, temp, kw)

    def _local_cross(self, cat1, cat2, cat3, ordered, kw):
        # corr3base.py:1738-1818
        for k, cats in enumerate((cat1, cat2, cat3)):
            nm = 'npatch%d' % (k + 1)
            if getattr(self, nm) == 1:
                setattr(self, nm, self._npatch_of(cats))
        n1, n2, n3 = self.npatch1, self.npatch2, self.npatch3
        if (n1 != n2 and n1 != 1 and n2 != 1) or (n1 != n3 and n1 != 1 and n3 != 1):
            raise RuntimeError("Cross correlation requires all catalogs use the same patches.")
        temp = self._local_temp([cat1[0].coords, cat2[0].coords, cat3[0].coords], kw['metric'])
        ordered = 0 if ordered is False else 4 if ordered is True else ordered
        for ii, c1 in enumerate(cat1):
            i = self._patch_index(c1, ii)
            c2e = self._make_expanded_patch(c1, cat2)
            c3e = self._make_expanded_patch(c1, cat3)
            self._single_process(c1, c2e, c3e, (i, i, i), 1 if ordered in (0, 1) else 4, temp,
                                 kw)
        # the local rule runs with ordered 1 or 4 only: the other first vertices are passes
        # of their own
        o = 1 if ordered == 0 else 4
        if ordered in (0, 3) and self._letters[0] == self._letters[1]:
            for ii, c2 in enumerate(cat2):
                i = self._patch_index(c2, ii)
                c1e = self._make_expanded_patch(c2, cat1)
                c3e = self._make_expanded_patch(c2, cat3)
                self._single_process(c2, c1e, c3e, (i, i, i), o, temp, kw)
        if ordered in (0, 2) and self._letters[0] == self._letters[2]:
            for ii, c3 in enumerate(cat3):
                i = self._patch_index(c3, ii)
                c1e = self._make_expanded_patch(c3, cat1)
                c2e = self._make_expanded_patch(c3, cat2)
                self._single_process(c3, c2e, c1e, (i, i, i), o, temp, kw)

    def _meanw(self, cat):
        # catalog.py calculateMeanW: the weight per object over all patches
        cats = cat if isinstance(cat, list) else [cat]
        return np.sum([c.sumw for c in cats]) / np.sum([c.nobj for c in cats])

    def _do_finalize(self, cat1, cat2, cat3, corr_only):
        # corr3base.py:2252-2310
        if corr_only:
            w1 = self._meanw(cat1)
            if cat2 is None:
                w2 = w3 = w1
            elif cat3 is None:
                w2 = self._meanw(cat2)
                w3 = w2 if self._letters[1] == self._letters[2] else w1
            else:
                w2, w3 = self._meanw(cat2), self._meanw(cat3)
            self._meanwww = w1 * w2 * w3
        self._finalize_with(cat1, cat2, cat3)

    def _finalize_with(self, cat1, cat2, cat3):
        self.finalize()

    def _apply_units(self, mask):
        # corr3base.py:2896-2929
        sas = self.bin_type == 'LogSAS'
        if self.coords == 'spherical' and self.metric == 'Euclidean':
            if sas:
                # cos of the great-circle angle from the chord triangle's
                cosphi = np.cos(self.meanu[mask])
                cosphi -= 0.25 * self.meand2[mask] * self.meand3[mask]
                cosphi /= np.sqrt((1 - self.meand2[mask] ** 2 / 4) *
                                  (1 - self.meand3[mask] ** 2 / 4))
                self.meanu[mask] = np.arccos(np.clip(cosphi, -1, 1))
                self.meand1[mask] = 2. * np.arcsin(self.meand1[mask] / 2.)
                self.meanlogd1[mask] = np.log(2. * np.arcsin(np.exp(self.meanlogd1[mask]) / 2.))
            self.meand2[mask] = 2. * np.arcsin(self.meand2[mask] / 2.)
            self.meanlogd2[mask] = np.log(2. * np.arcsin(np.exp(self.meanlogd2[mask]) / 2.))
            self.meand3[mask] = 2. * np.arcsin(self.meand3[mask] / 2.)
            self.meanlogd3[mask] = np.log(2. * np.arcsin(np.exp(self.meanlogd3[mask]) / 2.))
        if sas:
            self.meanu[mask] /= self._phi_units
            self.meand1[mask] /= self._sep_units
            self.meanlogd1[mask] -= self._log_sep_units
        self.meand2[mask] /= self._sep_units
        self.meanlogd2[mask] -= self._log_sep_units
        self.meand3[mask] /= self._sep_units
        self.meanlogd3[mask] -= self._log_sep_units

    def _finalize(self):
        # corr3base.py:2312-2396: LogMultipole keeps its zetas unnormalised and broadcasts
        # the distances from n = 0; LogSAS (the sums of patch results) divides everything
        # by the weight
        mask1 = self.weightr != 0
        sas = self.bin_type == 'LogSAS'
        if sas:
            for a in self._z:
                if a is not None:
                    a[mask1] /= self.weightr[mask1]
        if self._corr_only:
            self.meand2[:] = self.d2nom
            self.meanlogd2[:] = self.logd2
            self.meand3[:] = self.d3nom
            self.meanlogd3[:] = self.logd3
            if sas:
                self.meanu[:] = self.phi
                self.meand1[:] = np.sqrt(self.d2nom ** 2 + self.d3nom ** 2
                                         - 2 * self.d2nom * self.d3nom * np.cos(self.phi))
                self.meanlogd1[:] = np.log(self.meand1)
            else:
                self.meanu[:] = 0
                self.meand1[:] = 0
                self.meanlogd1[:] = 0
            self.ntri[:] = self.weightr / self._meanwww
            return
        names = ('meand2', 'meanlogd2', 'meand3', 'meanlogd3')
        if sas:
            names += ('meand1', 'meanlogd1', 'meanu')
        for nm in names:
            getattr(self, nm)[mask1] /= self.weightr[mask1]
        self._apply_units(mask1)
        mask2 = self.weightr == 0
        self.meand2[mask2] = self.d2nom[mask2]
        self.meanlogd2[mask2] = self.logd2[mask2]
        self.meand3[mask2] = self.d3nom[mask2]
        self.meanlogd3[mask2] = self.logd3[mask2]
        if sas:
            self.meanu[mask2] = self.phi[mask2]
            self.meand1[mask2] = np.sqrt(self.d2nom[mask2] ** 2 + self.d3nom[mask2] ** 2
                                         - 2 * self.d2nom[mask2] * self.d3nom[mask2] *
                                         np.cos(self.phi[mask2]))
            self.meanlogd1[mask2] = np.log(self.meand1[mask2])
        else:
            self.meanu[mask2] = 0
            self.meand1[mask2] = 0
            self.meanlogd1[mask2] = 0
            M = self.max_n
            for nm in ('ntri', 'meand2', 'meanlogd2', 'meand3', 'meanlogd3'):
                a = getattr(self, nm)
                a[:, :, :] = a[:, :, M][:, :, None]

    # ---- covariances (corr3base.py:1085-1110, 2451-2487, 2678-2850, 2966-3350) ----
    # With the multipole algorithm every patch result is a local one, (i, i, i), and all
    # catalogs carry the same patches, so the triple lists below are the reference's for
    # npatch1 == npatch2 == npatch3.

    @property
    def rng(self):
        if self._rng is None:
            self._rng = np.random.default_rng()
        return self._rng

    def _row_arrays(self, corr_only):
        """The arrays _sum adds, in a fixed order (corr3base.py:2451-2481)."""
        arrs = [a for a in self._z if a is not None] + [self.weightr]
        if self.weighti.size:
            arrs.append(self.weighti)
        if not corr_only:
            arrs += [self.meand1, self.meanlogd1, self.meand2, self.meanlogd2, self.meand3,
                     self.meanlogd3, self.meanu, self.ntri]
        return arrs

    def _sum(self, others, corr_only):
        # corr3base.py:2451-2481: self = sum of w * other
        if _cov._SUM_HOOK is not None:      # a design matrix's rows, formed on the GPU
            _cov._SUM_HOOK(self, others, corr_only)
        else:
            cols = [c._row_arrays(corr_only) for c, _ in others]
            for q, a in enumerate(self._row_arrays(corr_only)):
                np.sum([b[q] if w == 1 else b[q] * w for b, (_, w) in zip(cols, others)],
                       axis=0, out=a)
        self._cov = None
        self._varzeta = None

    def _get_npatch(self):
        return max(self.npatch1, self.npatch2, self.npatch3)

    def _calculate_xi_from_pairs(self, pairs, corr_only):
        # corr3base.py:2969-2975 ("pairs" are triples (i, j, k, w))
        self._sum([(self.results[(i, j, k)], w) for i, j, k, w in pairs], corr_only)
        self._finalize()

    @property
    def _ok(self):
        ok = np.zeros((self.npatch1, self.npatch2, self.npatch3), dtype=bool)
        for (i, j, k) in self.results:
            ok[i, j, k] = True
        return ok

    def _keep_ok(self, pairs):
        ok = self._ok
        return [(i, j, k, w) for i, j, k, w in pairs if ok[i, j, k] and w != 0]

    def _check_cpw(self, method, cross_patch_weight):
        # corr3base.py:3332-3349 (the default's warning is not repeated here)
        if cross_patch_weight is None:
            cross_patch_weight = self.cross_patch_weight
        return 'simple' if cross_patch_weight is None else cross_patch_weight

    def _patch_keys(self):
        n = self._get_npatch()
        if not (self.npatch1 == self.npatch2 == self.npatch3 == n):
            raise NotImplementedError("3-point covariances need the same patches in every "
                                      "catalog")
        return sorted(self.results)

    def _jackknife_pairs(self, cross_patch_weight):
        # corr3base.py:2997-3097 (all three patch indices in play)
        cpw = cross_patch_weight
        if cpw not in ('simple', 'mean', 'match'):
            raise ValueError(f"cross_patch_weight = {cpw} is invalid for jackknife")
        keys = self._patch_keys()
        n = self._get_npatch()
        w1 = 1 - n / (2 + 2 ** 0.5 * (n - 1))
        w2 = 1 - 2 ** 0.5 * n / (3 * (n - 1 + 2 ** 0.5))
        rows = []
        for m in range(n):
            row = []
            for key in keys:
                out = sum(p != m for p in key)
                if cpw == 'simple':
                    if out == 3:
                        row.append((*key, 1))
                elif out > 0:
                    if cpw == 'mean':
                        row.append((*key, out / 3))
                    else:
                        i, j, k = key
                        row.append((*key, 1 if out == 3 else
                                    w2 if (i != j and i != k and j != k) else w1))
            rows.append(row)
        return rows

    def _sample_pairs(self, cross_patch_weight):
        # corr3base.py:3099-3155
        cpw = cross_patch_weight
        if cpw not in ('simple', 'mean'):
            raise ValueError(f"cross_patch_weight = {cpw} is invalid for sample")
        keys = self._patch_keys()
        rows = []
        for m in range(self._get_npatch()):
            if cpw == 'simple':
                rows.append([(*key, 1) for key in keys if key[0] == m])
            else:
                rows.append([(*key, sum(p == m for p in key) / 3) for key in keys
                             if m in key])
        return rows

    def _marked_pairs(self, index, cross_patch_weight):
        # corr3base.py:3162-3233: the drawn patches as first vertices, each as often as
        # drawn
        cpw = cross_patch_weight
        if cpw not in ('simple', 'mean'):
            raise ValueError(f"cross_patch_weight = {cpw} is invalid for marked_bootstrap")
        keys = self._patch_keys()
        index, counts = np.unique(index, return_counts=True)
        out = []
        for i, w in zip(index, counts):
            for key in keys:
                if key[0] == i:
                    out.append((*key, w if cpw == 'simple' else sum(p == i for p in key) / 3 * w))
        return out

    def _bootstrap_pairs(self, index, cross_patch_weight):
        # corr3base.py:3235-3330: each (i, i, i) once per draw of i, then the triples with
        # patches that differ, weighted by the draws of their patches
        cpw = cross_patch_weight
        if cpw not in ('simple', 'mean', 'geom'):
            raise ValueError(f"cross_patch_weight = {cpw} is invalid for bootstrap")
        keys = self._patch_keys()
        ok = set(keys)
        index, counts = np.unique(index, return_counts=True)
        cnt = dict(zip(index, counts))
        out = [(i, i, i, w) for i, w in zip(index, counts) if (i, i, i) in ok]
        for (i, j, k) in keys:
            if i == j == k:
                continue
            if cpw == 'mean':
                out.append((i, j, k, (cnt.get(i, 0) + cnt.get(j, 0) + cnt.get(k, 0)) / 3))
            elif i in cnt and j in cnt and k in cnt:
                wi, wj, wk = cnt[i], cnt[j], cnt[k]
                if cpw == 'simple':
                    w = wi * wj if (i == k or j == k) else wi * wk if i == j else wi * wj * wk
                else:
                    w = (wi * wj * wk) ** (1. / 3.)
                out.append((i, j, k, w))
        return out

    def estimate_cov(self, method, *, func=None, comm=None, num_bootstrap=None,
                     cross_patch_weight=None):
        """Covariance estimate: 'shot', 'jackknife', 'sample', 'bootstrap' or
        'marked_bootstrap' (corr3base.py:2678-2790)."""
        if num_bootstrap is None:
            num_bootstrap = self.num_bootstrap
        all_func = None if func is None else (lambda corrs: func(corrs[0]))
        return _cov.estimate_multi_cov([self], method, func=all_func, comm=comm,
                                       num_bootstrap=num_bootstrap,
                                       cross_patch_weight=cross_patch_weight)

    def build_cov_design_matrix(self, method, *, func=None, comm=None, num_bootstrap=None,
                                cross_patch_weight=None):
        """The design matrix behind a covariance estimate (corr3base.py:2792-2850)."""
        if num_bootstrap is None:
            num_bootstrap = self.num_bootstrap
        all_func = None if func is None else (lambda corrs: func(corrs[0]))
        return _cov.build_multi_cov_design_matrix([self], method, func=all_func, comm=comm,
                                                  num_bootstrap=num_bootstrap,
                                                  cross_patch_weight=cross_patch_weight)

    @property
    def cov(self):
        # corr3base.py:1090-1099
        if self._cov is None:
            self._cov = self.estimate_cov(self.var_method)
        return np.diag(self._cov) if self._cov.ndim == 1 else self._cov

    @property
    def cov_diag(self):
        # corr3base.py:1101-1112
        if self._cov is None:
            self._cov = self.estimate_cov(self.var_method)
        return self._cov if self._cov.ndim == 1 else self._cov.diagonal()

    def getStat(self):
        """The statistic as a 1-d array (corr3base.py:1849-1858)."""
        return self.zeta.ravel()

    def getWeight(self):
        """The weights of getStat, |weight| for the complex multipole weight
        (corr3base.py:1860-1868)."""
        return np.abs(self.weight.ravel())

    # ---- output files (corr3base.py:3352-3662) ----

    @property
    def _write_params(self):
        params = make_minimal_config(self.config, _VALID3)
        params['coords'] = self.coords
        params['metric'] = self.metric
        params['corr'] = self._letters
        return params

    @property
    def _write_col_names(self):
        sas = self.bin_type == 'LogSAS'
        names = ['d2_nom', 'd3_nom', 'phi_nom' if sas else 'n', 'meand1', 'meanlogd1',
                 'meand2', 'meanlogd2', 'meand3', 'meanlogd3'] + (['meanphi'] if sas else [])
        return names + list(self._write_class_col_names) + self._write_weight_names()

    @property
    def _write_data(self):
        sas = self.bin_type == 'LogSAS'
        data = [self.d2nom, self.d3nom, self.phi if sas else self.n, self.meand1,
                self.meanlogd1, self.meand2, self.meanlogd2, self.meand3, self.meanlogd3]
        if sas:
            data.append(self.meanu)
        data += list(self._write_class_data) + self._write_weight_data()
        return [np.asarray(c, dtype=float).ravel() for c in data]

    _write_class_col_names = ()
    _write_class_data = ()

    def _write_weight_names(self):
        if self.bin_type == 'LogMultipole':
            return ['weightr', 'weighti', 'ntri']
        return ['weight', 'ntri']

    def _write_weight_data(self):
        if self.bin_type == 'LogMultipole':
            return [self.weightr, self.weighti, self.ntri]
        return [self.weightr, self.ntri]

    def write(self, file_name, *, file_type=None, precision=None, write_patch_results=False,
              write_cov=False):
        """Write the correlation to a file (ASCII; corr3base.py:3449-3530).  With
        write_patch_results every patch result follows at 16 digits, so that the
        covariances of the object read back equal this one's."""
        precision = self.config.get('precision', 4) if precision is None else precision
        with make_writer(file_name, precision, file_type) as writer:
            self._write(writer, None, write_patch_results, write_cov=write_cov)

    def _write(self, writer, name, write_patch_results, write_cov=False, zero_tot=False):
        # corr3base.py:3403-3447
        if name is None and (write_patch_results or write_cov):
            name = 'main'
        col_names, data = self._write_col_names, self._write_data
        params = self._write_params
        if write_patch_results:
            params['npatch1'] = self.npatch1
            params['npatch2'] = self.npatch2
            params['npatch3'] = self.npatch3
            params['num_rows'] = int(np.prod(self.data_shape))
            num_patch_tri = len(self.results)
            if zero_tot:
                i = 0
                for key, corr in self.results.items():
                    if not corr.nonzero:
                        params[name + '_zp_%d' % i] = repr((key, float(corr.tot)))
                        num_patch_tri -= 1
                        i += 1
                params['num_zero_patch'] = i
            params['num_patch_tri'] = num_patch_tri
        if write_cov:
            params['cov_shape'] = self.cov.shape
        writer.write(col_names, data, params=params, ext=name)
        if write_patch_results:
            writer.set_precision(16)
            i = 0
            for key, corr in self.results.items():
                if zero_tot and not corr.nonzero:
                    continue
                pparams = corr._write_params
                pparams['key'] = repr(tuple(int(k) for k in key))
                writer.write(corr._write_col_names, corr._write_data, params=pparams,
                             ext=name + '_pp_%d' % i)
                i += 1
        if write_cov:
            writer.write_array(self.cov, ext='cov')

    def read(self, file_name, *, file_type=None):
        """Read values written by write() into this object, which must have the same
        configuration (corr3base.py:3600-3622)."""
        with make_reader(file_name, file_type) as reader:
            self._do_read(reader)

    def _do_read(self, reader, name=None, params=None):
        self._read(reader, name, params)

    @classmethod
    def from_file(cls, file_name, *, file_type=None, logger=None, rng=None):
        """A new object from a file written by write() (corr3base.py:3624-3662); from the
        base class, the class the file names."""
        with make_reader(file_name, file_type) as reader:
            params = reader.read_params(ext=None)
            letters = params.get('corr', None)
            if cls is Corr3:
                kinds = {'NNN': NNNCorrelation, 'GGG': GGGCorrelation}
                if letters not in kinds:
                    raise OSError("%s does not seem to be a valid treecorr output file."
                                  % file_name)
                cls = kinds[letters]
            elif letters != cls._letters:
                raise OSError("Trying to read a %sCorrelation output file with %s"
                              % (letters, cls.__name__))
            corr = cls(make_minimal_config(params, _VALID3), rng=rng)
            corr._do_read(reader, name=None, params=params)
        return corr

    def _read(self, reader, name=None, params=None):
        # corr3base.py:3532-3570
        if params is None:
            params = reader.read_params(ext=name)
        num_rows = params.get('num_rows', None)
        num_patch_tri = params.get('num_patch_tri', 0)
        num_zero_patch = params.get('num_zero_patch', 0)
        cov_shape = params.get('cov_shape', None)
        name = 'main' if num_patch_tri and name is None else name
        self._read_from_data(reader.read_data(max_rows=num_rows, ext=name), params)
        self.results = {}
        for i in range(num_zero_patch):
            key, tot = _io_literal(params[name + '_zp_%d' % i])
            self.results[key] = self._zero_copy(tot)
        for i in range(num_patch_tri):
            pname = name + '_pp_%d' % i
            corr = self.copy()
            corr.results = {}
            pparams = reader.read_params(ext=pname)
            corr._read_from_data(reader.read_data(max_rows=num_rows, ext=pname), pparams)
            self.results[_io_literal(pparams['key'])] = corr
        if cov_shape is not None:
            if isinstance(cov_shape, str):
                cov_shape = _io_literal(cov_shape)
            self._cov = reader.read_array(tuple(cov_shape), ext='cov')

    def _read_from_data(self, data, params):
        # corr3base.py:3572-3598; in place, so the engine object keeps valid pointers
        s = self.data_shape
        for nm in ('meand1', 'meanlogd1', 'meand2', 'meanlogd2', 'meand3', 'meanlogd3',
                   'ntri'):
            getattr(self, nm)[...] = data[nm].reshape(s)
        if self.bin_type == 'LogSAS':
            self.meanu[...] = data['meanphi'].reshape(s)
        if self.bin_type == 'LogMultipole':
            self.weightr[...] = data['weightr'].reshape(s)
            self.weighti[...] = data['weighti'].reshape(s)
        elif 'weight' in data.dtype.names:
            self.weightr[...] = data['weight'].reshape(s)
        self.coords = params['coords'].strip() if params.get('coords') else None
        self.metric = params['metric'].strip() if params.get('metric') else None
        self.npatch1 = params.get('npatch1', 1)
        self.npatch2 = params.get('npatch2', 1)
        self.npatch3 = params.get('npatch3', 1)

    def _calculate_varzeta(self, n):
        # corr3base.py:2483-2487
        size = int(np.prod(self.data_shape))
        self._varzeta = [np.zeros(self.data_shape) for _ in range(n)]
        if self._var_num != 0:
            for i in range(n):
                self._varzeta[i].ravel()[:] = self.cov_diag[i * size:(i + 1) * size].real

    # ---- LogSAS through the Fourier sums ----

    def _x_to_natural_phases(self):
        """Phases from the multipole's x projection (Porth et al.) to the centroid projection,
        per zeta (corr3base.py:2489-2551); None for a class without spin."""
        zs = self._zetas
        phases = [None] * len(zs)
        if not zs:
            return phases
        s = self.meand2
        t = self.meand3 * np.exp(1j * self.meanphi * self._phi_units)
        u = (1 + t / np.abs(t)) / 2
        q1 = (s + t) / 3.
        q2 = q1 - t
        q3 = q1 - s
        g1 = (u * np.conj(q1)) ** _spin(self._letters[0])
        g2 = (t * np.conj(q2)) ** _spin(self._letters[1])
        g3 = (s * np.conj(q3)) ** _spin(self._letters[2])
        g1 /= np.abs(g1)
        g2 /= np.abs(g2)
        g3 /= np.abs(g3)
        phases[0] = g1 * g2 * g3
        if len(zs) == 4:
            phases[1] = np.conj(g1) * g2 * g3
            phases[2] = g1 * np.conj(g2) * g3
            phases[3] = g1 * g2 * np.conj(g3)
        return phases

    def toSAS(self, *, target=None, **kwargs):
        """Convert the multipole sums to LogSAS binning (corr3base.py:2553-2677):
        N(d2, d3, phi) = 1/2pi sum_n N_n(d2, d3) exp(i n phi) at the phi bin centres, times
        the phi bin size; the zetas likewise, divided by the weight and rotated to the
        centroid projection.  The sums over n are one FP64 product on the GPU per array."""
        if self.bin_type != 'LogMultipole':
            raise TypeError("toSAS is invalid for bin_type = %s" % self.bin_type)
        if target is None:
            config = self.config.copy()
            config['bin_type'] = 'LogSAS'
            max_n = config.pop('max_n')
            if 'nphi_bins' not in kwargs and 'phi_bin_size' not in kwargs:
                config['nphi_bins'] = max_n
            config.update(kwargs)
            sas = self.__class__(config)
        else:
            if not isinstance(target, self.__class__):
                raise ValueError("target must be an instance of %s" % self.__class__)
            sas = target
            sas.clear()
        if not np.array_equal(sas.rnom1d, self.rnom1d):
            raise ValueError("toSAS cannot change sep parameters")
        sas.meand2[:, :, :] = self.meand2[:, :, 0][:, :, None]
        sas.meanlogd2[:, :, :] = self.meanlogd2[:, :, 0][:, :, None]
        sas.meand3[:, :, :] = self.meand3[:, :, 0][:, :, None]
        sas.meanlogd3[:, :, :] = self.meanlogd3[:, :, 0][:, :, None]
        sas.npatch1, sas.npatch2, sas.npatch3 = self.npatch1, self.npatch2, self.npatch3
        sas.coords = self.coords
        sas.metric = self.metric
        sas._var_num = self._var_num
        sas.meanu[:] = sas.phi / sas._phi_units
        sas.meand1[:] = np.sqrt(sas.meand2 ** 2 + sas.meand3 ** 2
                                - 2 * sas.meand2 * sas.meand3 * np.cos(sas.phi))
        sas.meanlogd1[:] = np.log(sas.meand1)
        # Eqn 26 of Porth et al. 2023: the weight's real part, Re(W E) = [Wr, Wi] [Er; -Ei]
        nb, nn, nphi = self.nbins, 2 * self.max_n + 1, sas.nphi_bins
        arg = self.n1d[:, None] * sas.phi1d[None, :]
        er, ei = np.cos(arg), np.sin(arg)
        scale = sas.phi_bin_size / (2 * np.pi)
        W = np.concatenate([self.weightr.reshape(nb * nb, nn), self.weighti.reshape(nb * nb, nn)],
                           axis=1)
        sas.weightr[:] = (_gemm(W, np.concatenate([er, -ei], axis=0)) * scale).reshape(
            nb, nb, nphi)
        phi_frac = (sas.max_phi - sas.min_phi) / (2 * np.pi)
        denom = np.sum(sas.weight, axis=2)
        denom[denom == 0] = 1
        ratio = self.ntri[:, :, 0] / denom * phi_frac
        sas.ntri[:] = sas.weight * ratio[:, :, None]
        # per-patch results: raw sums (the mean distances times the result's weight), the
        # weights of all results in one product
        keys = list(self.results)
        sas.results = {}
        if keys:
            Wk = np.concatenate([np.concatenate([self.results[k].weightr.reshape(nb * nb, nn),
                                                 self.results[k].weighti.reshape(nb * nb, nn)],
                                                axis=1) for k in keys], axis=0)
            Rw = _gemm(Wk, np.concatenate([er, -ei], axis=0)) * scale
            for j, k in enumerate(keys):
                temp = sas.copy()
                temp.results = {}
                temp.weightr[:] = Rw[j * nb * nb:(j + 1) * nb * nb].reshape(nb, nb, nphi)
                temp.ntri[:] = temp.weight * ratio[:, :, None]
                for nm in ('meand1', 'meand2', 'meand3', 'meanlogd1', 'meanlogd2', 'meanlogd3',
                           'meanu'):
                    getattr(temp, nm)[:] *= temp.weightr
                sas.results[k] = temp
        zs = self._zetas
        if not zs:
            return self._to_sas_extra(sas)
        # Z E for the four complex zetas at once: [Zr, Zi] [[Er, Ei], [-Ei, Er]]
        Z = np.concatenate([np.concatenate([z.real.reshape(nb * nb, nn),
                                            z.imag.reshape(nb * nb, nn)], axis=1) for z in zs],
                           axis=0)
        E = np.block([[er, ei], [-ei, er]])
        R = _gemm(Z, E) * scale
        mask = sas.weightr != 0
        phases = sas._x_to_natural_phases()
        for i in range(len(zs)):
            blk = R[i * nb * nb:(i + 1) * nb * nb]
            z = (blk[:, :nphi] + 1j * blk[:, nphi:]).reshape(nb, nb, nphi)
            z[mask] /= sas.weightr[mask]
            if phases[i] is not None:
                z = z * phases[i]
            sas._z[2 * i][:] = np.real(z)
            if sas._z[2 * i + 1] is not None:
                sas._z[2 * i + 1][:] = np.imag(z)
        if keys:
            # the results keep raw sums: rotated, not divided by the weight
            Zk = np.concatenate([np.concatenate([z.real.reshape(nb * nb, nn),
                                                 z.imag.reshape(nb * nb, nn)], axis=1)
                                 for k in keys for z in self.results[k]._zetas], axis=0)
            Rk = _gemm(Zk, E) * scale
            nz = len(zs)
            for j, k in enumerate(keys):
                temp = sas.results[k]
                for i in range(nz):
                    blk = Rk[(j * nz + i) * nb * nb:(j * nz + i + 1) * nb * nb]
                    z = (blk[:, :nphi] + 1j * blk[:, nphi:]).reshape(nb, nb, nphi)
                    if phases[i] is not None:
                        z = z * phases[i]
                    temp._z[2 * i][:] = np.real(z)
                    if temp._z[2 * i + 1] is not None:
                        temp._z[2 * i + 1][:] = np.imag(z)
        return self._to_sas_extra(sas)

    def _to_sas_extra(self, sas):
        return sas


class NNNCorrelation(Corr3):
    """Count-count-count correlation (treecorr/nnncorrelation.py)."""
    _d = 0
    _letters = 'NNN'

    def __init__(self, config=None, *, logger=None, **kwargs):
        super().__init__(config, logger=logger, **kwargs)
        self.tot = 0.
        self.zeta = None
        self.varzeta = None
        self._rrr_weight = None
        self._rrr = self._drr = self._rdd = None

    def copy(self):
        # nnncorrelation.py:88-96: the randoms are copied with it
        ret = super().copy()
        for nm in ('_rrr', '_drr', '_rdd'):
            o = getattr(self, nm)
            if o is not None and o is not True:
                setattr(ret, nm, o.copy())
        return ret

    def _zero_copy(self, tot):
        # nnncorrelation.py:98-128: an empty result that only carries tot
        ret = Corr3.copy(self)
        ret._clear()
        ret.results = {}
        ret.tot = tot
        ret._rrr = ret._drr = ret._rdd = None
        return ret

    def _mean_weight(self):
        mean_ntri = np.mean(self.ntri)
        return 1 if mean_ntri == 0 else np.mean(self.weight) / mean_ntri

    def calculateZeta(self, *, rrr, drr=None, rdd=None):
        """zeta = (DDD - RRR) / RRR, or (DDD - RDD + DRR - RRR) / RRR with the data-random
        cross correlations, each scaled by the ratio of the totals of possible triangles
        (nnncorrelation.py:228-396).  Returns (zeta, varzeta)."""
        if rrr.tot == 0:
            raise ValueError("rrr has tot=0.")
        if (rdd is not None) != (drr is not None):
            raise TypeError("Must provide both rdd and drr (or neither).")
        if self.bin_type == 'LogMultipole':
            raise TypeError("calculateZeta is not valid for LogMultipole binning.")
        rrrf = self.tot / rrr.tot
        if drr is not None:
            if drr.tot == 0:
                raise ValueError("drr has tot=0.")
            drrf = self.tot / drr.tot
            if rdd.tot == 0:
                raise ValueError("rdd has tot=0.")
            rddf = self.tot / rdd.tot
        denom = rrr.weight * rrrf
        if rdd is None:
            zeta = self.weight - denom
        else:
            zeta = self.weight - rdd.weight * rddf + drr.weight * drrf - denom
        denom[rrr.weight == 0] = 1.
        self.zeta = zeta / denom
        dddw = self._mean_weight()
        factor = 1 + rrrf * rrr._mean_weight() / dddw
        if rdd is not None:
            factor += drrf * drr._mean_weight() / dddw + rddf * rdd._mean_weight() / dddw
        self._var_num = dddw * factor ** 2
        self._rrr_weight = np.abs(rrr.weight) * rrrf
        self._rrr, self._drr, self._rdd = rrr, drr, rdd
        if len(self.results) > 0:
            if rrr.npatch1 != 1 and rrr.npatch1 != self.npatch1:
                raise RuntimeError("If using patches, RRR must be run with the same patches "
                                   "as DDD")
            if drr is not None and (len(drr.results) == 0 or drr.npatch1 != self.npatch1 or
                                    drr.npatch2 not in (self.npatch2, 1)):
                raise RuntimeError("DRR must be run with the same patches as DDD")
            if rdd is not None and (len(rdd.results) == 0 or rdd.npatch2 != self.npatch2 or
                                    rdd.npatch1 not in (self.npatch1, 1)):
                raise RuntimeError("RDD must be run with the same patches as DDD")
            # triples the randoms have and the data lack enter as empty results
            add = set()
            if rrr.npatch1 != 1:
                add |= set(rrr.results) - set(self.results)
            if drr is not None and drr.npatch2 != 1:
                add |= set(drr.results) - set(self.results)
            if rdd is not None and rdd.npatch1 != 1:
                add |= set(rdd.results) - set(self.results)
            for ijk in add:
                self.results[ijk] = self._zero_copy(0)
        self._cov = self.estimate_cov(self.var_method)
        self.varzeta = self.cov_diag.reshape(self.zeta.shape)
        return self.zeta, self.varzeta

    def _calculate_xi_from_pairs(self, pairs, corr_only):
        # nnncorrelation.py:398-460: DDD from the triples, the randoms from the same
        # triples when they have patches, else scaled by the area fraction the triples
        # cover
        super()._calculate_xi_from_pairs(pairs, corr_only)
        if self._rrr is None:
            return
        ddd = self.weight
        if len(self._rrr.results) > 0:
            self._rrr._calculate_xi_from_pairs(self._rrr._keep_ok(pairs), corr_only=True)
            ddd_tot = self.tot
        else:
            area_frac = np.sum([(self.results[(i, j, k)].tot * w) ** (1. / 3.)
                                for i, j, k, w in pairs if i == j == k])
            area_frac /= np.sum([c.tot ** (1. / 3.) for ijk, c in self.results.items()
                                 if ijk[0] == ijk[1] == ijk[2]])
            ddd_tot = np.sum([c.tot for c in self.results.values()]) * area_frac
        rrrf = ddd_tot / self._rrr.tot
        denom = self._rrr.weight * rrrf
        if self._drr is None:
            zeta = ddd - denom
        else:
            pairs2 = pairs
            if self._drr.npatch2 == 1:
                pairs2 = [(i, 0, 0, w) for i, j, k, w in pairs2 if i == j == k]
            self._drr._calculate_xi_from_pairs(self._drr._keep_ok(pairs2), corr_only=True)
            pairs3 = pairs
            if self._rdd.npatch1 == 1 and not all(p[0] == 0 for p in pairs):
                pairs3 = [(0, j, k, w) for i, j, k, w in pairs3 if i == j or i == k]
            self._rdd._calculate_xi_from_pairs(self._rdd._keep_ok(pairs3), corr_only=True)
            zeta = (ddd - self._rdd.weight * (ddd_tot / self._rdd.tot) +
                    self._drr.weight * (ddd_tot / self._drr.tot) - denom)
        denom[denom == 0] = 1
        self.zeta = zeta / denom
        self._rrr_weight = denom

    def _clear(self):
        super()._clear()
        self.tot = 0.

    def process_auto(self, cat, **kw):
        super().process_auto(cat, **kw)
        self.tot += (1. / 6.) * cat.sumw ** 3

    def process_cross12(self, cat1, cat2, **kw):
        super().process_cross12(cat1, cat2, **kw)
        self.tot += 0.5 * cat1.sumw * cat2.sumw ** 2

    def process_cross(self, cat1, cat2, cat3, **kw):
        super().process_cross(cat1, cat2, cat3, **kw)
        self.tot += cat1.sumw * cat2.sumw * cat3.sumw

    def finalize(self):
        self._finalize()

    def _to_sas_extra(self, sas):
        sas.tot = self.tot
        for k, v in self.results.items():
            sas.results[k].tot = v.tot
        return sas

    def __iadd__(self, other):
        # nnncorrelation.py:184-193
        super().__iadd__(other)
        self.tot += other.tot
        return self

    def _sum(self, others, corr_only):
        # nnncorrelation.py:155-168: tot from every triple, sums from the nonzero ones
        tot = np.sum([c.tot * w for c, w in others])
        others = [(c, w) for c, w in others if c.nonzero]
        if len(others) == 0:
            self._clear()
        else:
            super()._sum(others, corr_only)
        self.tot = tot

    # output files (nnncorrelation.py:459-770): DDD, and with the randoms zeta, RRR, DRR,
    # RDD scaled to DDD's total; with write_patch_results the randoms' own groups follow
    _write_rrr = _write_drr = _write_rdd = None
    _write_patch_results = False

    def write(self, file_name, *, rrr=None, drr=None, rdd=None, file_type=None,
              precision=None, write_patch_results=False, write_cov=False):
        """Write the correlation to a file (nnncorrelation.py:459-597).  Given rrr (and
        drr, rdd), zeta is computed by calculateZeta and the randoms' columns are added."""
        precision = self.config.get('precision', 4) if precision is None else precision
        self._write_rrr, self._write_drr, self._write_rdd = rrr, drr, rdd
        self._write_patch_results = write_patch_results
        try:
            with make_writer(file_name, precision, file_type) as writer:
                self._write(writer, None, write_patch_results, write_cov=write_cov,
                            zero_tot=True)
                if write_patch_results:
                    for nm, o in (('_rrr', rrr or self._rrr), ('_drr', drr or self._drr),
                                  ('_rdd', rdd or self._rdd)):
                        if o is not None:
                            o._write(writer, nm, True, zero_tot=True)
        finally:
            self._write_rrr = self._write_drr = self._write_rdd = None
            self._write_patch_results = False

    def _write(self, writer, name, write_patch_results, write_cov=False, zero_tot=True):
        super()._write(writer, name, write_patch_results, write_cov=write_cov, zero_tot=True)

    @property
    def _write_params(self):
        # nnncorrelation.py:673-681
        params = super()._write_params
        params['tot'] = float(self.tot)
        if self._write_patch_results:
            params['_rrr'] = self._rrr is not None
            params['_drr'] = self._drr is not None
            params['_rdd'] = self._rdd is not None
        return params

    @property
    def _write_class_col_names(self):
        return ['zeta', 'sigma_zeta'] if (self.zeta is not None or self._write_rrr) else []

    @property
    def _write_class_data(self):
        if self._write_rrr is not None:
            zeta, varzeta = self.calculateZeta(rrr=self._write_rrr, drr=self._write_drr,
                                               rdd=self._write_rdd)
            return [zeta, np.sqrt(varzeta)]
        if self._write_drr is not None or self._write_rdd is not None:
            raise TypeError("rrr must be provided if other combinations are not None")
        return [self.zeta, np.sqrt(self.varzeta)] if self.zeta is not None else []

    def _write_weight_names(self):
        if self._write_rrr is not None:
            return ['DDD', 'RRR'] + (['DRR', 'RDD'] if self._write_drr is not None else []) + \
                ['ntri']
        multi = ['weightr', 'weighti'] if self.weighti.size else []
        return multi + ['DDD', 'ntri']

    def _write_weight_data(self):
        rrr, drr, rdd = self._write_rrr, self._write_drr, self._write_rdd
        if rrr is not None:
            out = [self.weight, rrr.weight * (self.tot / rrr.tot)]
            if drr is not None:
                out += [drr.weight * (self.tot / drr.tot), rdd.weight * (self.tot / rdd.tot)]
            return out + [self.ntri]
        if self.weighti.size:
            return [self.weightr, self.weighti, np.abs(self.weight), self.ntri]
        return [self.weightr, self.ntri]

    def _do_read(self, reader, name=None, params=None):
        # nnncorrelation.py:722-736: the randoms' groups follow when they were written
        self._read(reader, name, params)
        for nm in ('_rrr', '_drr', '_rdd'):
            if getattr(self, nm):
                o = Corr3.copy(self)
                o._rrr = o._drr = o._rdd = None
                o._read(reader, name=nm)
                setattr(self, nm, o)

    def _read_from_data(self, data, params):
        # nnncorrelation.py:754-770
        super()._read_from_data(data, params)
        s = self.data_shape
        if not self.weighti.size:
            self.weightr[...] = data['DDD'].reshape(s)
        if 'zeta' in data.dtype.names:
            self.zeta = data['zeta'].reshape(s)
            self.varzeta = data['sigma_zeta'].reshape(s) ** 2
        self.tot = params['tot']
        self._rrr = params.get('_rrr', None) or None
        self._drr = params.get('_drr', None) or None
        self._rdd = params.get('_rdd', None) or None

    def getStat(self):
        # nnncorrelation.py:200-207
        if self.zeta is None:
            raise RuntimeError("You need to call calculateZeta before calling estimate_cov.")
        return self.zeta.ravel()

    def getWeight(self):
        # nnncorrelation.py:209-218: the scaled RRR of calculateZeta
        if self._rrr_weight is not None:
            return self._rrr_weight.ravel()
        return self.tot


class GGGCorrelation(Corr3):
    """Shear-shear-shear correlation (treecorr/gggcorrelation.py)."""
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

    def _finalize_with(self, cat1, cat2, cat3):
        # corr3base.py:2252-2292: one variance per vertex
        v1 = calculateVarG(cat1)
        v2 = v1 if cat2 is None else calculateVarG(cat2)
        v3 = v2 if cat3 is None else calculateVarG(cat3)
        self.finalize(v1, v2, v3)

    def finalize(self, varg1, varg2, varg3):
        # gggcorrelation.py:162-178
        self._finalize()
        self._var_num = 4 * varg1 * varg2 * varg3
        if self.bin_type in ('LogSAS', 'LogMultipole'):
            self._var_num *= 2
        self._cov = None
        self._varzeta = None

    vargam0 = property(lambda self: self._vargam(0))
    vargam1 = property(lambda self: self._vargam(1))
    vargam2 = property(lambda self: self._vargam(2))
    vargam3 = property(lambda self: self._vargam(3))

    def _vargam(self, i):
        # gggcorrelation.py:180-201
        if self._varzeta is None:
            self._calculate_varzeta(4)
        return self._varzeta[i]

    # gggcorrelation.py:247-275
    _write_class_col_names = ('gam0r', 'gam0i', 'gam1r', 'gam1i', 'gam2r', 'gam2i', 'gam3r',
                              'gam3i', 'sigma_gam0', 'sigma_gam1', 'sigma_gam2', 'sigma_gam3')

    @property
    def _write_class_data(self):
        return list(self._z) + [np.sqrt(self.vargam0), np.sqrt(self.vargam1),
                                np.sqrt(self.vargam2), np.sqrt(self.vargam3)]

    def _read_from_data(self, data, params):
        super()._read_from_data(data, params)
        s = self.data_shape
        for i, nm in enumerate(('gam0r', 'gam0i', 'gam1r', 'gam1i', 'gam2r', 'gam2i', 'gam3r',
                                'gam3i')):
            self._z[i][...] = data[nm].reshape(s)
        self._varzeta = [data['sigma_gam%d' % i].reshape(s) ** 2 for i in range(4)]

    def getStat(self):
        """gam0, gam1, gam2, gam3 raveled and concatenated (gggcorrelation.py:203-215);
        complex."""
        return np.concatenate([self.gam0.ravel(), self.gam1.ravel(), self.gam2.ravel(),
                               self.gam3.ravel()])

    def getWeight(self):
        # gggcorrelation.py:217-222
        return np.concatenate([np.abs(self.weight.ravel())] * 4)
