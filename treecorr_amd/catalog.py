"""Catalog: the input arrays of a correlation (treecorr/catalog.py, restated).

Mirrors the parts of treecorr.Catalog that feed the ball-tree engine:
positions (x,y[,z] or ra,dec[,r] with units), weights w/wpos, scalar k, shear
g1/g2, flips, zero-weight removal (catalog.py:1343-1466), ra/dec -> xyz with the
reference's GenerateXYZ (catalog.py:1609-1622), the derived sums/variances used
at finalize (catalog.py:1200-1330), patches given as integer labels, and a
one-entry cache of built fields per kind (catalog.py:2230-2316).
"""
import ctypes

import numpy as np

from . import _lib
from .config import parse_unit


# The float64 columns a catalog may hold.
_COLUMNS = ('_x', '_y', '_z', '_ra', '_dec', '_r', '_w', '_wpos', '_k', '_g1', '_g2', '_z1',
            '_z2', '_v1', '_v2', '_t1', '_t2', '_q1', '_q2')
# the complex fields' columns: spin 0, 1, 2, 3, 4 (treecorr/catalog.py)
_COMPLEX = {'Z': ('z1', 'z2'), 'V': ('v1', 'v2'), 'G': ('g1', 'g2'), 'T': ('t1', 't2'),
            'Q': ('q1', 'q2')}


class Catalog:
    def __init__(self, file_name=None, config=None, *, x=None, y=None, z=None, ra=None,
                 dec=None, r=None, w=None, wpos=None, k=None, g1=None, g2=None, z1=None,
                 z2=None, v1=None, v2=None, t1=None, t2=None, q1=None, q2=None, patch=None,
                 x_units=None, y_units=None, ra_units=None, dec_units=None, flip_g1=False,
                 flip_g2=False, keep_zero_weight=False, name='', patch_centers=None, **kwargs):
        if file_name is not None:
            raise NotImplementedError("treecorr_amd.Catalog takes arrays; file input is not part "
                                      "of the GPU engine")
        cfg = dict(config or {})
        cfg.update(kwargs)
        self.name = name
        self.config = cfg

        def arr(a):
            return None if a is None else np.array(a, dtype=float, copy=True).ravel()
        self._x, self._y, self._z = arr(x), arr(y), arr(z)
        self._ra, self._dec, self._r = arr(ra), arr(dec), arr(r)
        self._w, self._wpos = arr(w), arr(wpos)
        self._k, self._g1, self._g2 = arr(k), arr(g1), arr(g2)
        self._z1, self._z2 = arr(z1), arr(z2)
        self._v1, self._v2 = arr(v1), arr(v2)
        self._t1, self._t2 = arr(t1), arr(t2)
        self._q1, self._q2 = arr(q1), arr(q2)
        # catalog.py:862-875: an integer patch makes this the catalog of that one patch
        single = None
        if patch is not None and np.ndim(patch) == 0:
            single, patch = int(patch), None
        self._patch = None if patch is None else np.asarray(patch, dtype=int).ravel()
        if patch_centers is not None and patch is not None:
            raise ValueError("Cannot provide both patch and patch_centers")
        self._centers = None if patch_centers is None else \
            np.ascontiguousarray(patch_centers, dtype=float)

        if self._x is not None or self._y is not None:
            if self._x is None or self._y is None:
                raise TypeError("x and y must both be given")
            if self._ra is not None or self._dec is not None:
                raise TypeError("ra and dec may not be given with x,y")
            if x_units is not None or y_units is not None:
                if x_units is None or y_units is None:
                    raise TypeError("x_units and y_units must both be given")
                self._x *= parse_unit(x_units)
                self._y *= parse_unit(y_units)
        elif self._ra is not None or self._dec is not None:
            if self._ra is None or self._dec is None:
                raise TypeError("ra and dec must both be given")
            if ra_units is None or dec_units is None:
                raise TypeError("ra_units and dec_units are required with ra, dec")
            if self._z is not None:
                raise TypeError("z may not be given with ra, dec")
            self._ra = self._ra * parse_unit(ra_units)
            self._dec = self._dec * parse_unit(dec_units)
        else:
            raise TypeError("Either x,y or ra,dec must be given")

        n = len(self._x) if self._x is not None else len(self._ra)
        for nm in _COLUMNS[1:] + ('_patch',):
            a = getattr(self, nm)
            if a is not None and len(a) != n:
                raise ValueError("%s has the wrong length" % nm[1:])
        for nm in _COLUMNS:
            a = getattr(self, nm)
            if a is not None and np.any(np.isnan(a)):
                raise ValueError("NaN values found in %s" % nm[1:])
        if flip_g1 and self._g1 is not None:
            self._g1 = -self._g1
        if flip_g2 and self._g2 is not None:
            self._g2 = -self._g2

        self._generate_xyz()

        # catalog.py:1432-1464: wpos==0 forces w=0; drop zero-weight objects
        if self._wpos is not None and np.any(self._wpos == 0.):
            if self._w is None:
                self._w = np.ones(n)
            self._w[self._wpos == 0.] = 0.
        if self._w is not None:
            self._nontrivial_w = True
            self._sumw = np.sum(self._w)
            if self._sumw == 0:
                raise ValueError("Catalog has invalid sumw == 0")
        else:
            self._nontrivial_w = False
            self._sumw = n
            self._w = np.ones(n)
        if self._nontrivial_w and not keep_zero_weight:
            wpos = self._wpos if self._wpos is not None else self._w
            if np.any(wpos == 0):
                self._select(np.where(wpos != 0)[0])

        self._fields = {}
        npatch = cfg.get('npatch')
        if single is not None:
            self._select_single_patch(single)
        elif self._centers is not None:
            self._assign_patches()
        elif self._patch is None and npatch is not None and int(npatch) != 1:
            # catalog.py:1473-1484: patches from k-means on an N field whose top level
            # has about npatch cells
            npatch = int(npatch)
            if npatch > self.ntot:
                raise ValueError("npatch is larger than the number of objects in catalog")
            max_top = int.bit_length(npatch) - 1
            c = 'spherical' if self._ra is not None else self.coords
            fld = self.getNField(max_top=max_top, coords=c)
            self._patch, self._centers = fld.run_kmeans(
                npatch, init=cfg.get('kmeans_init', 'tree'), alt=bool(cfg.get('kmeans_alt', False)),
                rng=cfg.get('rng'))
            self._npatch = npatch
            self._fields = {}
        else:
            self._npatch = 1 if self._patch is None else int(self._patch.max()) + 1

    def _select_single_patch(self, p):
        # catalog.py:1576-1594, 1624-1630: keep the objects of patch p -- those for which
        # no other centre is strictly nearer (SelectPatch, on the GPU) when centres are
        # given, else all of them
        if self._centers is not None:
            c = self._centers
            dim = 2 if self.coords == 'flat' else 3
            if c.ndim != 2 or c.shape[1] != dim:
                raise ValueError("Centers array has wrong shape.")
            self._generate_xyz()
            use = np.empty(self.ntot, dtype=np.int32)
            z = None if dim == 2 else np.ascontiguousarray(self._z, dtype=float)
            _lib.check(_lib.lib().tc_select_patch(
                p, _lib.dptr(c.ravel()), len(c),
                _lib.dptr(np.ascontiguousarray(self._x, dtype=float)),
                _lib.dptr(np.ascontiguousarray(self._y, dtype=float)), _lib.dptr(z), self.ntot,
                use.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))
            self._select(np.where(use)[0])
            self._npatch = len(c)
        elif self.config.get('npatch') is not None:       # catalog.py:883-886
            self._npatch = int(self.config['npatch'])
            if self._npatch < 1:
                raise ValueError("npatch must be >= 1")
        elif p:
            raise ValueError("Either npatch or patch_centers is required when providing "
                             "an integer patch value.")
        else:
            self._npatch = 1
        self._patch = None
        self._single_patch = p

    def _assign_patches(self):
        # catalog.py:1466-1472, 1556-1566: the nearest of the given centres (QuickAssign
        # on the GPU); npatch is the number of centres, empty patches included
        c = self._centers
        dim = 2 if self.coords == 'flat' else 3
        if c.ndim != 2 or c.shape[1] != dim:
            raise ValueError("Centers array has wrong shape.")
        self._patch = np.empty(self.ntot, dtype=np.int32)
        z = None if dim == 2 else np.ascontiguousarray(self._z, dtype=float)
        ip = ctypes.POINTER(ctypes.c_int)
        _lib.check(_lib.lib().tc_quick_assign(
            _lib.dptr(c.ravel()), len(c), _lib.dptr(np.ascontiguousarray(self._x, dtype=float)),
            _lib.dptr(np.ascontiguousarray(self._y, dtype=float)), _lib.dptr(z), self.ntot,
            self._patch.ctypes.data_as(ip)))
        self._patch = self._patch.astype(int)
        self._npatch = len(c)

    @property
    def patch_centers(self):
        return self._centers

    # ---- construction helpers ----
    def _generate_xyz(self):
        # catalog.py:1609-1622 -> GenerateXYZ (src/KMeans.cpp:1013-1050, glibc sincos)
        if self._x is None:
            n = len(self._ra)
            x, y, z = np.empty(n), np.empty(n), np.empty(n)
            r = self._r
            _lib.check(_lib.lib().tc_generate_xyz(_lib.dptr(x), _lib.dptr(y), _lib.dptr(z),
                                                  _lib.dptr(np.ascontiguousarray(self._ra)),
                                                  _lib.dptr(np.ascontiguousarray(self._dec)),
                                                  _lib.dptr(r), n))
            self._x, self._y, self._z = x, y, z

    def _select(self, idx):
        for nm in _COLUMNS + ('_patch',):
            a = getattr(self, nm)
            if a is not None:
                setattr(self, nm, a[idx])

    # ---- arrays ----
    x = property(lambda self: self._x)
    y = property(lambda self: self._y)
    z = property(lambda self: self._z)
    ra = property(lambda self: self._ra)
    dec = property(lambda self: self._dec)
    r = property(lambda self: self._r)
    w = property(lambda self: self._w)
    wpos = property(lambda self: self._wpos)
    k = property(lambda self: self._k)
    g1 = property(lambda self: self._g1)
    g2 = property(lambda self: self._g2)
    z1 = property(lambda self: self._z1)
    z2 = property(lambda self: self._z2)
    v1 = property(lambda self: self._v1)
    v2 = property(lambda self: self._v2)
    t1 = property(lambda self: self._t1)
    t2 = property(lambda self: self._t2)
    q1 = property(lambda self: self._q1)
    q2 = property(lambda self: self._q2)
    # catalog.py:1199-1204: a single-patch catalog reports its patch number
    patch = property(lambda self: self._single_patch if self._single_patch is not None
                     else self._patch)
    npatch = property(lambda self: self._npatch)
    nontrivial_w = property(lambda self: self._nontrivial_w)

    @property
    def coords(self):
        # catalog.py:1318-1328
        if self._ra is not None:
            return 'spherical' if self._r is None else '3d'
        return 'flat' if self._z is None else '3d'

    @property
    def ntot(self):
        return len(self._x)

    @property
    def nobj(self):
        return int(np.sum(self._w != 0)) if self._nontrivial_w else self.ntot

    @property
    def sumw(self):
        c = self.__dict__.setdefault('_cache', {})
        if 'sumw' not in c:
            c['sumw'] = float(np.sum(self._w))
        return c['sumw']

    @property
    def sumw2(self):
        return float(np.sum(self._w ** 2)) if self._nontrivial_w else float(self.ntot)

    @property
    def meanw(self):
        return self.sumw / self.nobj if self._nontrivial_w else 1.

    def _weighted_var(self, v):
        # catalog.py:1213-1229
        if v is None:
            return 0.
        if self._nontrivial_w:
            use = self._w != 0
            meanv = np.sum(self._w[use] * v[use]) / self.sumw
            return np.sum(self._w[use] ** 2 * (v[use] - meanv) ** 2) / self.sumw
        return np.sum((v - np.mean(v)) ** 2) / self.nobj

    @property
    def vark(self):
        return self._weighted_var(self._k)

    @property
    def varg(self):
        return (self._weighted_var(self._g1) + self._weighted_var(self._g2)) / 2.

    def _var_complex(self, letter):
        # catalog.py:1229-1275: the mean of the two components' variances
        a, b = _COMPLEX[letter]
        return (self._weighted_var(getattr(self, '_' + a)) +
                self._weighted_var(getattr(self, '_' + b))) / 2.

    varz = property(lambda self: self._var_complex('Z'))
    varv = property(lambda self: self._var_complex('V'))
    vart = property(lambda self: self._var_complex('T'))
    varq = property(lambda self: self._var_complex('Q'))

    def _get_center_size(self):
        # catalog.py:1330-1341 (computed once per catalog)
        if getattr(self, '_cen_s', None) is None:
            mx, my = np.mean(self._x), np.mean(self._y)
            mz = 0
            dsq = (self._x - mx) ** 2 + (self._y - my) ** 2
            if self._z is not None:
                mz = np.mean(self._z)
                dsq += (self._z - mz) ** 2
            self._cen_s = (mx, my, mz, np.max(dsq) ** 0.5)
        return self._cen_s

    # ---- patches ----
    def get_patches(self):
        """A list with one Catalog per non-empty patch (or [self] without patches),
        made once so that their fields are built once (catalog.py:2966-3050)."""
        if self._patch is None:
            return [self]
        if getattr(self, '_patches', None) is not None:
            return self._patches
        # One stable counting sort of every column by patch (native, host threads);
        # each patch catalog then holds views of its contiguous slice.
        names = [nm for nm in _COLUMNS if getattr(self, nm) is not None]
        cols = [np.ascontiguousarray(getattr(self, nm), dtype=float) for nm in names]
        split = [np.empty_like(c) for c in cols]
        counts = np.zeros(self._npatch, dtype=np.int64)
        labels = np.ascontiguousarray(self._patch, dtype=np.int32)
        dp = ctypes.POINTER(ctypes.c_double)
        _lib.check(_lib.lib().tc_patch_split(
            labels.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), len(labels), self._npatch,
            len(cols), (dp * len(cols))(*[c.ctypes.data_as(dp) for c in cols]),
            (dp * len(cols))(*[c.ctypes.data_as(dp) for c in split]),
            counts.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong))))
        bounds = np.concatenate([[0], np.cumsum(counts)])
        # kept for building all patch fields in one pass (field.build_patch_fields)
        self._split_cols = dict(zip(names, split))
        self._split_bounds = np.concatenate([[0], np.cumsum(counts[counts > 0])])
        self._split_patches = [p for p in range(self._npatch) if counts[p] > 0]
        out = []
        for p in range(self._npatch):
            if counts[p] == 0:
                continue
            sub = Catalog.__new__(Catalog)
            sub.__dict__.update(self.__dict__)
            sub._fields = {}
            sub._cen_s = None
            sub._cache = {}
            sub._patches = None
            sub._split_parent = self
            for nm in ('_split_cols', '_split_bounds', '_split_patches'):
                sub.__dict__.pop(nm, None)
            for nm, a in zip(names, split):
                setattr(sub, nm, a[bounds[p]:bounds[p + 1]])
            sub._patch = None
            sub._npatch = self._npatch    # a patch catalog knows the full npatch
            sub._single_patch = p
            sub.name = '%s patch %d' % (self.name, p)
            out.append(sub)
        self._patches = out
        return out

    _single_patch = None
    _split_parent = None     # the catalog whose get_patches made this patch catalog

    # ---- fields ----
    def _get_field(self, kind, cls, **kw):
        key = (kind, tuple(sorted(kw.items())))
        f = self._fields.get(kind)
        if f is not None and f[0] == key:
            return f[1]
        from . import field as _field
        fld = getattr(_field, cls)(self, **kw)
        self._fields[kind] = (key, fld)   # one cached field per kind
        return fld

    def getNField(self, *, min_size=0, max_size=None, split_method='mean', brute=False,
                  min_top=None, max_top=10, coords=None):
        return self._get_field('N', 'NField', min_size=min_size, max_size=max_size,
                               split_method=split_method, brute=brute, min_top=min_top,
                               max_top=max_top, coords=coords)

    def getKField(self, *, min_size=0, max_size=None, split_method='mean', brute=False,
                  min_top=None, max_top=10, coords=None):
        if self._k is None:
            raise TypeError("k is not defined for this catalog")
        return self._get_field('K', 'KField', min_size=min_size, max_size=max_size,
                               split_method=split_method, brute=brute, min_top=min_top,
                               max_top=max_top, coords=coords)

    def _get_complex_field(self, letter, **kw):
        a, b = _COMPLEX[letter]
        if getattr(self, '_' + a) is None or getattr(self, '_' + b) is None:
            raise TypeError("%s,%s are not defined for this catalog" % (a, b))
        return self._get_field(letter, letter + 'Field', **kw)

    def getZField(self, *, min_size=0, max_size=None, split_method='mean', brute=False,
                  min_top=None, max_top=10, coords=None):
        return self._get_complex_field('Z', min_size=min_size, max_size=max_size,
                                       split_method=split_method, brute=brute,
                                       min_top=min_top, max_top=max_top, coords=coords)

    def getVField(self, *, min_size=0, max_size=None, split_method='mean', brute=False,
                  min_top=None, max_top=10, coords=None):
        return self._get_complex_field('V', min_size=min_size, max_size=max_size,
                                       split_method=split_method, brute=brute,
                                       min_top=min_top, max_top=max_top, coords=coords)

    def getTField(self, *, min_size=0, max_size=None, split_method='mean', brute=False,
                  min_top=None, max_top=10, coords=None):
        return self._get_complex_field('T', min_size=min_size, max_size=max_size,
                                       split_method=split_method, brute=brute,
                                       min_top=min_top, max_top=max_top, coords=coords)

    def getQField(self, *, min_size=0, max_size=None, split_method='mean', brute=False,
                  min_top=None, max_top=10, coords=None):
        return self._get_complex_field('Q', min_size=min_size, max_size=max_size,
                                       split_method=split_method, brute=brute,
                                       min_top=min_top, max_top=max_top, coords=coords)

    def getGField(self, *, min_size=0, max_size=None, split_method='mean', brute=False,
                  min_top=None, max_top=10, coords=None):
        if self._g1 is None or self._g2 is None:
            raise TypeError("g1,g2 are not defined for this catalog")
        return self._get_field('G', 'GField', min_size=min_size, max_size=max_size,
                               split_method=split_method, brute=brute, min_top=min_top,
                               max_top=max_top, coords=coords)


def calculateVarK(cat_list):
    cats = cat_list if isinstance(cat_list, list) else [cat_list]
    if len(cats) == 1:
        return cats[0].vark
    k = np.concatenate([c.k for c in cats])
    w = np.concatenate([c.w for c in cats])
    use = w != 0
    meank = np.sum(w[use] * k[use]) / np.sum(w)
    return np.sum(w[use] ** 2 * (k[use] - meank) ** 2) / np.sum(w)


def _calculate_var_complex(cat_list, letter):
    """The variance per component of a complex field over catalogs
    (catalog.py:3402-3470)."""
    cats = cat_list if isinstance(cat_list, list) else [cat_list]
    if len(cats) == 1:
        return cats[0]._var_complex(letter)
    w = np.concatenate([c.w for c in cats])
    use = w != 0
    out = 0.
    for comp in _COMPLEX[letter]:
        v = np.concatenate([getattr(c, comp) for c in cats])
        meanv = np.sum(w[use] * v[use]) / np.sum(w)
        out += np.sum(w[use] ** 2 * (v[use] - meanv) ** 2) / np.sum(w)
    return out / 2.


def calculateVarZ(cat_list):
    return _calculate_var_complex(cat_list, 'Z')


def calculateVarV(cat_list):
    return _calculate_var_complex(cat_list, 'V')


def calculateVarT(cat_list):
    return _calculate_var_complex(cat_list, 'T')


def calculateVarQ(cat_list):
    return _calculate_var_complex(cat_list, 'Q')


def calculateVarG(cat_list):
    cats = cat_list if isinstance(cat_list, list) else [cat_list]
    if len(cats) == 1:
        return cats[0].varg
    w = np.concatenate([c.w for c in cats])
    use = w != 0
    out = 0.
    for comp in ('g1', 'g2'):
        g = np.concatenate([getattr(c, comp) for c in cats])
        meang = np.sum(w[use] * g[use]) / np.sum(w)
        out += np.sum(w[use] ** 2 * (g[use] - meang) ** 2) / np.sum(w)
    return out / 2.
