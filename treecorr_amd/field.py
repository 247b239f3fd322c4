"""Fields: ball trees built on the GPU (treecorr/field.py NField/KField/GField).

A Field owns a tc_field handle of the C ABI.  Construction uploads the
catalog columns and builds TreeCorr's mean-split tree on the device
(src/Field.cpp:163-277, src/Cell.cpp:556-615); the tree stays resident in HBM.
"""
import ctypes
import os

import numpy as np

from . import _lib

FLAT, THREED, SPHERE = 1, 2, 3
_COORDS = {'flat': FLAT, '3d': THREED, 'spherical': SPHERE}
_SPLIT = {'middle': 0, 'median': 1, 'mean': 2, 'random': 3}

NODE_DTYPE = np.dtype([('x', 'f8'), ('y', 'f8'), ('z', 'f8'), ('size', 'f4'), ('w', 'f4'),
                       ('a', 'f4'), ('b', 'f4'), ('child', 'i4'), ('n', 'i4')])


def coord_enum(coords):
    return _COORDS[coords]


def default_min_top():
    """Field._determine_top default, treecorr/field.py:94-104: max(3, ceil(log2(ncpu)))."""
    from .util import get_omp_threads
    n_cpu = get_omp_threads()
    return max(3, int.bit_length(n_cpu - 1))


class Field:
    _d = None

    def __init__(self, cat, *, min_size=0, max_size=None, split_method='mean', brute=False,
                 min_top=None, max_top=10, coords=None, rng=None, logger=None, _handle=None):
        self.ntot = cat.ntot
        self.min_size = float(min_size) if not brute else 0.
        self.max_size = float(max_size) if max_size is not None else np.inf
        self.split_method = split_method
        self._sm = _SPLIT[split_method]
        self.brute = bool(brute)
        if min_top is None:
            min_top = default_min_top()
        max_top = int(max_top)
        self.min_top = min(int(min_top), max_top)
        self.max_top = max_top
        self.coords = coords if coords is not None else cat.coords
        self._coords = coord_enum(self.coords)
        self._cat = cat
        if _handle is not None:     # built with the other patches, see build_patch_fields
            self.data = _handle
            return
        x = np.ascontiguousarray(cat.x, dtype=float)
        y = np.ascontiguousarray(cat.y, dtype=float)
        z = None if cat.z is None else np.ascontiguousarray(cat.z, dtype=float)
        w = np.ascontiguousarray(cat.w, dtype=float)
        wpos = None if cat.wpos is None else np.ascontiguousarray(cat.wpos, dtype=float)
        d1, d2 = self._payload(cat)
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().tc_field_create(
            self._d, self._coords, _lib.dptr(x), _lib.dptr(y), _lib.dptr(z), _lib.dptr(d1),
            _lib.dptr(d2), _lib.dptr(w), _lib.dptr(wpos), len(x), self.min_size, self.max_size,
            self._sm, 0, int(self.brute), self.min_top, self.max_top, ctypes.byref(h)))
        self.data = h

    def _payload(self, cat):
        return None, None

    def __del__(self):
        h = getattr(self, 'data', None)
        if h is not None and h.value:
            try:
                _lib.lib().tc_field_destroy(h)
            except Exception:
                pass
            self.data = None

    @property
    def nTopLevelNodes(self):
        n = ctypes.c_longlong()
        _lib.check(_lib.lib().tc_field_ntop(self.data, ctypes.byref(n)))
        return n.value

    @property
    def nobj(self):
        n = ctypes.c_longlong()
        _lib.check(_lib.lib().tc_field_nobj(self.data, ctypes.byref(n)))
        return n.value

    @property
    def size(self):
        s = ctypes.c_double()
        _lib.check(_lib.lib().tc_field_size(self.data, ctypes.byref(s)))
        return s.value

    @property
    def center(self):
        c = np.zeros(3)
        _lib.check(_lib.lib().tc_field_center(self.data, _lib.dptr(c)))
        return c

    def nodes(self):
        """(nodes structured array, top-level node ids) copied from the device."""
        nn = ctypes.c_longlong()
        _lib.check(_lib.lib().tc_field_nnodes(self.data, ctypes.byref(nn)))
        nodes = np.zeros(nn.value, dtype=NODE_DTYPE)
        top = np.zeros(self.nTopLevelNodes, dtype=np.int32)
        _lib.check(_lib.lib().tc_field_copy_nodes(self.data, nodes.ctypes.data_as(ctypes.c_void_p),
                                                  top.ctypes.data_as(ctypes.c_void_p)))
        return nodes, top


    # ---- k-means patches (treecorr/field.py:286-590) ----
    def run_kmeans(self, npatch, *, max_iter=200, tol=1.e-5, init='tree', alt=False, rng=None):
        """Patch labels and centres from k-means (field.py:286-404)."""
        centers = self.kmeans_initialize_centers(npatch, init=init, rng=rng)
        self.kmeans_refine_centers(centers, max_iter=max_iter, tol=tol, alt=alt)
        patches = self.kmeans_assign_patches(centers)
        return patches, centers

    def kmeans_initialize_centers(self, npatch, init='tree', *, rng=None):
        """Initial centres from the tree or at random objects (field.py:406-452); the
        same centres as the reference for the same rng."""
        from . import kmeans
        if npatch > self.ntot:
            raise ValueError("Invalid npatch.  Cannot be greater than self.ntot.")
        if npatch < 1:
            raise ValueError("Invalid npatch.  Cannot be less than 1.")
        seed = 0 if rng is None else int(rng.random() * 2**63)
        T = kmeans.TreeView.from_device_nodes(*self.nodes())
        if init == 'tree':
            centers = kmeans.init_tree(T, int(npatch), seed)
        elif init == 'random':
            centers = kmeans.init_random(T, int(npatch), seed)
        elif init == 'kmeans++':
            centers = kmeans.init_kmpp(T, int(npatch), seed, self._coords == THREED)
        else:
            raise ValueError("Invalid init: %s. " % init +
                             "Must be one of 'tree', 'random', or 'kmeans++.'")
        return centers[:, :2].copy() if self._coords == FLAT else centers

    def _kmeans_objects(self):
        # the tree's leaves: positions as the tree holds them (unit vectors on the
        # sphere), weights as the cells hold them (float)
        cat = self._cat
        x = np.ascontiguousarray(cat.x, dtype=float)
        y = np.ascontiguousarray(cat.y, dtype=float)
        z = None if self._coords == FLAT else np.ascontiguousarray(cat.z, dtype=float)
        if self._coords == SPHERE:
            r = np.sqrt(x * x + y * y + z * z)
            x, y, z = x / r, y / r, z / r
        w = np.ascontiguousarray(np.asarray(cat.w, dtype=np.float32), dtype=float)
        return x, y, z, w

    def kmeans_refine_centers(self, centers, *, max_iter=200, tol=1.e-5, alt=False):
        """Lloyd iterations on the GPU (field.py:454-560); centers is updated in place."""
        from . import kmeans
        x, y, z, w = self._kmeans_objects()
        if alt:
            return kmeans.refine_alt(self.data, centers, x, y, z, w, self.size,
                                     max_iter=max_iter, tol=tol)
        return kmeans.refine(centers, x, y, z, w, self.size, self.coords, max_iter=max_iter,
                             tol=tol)

    def kmeans_assign_patches(self, centers):
        """Each object's patch: its nearest centre (field.py:562-590)."""
        from . import kmeans
        x, y, z, _ = self._kmeans_objects()
        return kmeans.assign(centers, x, y, z).astype(int)


class NField(Field):
    _d = 0


class KField(Field):
    _d = 1

    def _payload(self, cat):
        return np.ascontiguousarray(cat.k, dtype=float), None


class _ComplexField(Field):
    _cols = None

    def _payload(self, cat):
        return tuple(np.ascontiguousarray(getattr(cat, c), dtype=float) for c in self._cols)


class ZField(_ComplexField):
    """Spin-0 complex field (treecorr/field.py ZField)."""
    _d = 2
    _cols = ('z1', 'z2')


class VField(_ComplexField):
    """Spin-1 (vector) field (treecorr/field.py VField)."""
    _d = 3
    _cols = ('v1', 'v2')


class TField(_ComplexField):
    """Spin-3 field (treecorr/field.py TField)."""
    _d = 5
    _cols = ('t1', 't2')


class QField(_ComplexField):
    """Spin-4 field (treecorr/field.py QField)."""
    _d = 6
    _cols = ('q1', 'q2')


class GField(Field):
    _d = 4

    def _payload(self, cat):
        return (np.ascontiguousarray(cat.g1, dtype=float),
                np.ascontiguousarray(cat.g2, dtype=float))


_KIND_CLS = {'N': 'NField', 'K': 'KField', 'Z': 'ZField', 'V': 'VField', 'G': 'GField',
             'T': 'TField', 'Q': 'QField'}
_PAYLOAD = {'K': ('_k', None), 'Z': ('_z1', '_z2'), 'V': ('_v1', '_v2'), 'G': ('_g1', '_g2'),
            'T': ('_t1', '_t2'), 'Q': ('_q1', '_q2')}


def build_patch_fields(kind, cats, **kw):
    """Fields of the patch catalogs cats in one device build (tc_field_create_multi),
    when they are all the non-empty patches of one catalog, in order: the patch trees
    are built in one level-synchronous pass over the parent's columns, already
    ordered by patch.  Each field lands in its catalog's field cache, as
    Catalog.get<kind>Field would build it.  Returns None when cats are not such a
    list."""
    parent = getattr(cats[0], '_split_parent', None)
    if parent is None or any(getattr(c, '_split_parent', None) is not parent for c in cats):
        return None
    if [c._single_patch for c in cats] != parent._split_patches:
        return None
    cols = parent._split_cols
    if kind == 'K' and cols.get('_k') is None:
        raise TypeError("k is not defined for this catalog")
    if kind == 'G' and (cols.get('_g1') is None or cols.get('_g2') is None):
        raise TypeError("g1,g2 are not defined for this catalog")
    cls = globals()[_KIND_CLS[kind]]
    proto = cls(cats[0], _handle=ctypes.c_void_p(), **kw)   # the parameters as a field holds them
    n1, n2 = _PAYLOAD.get(kind, (None, None))
    d1 = cols.get(n1) if n1 else None
    d2 = cols.get(n2) if n2 else None
    if (n1 and d1 is None) or (n2 and d2 is None):
        raise TypeError("the %s field's columns are not defined for this catalog" % kind)
    bounds = np.ascontiguousarray(parent._split_bounds, dtype=np.int64)
    n = int(bounds[-1])
    out = (ctypes.c_void_p * len(cats))()
    _lib.check(_lib.lib().tc_field_create_multi(
        cls._d, proto._coords, _lib.dptr(cols['_x']), _lib.dptr(cols['_y']),
        _lib.dptr(cols.get('_z')), _lib.dptr(d1), _lib.dptr(d2), _lib.dptr(cols['_w']),
        _lib.dptr(cols.get('_wpos')), n, bounds.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)),
        len(cats), proto.min_size, proto.max_size, proto._sm, 0, int(proto.brute), proto.min_top,
        proto.max_top, out))
    proto.data = None
    fields = [cls(c, _handle=ctypes.c_void_p(out[r]), **kw) for r, c in enumerate(cats)]
    key = (kind, tuple(sorted(kw.items())))
    for c, f in zip(cats, fields):
        c._fields[kind] = (key, f)
    return fields


def preorder(nodes, top):
    """Flatten a device tree into TreeCorr's pre-order cell list (for comparisons)."""
    order = []
    stack = list(reversed([int(t) for t in top]))
    while stack:
        i = stack.pop()
        order.append(i)
        c = int(nodes['child'][i])
        if c >= 0:
            stack.append(c + 1)
            stack.append(c)
    return np.array(order, dtype=np.int64)
