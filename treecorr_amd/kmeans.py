"""k-means patches: Field.run_kmeans and its three steps (treecorr/field.py:286-530).

The initial centres come from the field's tree on the host, restating the
reference's own choices (src/KMeans.cpp InitializeCentersTree :28-109,
InitializeCentersRand :111-157, SelectRandomFrom src/Corr2.cpp:622-662), with the
same random stream: glibc's rand() seeded by srand(seed) (urand, src/Cell.cpp:59-82),
so a seeded run starts from the reference's centres exactly.  The Lloyd iterations
run on the GPU over the objects (tc_kmeans_run); the final labels are the nearest
centres (tc_quick_assign).
"""
import ctypes
import os

import numpy as np

from . import _lib

_LIBC = None
_SEEDED = False
RAND_MAX = 2147483647        # glibc


def _libc():
    global _LIBC
    if _LIBC is None:
        _LIBC = ctypes.CDLL(None)
        _LIBC.rand.restype = ctypes.c_int
        _LIBC.srand.argtypes = [ctypes.c_uint]
    return _LIBC


def urand(seed=0):
    """urand (src/Cell.cpp:59-82): srand(seed) when seed != 0, else seeded from the
    system's entropy once; then rand()/RAND_MAX."""
    global _SEEDED
    L = _libc()
    if seed != 0:
        L.srand(ctypes.c_uint(seed & 0xffffffff).value)
        _SEEDED = True
    elif not _SEEDED:
        L.srand(int.from_bytes(os.urandom(4), 'little'))
        _SEEDED = True
    return L.rand() / RAND_MAX


def select_random_from(m, n):
    """n distinct integers from [0, m) (SelectRandomFrom, src/Corr2.cpp:622-662)."""
    sel = [0] * n
    if m > 3 * n:
        # Floyd's algorithm: each new value goes to the next slot
        chosen = set()
        while len(chosen) < n:
            j = int(urand() * m)
            if j == m:
                j = m - 1
            if j not in chosen:
                chosen.add(j)
                sel[len(chosen) - 1] = j
    else:
        # Fisher-Yates through the first n
        full = list(range(m))
        for i in range(n):
            j = int(urand() * (m - i)) + i
            if j == m:
                j = m - 1
            full[i], full[j] = full[j], full[i]
        sel = full[:n]
    return sel


class TreeView:
    """A tree as arrays: pos (nnode x 3), n, left, right (-1 at leaves), top (the
    top-level cells in TreeCorr's order)."""

    def __init__(self, pos, n, left, right, top, size=None):
        self.pos, self.n, self.left, self.right, self.top = pos, n, left, right, top
        self.size = size          # the cells' sizes, float32 as the cells hold them

    @classmethod
    def from_device_nodes(cls, nodes, top):
        """From the device layout (two children adjacent at child, child + 1)."""
        pos = np.column_stack([nodes['x'], nodes['y'], nodes['z']])
        left = np.asarray(nodes['child'], dtype=np.int64)
        right = np.where(left >= 0, left + 1, -1)
        return cls(pos, np.asarray(nodes['n'], dtype=np.int64), left, right,
                   np.asarray(top, dtype=np.int64), np.asarray(nodes['size'], dtype=np.float32))


def _init_tree_rec(T, centers, cell, first, ncenters):
    # InitializeCentersTree (recursive), src/KMeans.cpp:28-59
    if ncenters == 1:
        centers[first] = T.pos[cell]
    elif T.left[cell] >= 0:
        m1 = ncenters // 2
        m2 = ncenters - m1
        if urand() < 0.5:
            m1, m2 = m2, m1
        _init_tree_rec(T, centers, T.left[cell], first, m1)
        _init_tree_rec(T, centers, T.right[cell], first + m1, m2)
    else:
        for i in range(ncenters):
            centers[first + i] = T.pos[cell] * (1. + urand() * 1.e-8)


def init_tree(T, npatch, seed):
    """Initial centres from the tree (InitializeCentersTree, src/KMeans.cpp:61-109).
    Returns an (npatch, 3) array."""
    centers = np.zeros((npatch, 3))
    ncells = len(T.top)
    urand(seed)
    if ncells > npatch:
        sel = select_random_from(ncells, npatch)
        for i in range(npatch):
            centers[i] = T.pos[T.top[sel[i]]]
    else:
        n1 = npatch // ncells
        k2 = npatch % ncells
        k1 = ncells - k2
        nvalues = [n1] * k1 + [n1 + 1] * k2
        for k in range(ncells, 1, -1):
            j = int(urand() * k)
            if j != k - 1:
                nvalues[j], nvalues[k - 1] = nvalues[k - 1], nvalues[j]
        first = 0
        for k in range(ncells):
            _init_tree_rec(T, centers, T.top[k], first, nvalues[k])
            first += nvalues[k]
    return centers


def _leaf_number(T, cell, m):
    # BaseCell::getLeafNumber: descend by the counts of the left child
    while T.left[cell] >= 0:
        nl = T.n[T.left[cell]]
        if m < nl:
            cell = T.left[cell]
        else:
            m -= nl
            cell = T.right[cell]
    return cell


def init_random(T, npatch, seed):
    """Initial centres at random objects (InitializeCentersRand, src/KMeans.cpp:111-157)."""
    centers = np.zeros((npatch, 3))
    ntot = int(sum(T.n[c] for c in T.top))
    urand(seed)
    index = select_random_from(ntot, npatch)
    for i in range(npatch):
        m = index[i]
        for c in T.top:
            if m < T.n[c]:
                centers[i] = T.pos[_leaf_number(T, c, m)]
                break
            m -= T.n[c]
        for i1 in range(i):
            if np.array_equal(centers[i1], centers[i]):
                centers[i] = centers[i] * (1. + urand() * 1.e-8)
    return centers


class _LeafAlreadyUsed(Exception):
    pass


def _nsq(a, b):
    d = a - b
    return (d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]


def _kmpp_descend(T, cell, centers, ncen):
    # InitializeCentersKMPP (recursive), src/KMeans.cpp:159-218: from a chosen top cell,
    # walk down picking the daughter with probability proportional to its squared distance
    # to the nearest centre so far, to a leaf not yet used as a centre
    if T.size[cell] == 0:
        p = T.pos[cell]
        for j in range(ncen):
            if np.array_equal(p, centers[j]):
                raise _LeafAlreadyUsed()
        return p
    left, right = T.left[cell], T.right[cell]
    dsq1 = _nsq(T.pos[left], centers[0])
    dsq2 = _nsq(T.pos[right], centers[0])
    for j in range(1, ncen):
        d1 = _nsq(T.pos[left], centers[j])
        d2 = _nsq(T.pos[right], centers[j])
        if d1 < dsq1:
            dsq1 = d1
        if d2 < dsq2:
            dsq2 = d2
    if dsq1 + dsq2 == 0.:
        raise _LeafAlreadyUsed()
    u = urand() * (dsq1 + dsq2)
    first, second = (left, right) if u < dsq1 else (right, left)
    try:
        return _kmpp_descend(T, first, centers, ncen)
    except _LeafAlreadyUsed:
        return _kmpp_descend(T, second, centers, ncen)


def init_kmpp(T, npatch, seed, threed):
    """Initial centres by k-means++ over the tree (InitializeCentersKMPP,
    src/KMeans.cpp:220-313): the first at a random object; each next one from a top cell
    drawn with probability (n - centres already there) (d^2 + c s^2) (d the distance to the
    nearest centre, s the cell's size, c = 3/5 in 3-d, 1/2 otherwise), then by the descent
    above.  Returns an (npatch, 3) array."""
    centers = np.zeros((npatch, 3))
    top = [int(c) for c in T.top]
    ncells = len(top)
    urand(seed)
    ntot = int(sum(T.n[c] for c in top))
    per_cell = [0] * ncells
    m = int(urand() * ntot)
    for k, c in enumerate(top):
        if m < T.n[c]:
            centers[0] = T.pos[_leaf_number(T, c, m)]
            per_cell[k] += 1
            break
        m -= T.n[c]
    coef = 0.6 if threed else 0.5
    for i in range(1, npatch):
        p = [0.] * ncells
        sump = 0.
        for k, c in enumerate(top):
            dsq1 = _nsq(centers[0], T.pos[c])
            for i1 in range(1, i):
                dsq2 = _nsq(centers[i1], T.pos[c])
                if dsq2 < dsq1:
                    dsq1 = dsq2
            s = float(T.size[c])                  # the cell's float size, as a double
            ssq = s * s
            p[k] = float(T.n[c] - per_cell[k]) * (dsq1 + coef * ssq)
            sump += p[k]
        u = urand()
        for k, c in enumerate(top):
            p[k] /= sump
            if u < p[k]:
                centers[i] = _kmpp_descend(T, c, centers, i)
                per_cell[k] += 1
                break
            u -= p[k]
    return centers


def refine(centers, x, y, z, w, size, coords, max_iter=200, tol=1.e-5):
    """Lloyd iterations on the GPU (KMeansRun, src/KMeans.cpp:773-833): centres become
    the w-weighted means of their objects (normalised to the unit sphere for spherical
    coordinates) until the rms shift is below tol * size.  Returns the iterations run."""
    dim = centers.shape[1]
    c = np.ascontiguousarray(centers, dtype=float)
    niter = ctypes.c_int()
    _lib.check(_lib.lib().tc_kmeans_run(
        _lib.dptr(c), c.shape[0], dim, _lib.dptr(x), _lib.dptr(y),
        _lib.dptr(z) if dim == 3 else None, _lib.dptr(w), len(x), int(max_iter), float(tol),
        float(size), 1 if coords == 'spherical' else 0, ctypes.byref(niter)))
    centers[:] = c
    return niter.value


def refine_alt(field_handle, centers, x, y, z, w, size, max_iter=200, tol=1.e-5):
    """The alternate k-means (KMeansRun with alt=True, src/KMeans.cpp:774-833): each
    iteration estimates the patches' inertia on the field's tree, then assigns every
    object to its least |p - c|^2 + I (the inertia over the mean patch weight / 3)."""
    dim = centers.shape[1]
    c = np.ascontiguousarray(centers, dtype=float)
    niter = ctypes.c_int()
    _lib.check(_lib.lib().tc_kmeans_run_alt(
        field_handle, _lib.dptr(c), c.shape[0], _lib.dptr(x), _lib.dptr(y),
        _lib.dptr(z) if dim == 3 else None, _lib.dptr(w), len(x), int(max_iter), float(tol),
        float(size), ctypes.byref(niter)))
    centers[:] = c
    return niter.value


def assign(centers, x, y, z):
    """Each object's nearest centre, the lowest index on ties."""
    n = len(x)
    out = np.empty(n, dtype=np.int32)
    dim = centers.shape[1]
    c = np.ascontiguousarray(centers, dtype=float)
    _lib.check(_lib.lib().tc_quick_assign(_lib.dptr(c), c.shape[0], _lib.dptr(x), _lib.dptr(y),
                                          _lib.dptr(z) if dim == 3 else None, n,
                                          out.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))
    return out
