"""Device tree build vs the reference's own cells (golden dumps) and vs the oracle."""
import json
import os

import numpy as np
import pytest

import cases as C

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')
CNAME = {1: 'flat', 2: '3d', 3: 'spherical'}
SNAME = {0: 'middle', 1: 'median', 2: 'mean'}


def device_tree(tca, d, coords, cat, kw):
    """Build a device tree through the C ABI and return it in TreeCorr's pre-order."""
    import ctypes
    import capi
    from treecorr_amd import _lib
    from treecorr_amd.field import NODE_DTYPE, preorder
    f = capi.Field(d, coords, cat, **kw)
    L = _lib.lib()
    nn = ctypes.c_longlong()
    _lib.check(L.tc_field_nnodes(f.h, ctypes.byref(nn)))
    nodes = np.zeros(nn.value, dtype=NODE_DTYPE)
    top = np.zeros(f.ntop, dtype=np.int32)
    _lib.check(L.tc_field_copy_nodes(f.h, nodes.ctypes.data_as(ctypes.c_void_p),
                                     top.ctypes.data_as(ctypes.c_void_p)))
    order = preorder(nodes, top)
    return f, nodes[order], len(top)


def assert_same_tree(nd, ntop, ref):
    assert len(nd) == len(ref['n'])
    assert ntop == len(ref['top'])
    np.testing.assert_array_equal(nd['x'], ref['pos'][:, 0])
    np.testing.assert_array_equal(nd['y'], ref['pos'][:, 1])
    np.testing.assert_array_equal(nd['z'], ref['pos'][:, 2])
    np.testing.assert_array_equal(nd['size'], ref['size'])
    np.testing.assert_array_equal(nd['w'], ref['w'])
    np.testing.assert_array_equal(nd['n'], ref['n'])
    leaf = ref['left'] < 0
    np.testing.assert_array_equal(nd['child'] < 0, leaf)


@pytest.mark.gpu
@pytest.mark.parametrize('case', C.TREE_CASES, ids=[c[0] for c in C.TREE_CASES])
def test_device_tree_matches_reference_cells(gpu, case):
    z = np.load(os.path.join(GOLD, case[0] + '.npz'), allow_pickle=False)
    spec = json.loads(str(z['spec']))
    ref = {k: z[k] for k in z.files if k != 'spec'}
    cat, kw = C.tree_case(case)
    assert C.checksum(cat) == spec['checksum']
    d, coords = case[1], case[2]
    f, nd, ntop = device_tree(gpu, d, coords, cat, kw)
    assert_same_tree(nd, ntop, ref)
    if d == 1:
        np.testing.assert_array_equal(nd['a'], ref['wk'])
    if d == 4:
        np.testing.assert_array_equal(nd['a'], ref['g'][:, 0])
        np.testing.assert_array_equal(nd['b'], ref['g'][:, 1])


@pytest.mark.gpu
@pytest.mark.parametrize('d,coords,n', [(4, 3, 200000), (1, 2, 100000), (0, 1, 150000),
                                        (2, 3, 100000), (2, 2, 60000), (3, 3, 80000),
                                        (5, 2, 60000), (6, 3, 60000), (3, 1, 40000)])
def test_device_tree_matches_oracle_large(gpu, oracle_lib, d, coords, n):
    O = oracle_lib
    cat, scale = C.make_catalog(coords, n, 40 + d, 0.3)
    kw = dict(min_size=scale * 2e-5, max_size=scale * 0.02, split_method=2, brute=False,
              min_top=5, max_top=10)
    f, nd, ntop = device_tree(gpu, d, coords, cat, kw)
    ref = O.Field(d, coords, cat['x'], cat['y'], cat['z'], k=cat['k'], g1=cat['g1'],
                  g2=cat['g2'], w=cat['w'], **kw).dump()
    assert_same_tree(nd, ntop, ref)
    if d >= 2:
        np.testing.assert_array_equal(nd['a'], ref['g'][:, 0])
        np.testing.assert_array_equal(nd['b'], ref['g'][:, 1])
    # Field center and size (Field::_center/_sizesq)
    import ctypes
    from treecorr_amd import _lib
    ofield = O.Field(d, coords, cat['x'], cat['y'], cat['z'], k=cat['k'], g1=cat['g1'],
                     g2=cat['g2'], w=cat['w'], **kw)
    size = ctypes.c_double()
    _lib.check(_lib.lib().tc_field_size(f.h, ctypes.byref(size)))
    assert size.value == np.sqrt(ofield.sizesq)


@pytest.mark.gpu
def test_device_tree_degenerate_split_uses_median(gpu, oracle_lib):
    """Objects whose mean split leaves one side empty take std::nth_element's order
    (src/Cell.cpp:533-550): duplicated coordinates along the split axis."""
    O = oracle_lib
    rng = np.random.default_rng(5)
    n = 300
    x = np.repeat(rng.uniform(0, 1, 30), 10)      # many exact duplicates
    y = np.repeat(rng.uniform(0, 1e-3, 30), 10)
    x[:10] = 1.0
    x[10:20] = np.nextafter(1.0, 2.0)              # 1-ulp pair: mean rounds to one side
    y[:20] = 0.5
    cat = dict(x=x, y=y, z=None, w=np.ones(n), k=None, g1=rng.normal(0, .2, n),
               g2=rng.normal(0, .2, n))
    kw = dict(min_size=0., max_size=np.inf, split_method=2, brute=False, min_top=2, max_top=10)
    f, nd, ntop = device_tree(gpu, 4, 1, cat, kw)
    ref = O.Field(4, 1, x, y, None, g1=cat['g1'], g2=cat['g2'], w=cat['w'], **kw).dump()
    assert_same_tree(nd, ntop, ref)


def multi_trees(d, coords, cat, labels, kw):
    """tc_field_create_multi over the catalog ordered by label: one field per label."""
    import ctypes
    from treecorr_amd import _lib
    from treecorr_amd.field import NODE_DTYPE, preorder
    L = _lib.lib()
    order = np.argsort(labels, kind='stable')
    counts = np.bincount(labels)
    assert np.all(counts > 0)
    bounds = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    cols = {k: (None if cat.get(k) is None else np.ascontiguousarray(cat[k][order]))
            for k in ('x', 'y', 'z', 'k', 'g1', 'g2', 'w')}
    if coords == 1:
        cols['z'] = None
    d1 = cols['k'] if d == 1 else cols['g1'] if d == 4 else None
    d2 = cols['g2'] if d == 4 else None
    nroot = len(counts)
    out = (ctypes.c_void_p * nroot)()
    _lib.check(L.tc_field_create_multi(
        d, coords, _lib.dptr(cols['x']), _lib.dptr(cols['y']), _lib.dptr(cols['z']),
        _lib.dptr(d1), _lib.dptr(d2), _lib.dptr(cols['w']), None, len(order),
        bounds.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)), nroot, kw['min_size'],
        kw['max_size'], kw['split_method'], 0, int(kw['brute']), kw['min_top'], kw['max_top'],
        out))
    try:
        nn = ctypes.c_longlong()
        _lib.check(L.tc_field_nnodes(ctypes.c_void_p(out[0]), ctypes.byref(nn)))
        trees = []
        for r in range(nroot):
            h = ctypes.c_void_p(out[r])
            nt = ctypes.c_longlong()
            _lib.check(L.tc_field_ntop(h, ctypes.byref(nt)))
            nodes = np.zeros(nn.value, dtype=NODE_DTYPE)
            top = np.zeros(nt.value, dtype=np.int32)
            _lib.check(L.tc_field_copy_nodes(h, nodes.ctypes.data_as(ctypes.c_void_p),
                                             top.ctypes.data_as(ctypes.c_void_p)))
            size = ctypes.c_double()
            _lib.check(L.tc_field_size(h, ctypes.byref(size)))
            trees.append((nodes[preorder(nodes, top)], len(top), size.value,
                           {k: (None if v is None else v[bounds[r]:bounds[r + 1]])
                            for k, v in cols.items()}))
        return trees
    finally:
        for r in range(nroot):
            L.tc_field_destroy(ctypes.c_void_p(out[r]))


@pytest.mark.gpu
@pytest.mark.parametrize('d,coords,n,npatch,check', [
    (4, 3, 60000, 12, 12),      # 512-thread level 0
    (0, 1, 40000, 300, 40),     # 256-thread level 0
    (1, 2, 12000, 3000, 60),    # > 2048 roots: one thread or one wave per root at level 0
])
def test_multi_root_build_equals_per_patch_trees(gpu, oracle_lib, d, coords, n, npatch, check):
    """All patch trees built in one pass (the patches as level-0 segments) are the
    trees the oracle builds for each patch alone: same cells bit for bit, same top
    cells, same field size; tiny patches (one or two objects) included."""
    O = oracle_lib
    cat, scale = C.make_catalog(coords, n, 500 + npatch, 0.3)
    rng = np.random.default_rng(npatch)
    labels = rng.integers(0, npatch, n)
    labels[:npatch] = np.arange(npatch)      # every patch non-empty
    kw = dict(min_size=scale * 2e-4, max_size=scale * 0.05, split_method=2, brute=False,
              min_top=3, max_top=10)
    trees = multi_trees(d, coords, cat, labels, kw)
    counts = np.bincount(labels)
    pick = set(rng.choice(npatch, min(check, npatch), replace=False).tolist())
    pick |= set(np.where(counts <= 2)[0][:10].tolist())
    for r in sorted(pick):
        nd, ntop, size, c = trees[r]
        f = O.Field(d, coords, c['x'], c['y'], c['z'], k=c['k'], g1=c['g1'], g2=c['g2'],
                    w=c['w'], **kw)
        assert_same_tree(nd, ntop, f.dump())
        assert size == np.sqrt(f.sizesq), r


class _Speculation:
    """tc_set_tree_speculation for the duration of a test; counts the exact rebuilds."""

    def __init__(self, mode):
        from treecorr_amd import _lib
        self.L = _lib.lib()
        self.mode = mode

    def __enter__(self):
        self.before = self.L.tc_set_tree_speculation(-1)
        assert self.L.tc_set_tree_speculation(self.mode) == 0
        return self

    def __exit__(self, *exc):
        self.redo = self.L.tc_set_tree_speculation(-1) - self.before
        self.L.tc_set_tree_speculation(1)


@pytest.mark.gpu
@pytest.mark.parametrize('mode', [3, 2])
@pytest.mark.parametrize('d,coords,n,sm', [(4, 3, 60000, 2), (4, 2, 50000, 2),
                                           (1, 2, 30000, 0), (0, 1, 40000, 2)])
def test_speculative_build_matches_oracle(gpu, oracle_lib, mode, d, coords, n, sm):
    """Top levels partitioned from approximate centroids, then checked exactly: the
    tree is the oracle's bit for bit.  Mode 3 passes its check (random objects sit
    nowhere near a split plane's last bits), so no rebuild runs; mode 2 forces the
    rebuild, whose tree must be the same."""
    O = oracle_lib
    cat, scale = C.make_catalog(coords, n, 70 + d + coords, 0.3)
    kw = dict(min_size=scale * 2e-4, max_size=scale * 0.02, split_method=sm, brute=False,
              min_top=4, max_top=10)
    with _Speculation(mode) as s:
        f, nd, ntop = device_tree(gpu, d, coords, cat, kw)
    assert s.redo == (1 if mode == 2 else 0)
    ref = O.Field(d, coords, cat['x'], cat['y'], cat['z'], k=cat['k'], g1=cat['g1'],
                  g2=cat['g2'], w=cat['w'], **kw)
    assert_same_tree(nd, ntop, ref.dump())
    import ctypes
    from treecorr_amd import _lib
    size = ctypes.c_double()
    _lib.check(_lib.lib().tc_field_size(f.h, ctypes.byref(size)))
    assert size.value == np.sqrt(ref.sizesq)


@pytest.mark.gpu
def test_speculative_multi_root_build(gpu, oracle_lib):
    """Speculation over a multi-root build (12 patches at level 0)."""
    O = oracle_lib
    d, coords, n, npatch = 4, 3, 60000, 12
    cat, scale = C.make_catalog(coords, n, 512, 0.3)
    rng = np.random.default_rng(npatch)
    labels = rng.integers(0, npatch, n)
    kw = dict(min_size=scale * 2e-4, max_size=scale * 0.05, split_method=2, brute=False,
              min_top=3, max_top=10)
    with _Speculation(3) as s:
        trees = multi_trees(d, coords, cat, labels, kw)
    assert s.redo == 0
    for r in (0, 5, 11):
        nd, ntop, size, c = trees[r]
        f = O.Field(d, coords, c['x'], c['y'], c['z'], k=c['k'], g1=c['g1'], g2=c['g2'],
                    w=c['w'], **kw)
        assert_same_tree(nd, ntop, f.dump())
        assert size == np.sqrt(f.sizesq), r


@pytest.mark.gpu
def test_speculative_build_degenerate_split(gpu, oracle_lib):
    """The root's mean split within an ulp of a cluster of objects (20 objects at x = 1
    and the next double, the others balanced at 0.5 and 1.5): the validation by rounding
    bounds cannot vouch for the decision, so the build waits for the exact pass (and
    rebuilds if a decision differs); whichever side the approximate centroid puts them,
    the tree is the reference's."""
    O = oracle_lib
    rng = np.random.default_rng(5)
    n = 300
    x = np.concatenate([np.full(10, 1.0), np.full(10, np.nextafter(1.0, 2.0)),
                        np.full(140, 0.5), np.full(140, 1.5)])
    y = rng.uniform(0, 1e-3, n)
    cat = dict(x=x, y=y, z=None, w=np.ones(n), k=None, g1=rng.normal(0, .2, n),
               g2=rng.normal(0, .2, n))
    kw = dict(min_size=0., max_size=np.inf, split_method=2, brute=False, min_top=2, max_top=10)
    from treecorr_amd import _lib
    d0 = _lib.lib().tc_set_tree_speculation(-2)
    with _Speculation(3):
        f, nd, ntop = device_tree(gpu, 4, 1, cat, kw)
    # objects within the rounding bound of the split value: the build waits for the
    # exact pass (and rebuilds if a decision differs)
    assert _lib.lib().tc_set_tree_speculation(-2) == d0
    ref = O.Field(4, 1, x, y, None, g1=cat['g1'], g2=cat['g2'], w=cat['w'], **kw).dump()
    assert_same_tree(nd, ntop, ref)


@pytest.mark.gpu
def test_speculative_build_full_size_equals_exact_build(gpu):
    """At the headline size (2,000,000 galaxies on the sphere, GG, the benchmark's tree
    parameters) the default speculative build gives the same tree, node for node, as
    the build without speculation, and its check passes (no rebuild)."""
    cat, scale = C.make_catalog(C.SPHERE, 2000000, 2026, 0.3)
    kw = dict(min_size=scale * 1e-4, max_size=scale * 0.01, split_method=2, brute=False,
              min_top=4, max_top=10)
    from treecorr_amd import _lib
    L = _lib.lib()
    trees = []
    for mode in (1, 0):
        d0 = L.tc_set_tree_speculation(-2)
        with _Speculation(mode) as s:
            f, nd, ntop = device_tree(gpu, 4, 3, cat, kw)
        assert s.redo == 0
        # mode 1: the decisions pass the rounding-bound validation, so the build returned
        # before its exact pass (the node dump above waited for it)
        assert L.tc_set_tree_speculation(-2) - d0 == (1 if mode == 1 else 0)
        trees.append((nd, ntop))
        del f
    (a, ta), (b, tb) = trees
    assert ta == tb and len(a) == len(b)
    for k in ('x', 'y', 'z', 'size', 'w', 'a', 'b', 'n'):
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)
    np.testing.assert_array_equal(a['child'] < 0, b['child'] < 0)
