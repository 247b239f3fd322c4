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
@pytest.mark.parametrize('d,coords,n', [(4, 3, 200000), (1, 2, 100000), (0, 1, 150000)])
def test_device_tree_matches_oracle_large(gpu, oracle_lib, d, coords, n):
    O = oracle_lib
    cat, scale = C.make_catalog(coords, n, 40 + d, 0.3)
    kw = dict(min_size=scale * 2e-5, max_size=scale * 0.02, split_method=2, brute=False,
              min_top=5, max_top=10)
    f, nd, ntop = device_tree(gpu, d, coords, cat, kw)
    ref = O.Field(d, coords, cat['x'], cat['y'], cat['z'], k=cat['k'], g1=cat['g1'],
                  g2=cat['g2'], w=cat['w'], **kw).dump()
    assert_same_tree(nd, ntop, ref)
    if d == 4:
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
