"""k-means initial centres: the host restatement of the reference's tree and random
initialisations (src/KMeans.cpp:28-157, SelectRandomFrom src/Corr2.cpp:622-662) on the
oracle's tree, against the compiled reference's KMeansInitTree/KMeansInitRand on its
own tree of the same catalog, with the same seed.  The two trees are the same cell for
cell (the tree tests pin that), and both draw from glibc's rand() seeded by srand, so
the centres must be identical."""
import numpy as np
import pytest

import cases as C


def _trees(O, R, coords, n, seed, max_top):
    cat, scale = C.make_catalog(coords, n, seed, 0.3)
    kw = dict(min_size=0., max_size=np.inf, split_method=2, brute=False, min_top=0,
              max_top=max_top)
    of = O.Field(0, coords, cat['x'], cat['y'], cat['z'], w=cat['w'], **kw)
    rf = R.make_field(0, coords, cat['x'], cat['y'], cat['z'], w=cat['w'], **kw)
    return of.dump(), rf


@pytest.mark.parametrize('coords,npatch,max_top', [(3, 20, 4), (1, 7, 2), (2, 40, 5),
                                                   (3, 3, 4)])
@pytest.mark.parametrize('init', ['tree', 'random', 'kmeans++'])
def test_kmeans_init_matches_compiled_reference(oracle_lib, coords, npatch, max_top, init):
    from oracle import ref as R
    from treecorr_amd import kmeans
    if R.module() is None:
        pytest.skip('oracle/_ref not built')
    d, rf = _trees(oracle_lib, R, coords, 5000, 90 + npatch, max_top)
    T = kmeans.TreeView(d['pos'], d['n'], d['left'], d['right'], d['top'], d['size'])
    dim = 2 if coords == 1 else 3
    name = {'tree': 'KMeansInitTree', 'random': 'KMeansInitRand', 'kmeans++': 'KMeansInitKMPP'}
    for seed in (12345, 987654321):
        ref = np.zeros((npatch, dim))
        getattr(rf, name[init])(ref, npatch, seed)
        if init == 'kmeans++':
            got = kmeans.init_kmpp(T, npatch, seed, coords == 2)[:, :dim]
        else:
            fn = kmeans.init_tree if init == 'tree' else kmeans.init_random
            got = fn(T, npatch, seed)[:, :dim]
        np.testing.assert_array_equal(got, ref)


def test_select_random_from_is_a_selection():
    """Floyd (m > 3n) and Fisher-Yates branches give n distinct values in [0, m)."""
    from treecorr_amd import kmeans
    kmeans.urand(7)
    for m, n in ((1000, 10), (20, 15), (5, 5)):
        s = kmeans.select_random_from(m, n)
        assert len(s) == n and len(set(s)) == n and min(s) >= 0 and max(s) < m
