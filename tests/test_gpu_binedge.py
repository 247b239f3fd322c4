"""Bin decisions at the Log binning's edges on the GPU: pairs within a few ulp of every
bin edge and of singleBin's refined-test threshold (tests/binedge.py) land where the
reference's arithmetic with the host libm's log puts them (src/Corr2.cpp:492-513,
include/BinType.h:113-150).  The device's own log is within one ulp; its edge guard
takes the correctly rounded log for such pairs, which these cases show equals the libm
one (tests/test_binedge.py)."""
import numpy as np
import pytest

import binedge as B

RTOL = 1e-13


def close(a, b):
    return np.allclose(a, b, rtol=RTOL, atol=0.)


@pytest.mark.gpu
def test_gpu_leaf_pairs_at_bin_edges(gpu, oracle_lib):
    import capi
    O = oracle_lib
    for lay in B.LEAF_LAYOUTS:
        L = B.layout(*lay)
        rs = B.leaf_pairs(L)
        cat = B.leaf_catalog(L, rs)
        print('leaf edges', lay, len(rs), flush=True)
        f = capi.Field(0, 1, cat, min_size=0., max_size=0.)
        out = capi.corr2(0, f, None, L['nbins'], L['min_sep'], L['max_sep'], L['bin_size'],
                         0., 1.)
        exp, _ = B.expect_leaf(L, rs)
        np.testing.assert_array_equal(out['npairs'], exp['npairs'])
        for k in ('weight', 'meanr', 'meanlogr'):
            assert close(out[k], exp[k]), (lay, k)
        fo = O.Field(0, 1, cat['x'], cat['y'], w=cat['w'], min_size=0., max_size=0.)
        ref = O.corr2(0, 1, fo, None, L['nbins'], L['min_sep'], L['max_sep'], L['bin_size'],
                      0., 1., exact=True)
        for k in ('visits', 'accepted', 'leafpairs', 'refined'):
            assert out['counters'][k] == ref['counters'][k], (lay, k)


@pytest.mark.gpu
def test_gpu_refined_test_at_its_threshold(gpu, oracle_lib):
    import capi
    O = oracle_lib
    P = B.REFINE
    cases = B.refine_cases()
    print('refined-test cases', len(cases), flush=True)
    nfit = 0
    for s, r in cases:
        cat = B.refine_catalog(s, r)
        f = capi.Field(0, 1, cat, min_size=0., max_size=np.inf, min_top=0)
        out = capi.corr2(0, f, None, P['nbins'], P['min_sep'], P['max_sep'], P['bin_size'],
                         P['b'], P['a'])
        exp, fits = B.expect_refine(s, r)
        nfit += fits
        np.testing.assert_array_equal(out['npairs'], exp['npairs'])
        for k in ('weight', 'meanr', 'meanlogr'):
            assert close(out[k], exp[k]), (s, r, fits, k)
        fo = O.Field(0, 1, cat['x'], cat['y'], w=cat['w'], min_size=0., max_size=np.inf,
                     min_top=0)
        ref = O.corr2(0, 1, fo, None, P['nbins'], P['min_sep'], P['max_sep'], P['bin_size'],
                      P['b'], P['a'], exact=True)
        for k in ('visits', 'accepted', 'leafpairs', 'refined'):
            assert out['counters'][k] == ref['counters'][k], (s, r, k)
    assert 0 < nfit < len(cases)
