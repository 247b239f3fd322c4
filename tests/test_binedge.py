"""The oracle at the Log binning's decision points: pairs within a few ulp of every bin
edge and of singleBin's refined-test threshold (tests/binedge.py) give the sums that
the reference's arithmetic gives with the host libm's log."""
import numpy as np
import pytest

import binedge as B

RTOL = 1e-13


def close(a, b):
    return np.allclose(a, b, rtol=RTOL, atol=0.)


def test_leaf_sweeps_cross_every_edge():
    """Each edge's ulp sweep puts pairs on both sides of it, so a last-bit change of
    the log at the right pair moves it to the other bin."""
    for lay in B.LEAF_LAYOUTS:
        L = B.layout(*lay)
        rs = B.leaf_pairs(L)
        _, ks = B.expect_leaf(L, rs)
        per = len(B.LEAF_OFFSETS)
        for e in range(L['nbins'] - 1):
            got = set(ks[e * per:(e + 1) * per])
            assert got == {e, e + 1}, (lay, e, got)


def test_refine_sweeps_flip():
    cases = B.refine_cases()
    per = len(B.REFINE_OFFSETS)
    fits = [B.expect_refine(s, r)[1] for s, r in cases]
    for i in range(0, len(cases), per):
        assert 0 < sum(fits[i:i + per]) < per


def test_libm_log_rounds_these_cases_correctly():
    """The host libm's log agrees with the correctly rounded log on every decision of
    these cases, so the expectations below are both the reference's and the correctly
    rounded ones (the device's guard takes the correctly rounded log at edges)."""
    for lay in B.LEAF_LAYOUTS:
        L = B.layout(*lay)
        rs = B.leaf_pairs(L)
        assert np.array_equal(B.expect_leaf(L, rs)[1], B.expect_leaf(L, rs, B.cr_log)[1])
    for s, r in B.refine_cases():
        assert B.expect_refine(s, r)[1] == B.expect_refine(s, r, logf=B.cr_log)[1]


def test_oracle_leaf_edges(oracle_lib):
    O = oracle_lib
    for lay in B.LEAF_LAYOUTS:
        L = B.layout(*lay)
        rs = B.leaf_pairs(L)
        cat = B.leaf_catalog(L, rs)
        f = O.Field(0, 1, cat['x'], cat['y'], w=cat['w'], min_size=0., max_size=0.)
        ref = O.corr2(0, 1, f, None, L['nbins'], L['min_sep'], L['max_sep'], L['bin_size'],
                      0., 1., exact=True)
        exp, _ = B.expect_leaf(L, rs)
        np.testing.assert_array_equal(ref['npairs'], exp['npairs'])
        for k in ('weight', 'meanr', 'meanlogr'):
            assert close(ref[k], exp[k]), (lay, k)


def test_oracle_refine_edges(oracle_lib):
    O = oracle_lib
    P = B.REFINE
    for s, r in B.refine_cases():
        cat = B.refine_catalog(s, r)
        f = O.Field(0, 1, cat["x"], cat["y"], w=cat["w"], min_size=0., max_size=np.inf,
                    min_top=0)
        ref = O.corr2(0, 1, f, None, P['nbins'], P['min_sep'], P['max_sep'], P['bin_size'],
                      P['b'], P['a'], exact=True)
        exp, fits = B.expect_refine(s, r)
        np.testing.assert_array_equal(ref['npairs'], exp['npairs'])
        for k in ('weight', 'meanr', 'meanlogr'):
            assert close(ref[k], exp[k]), (s, r, fits, k)
