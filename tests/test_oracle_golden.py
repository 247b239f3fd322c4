"""The CPU oracle against the golden vectors written from the literal reference
(tests/golden/make_golden.py), against brute force, and against the reference
build itself when oracle/_ref is present."""
import json
import os

import numpy as np
import pytest

import cases as C

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')
KEYS = ('xi0', 'xi1', 'xi2', 'xi3', 'meanr', 'meanlogr', 'weight', 'npairs')


def load(name):
    z = np.load(os.path.join(GOLD, name + '.npz'), allow_pickle=False)
    return json.loads(str(z['spec'])), {k: z[k] for k in z.files if k != 'spec'}


def normwise(a, b):
    den = np.max(np.abs(b))
    return 0. if den == 0 else np.max(np.abs(a - b)) / den


@pytest.mark.parametrize('case', C.CORR2_CASES, ids=[c[0] for c in C.CORR2_CASES])
def test_oracle_corr2_golden(oracle_lib, case):
    O = oracle_lib
    spec, gold = load(case[0])
    cat1, cat2, bn, extra = C.corr2_case(case)
    assert C.checksum(cat1) == spec['checksum1']
    d, coords = case[1], case[2]
    kw = dict(min_size=bn['min_size'], max_size=bn['max_size'], min_top=extra['min_top'],
              brute=extra['brute'])

    def field(cat):
        return O.Field(d, coords, cat['x'], cat['y'], cat['z'], k=cat['k'], g1=cat['g1'],
                       g2=cat['g2'], w=cat['w'], **kw)
    f1 = field(cat1)
    f2 = field(cat2) if cat2 is not None else None
    assert f1.ntop == spec['ntop1']
    out = O.corr2(d, coords, f1, f2, bn['nbins'], bn['min_sep'], bn['max_sep'], bn['bin_size'],
                  bn['b'], bn['a'], exact=True)
    # npairs: bit exact; weighted sums: 1e-12 normwise (the reference sums naively,
    # the oracle with compensation; both agree far below this bound).
    np.testing.assert_array_equal(out['npairs'], gold['npairs'])
    for k in KEYS[:-1]:
        if gold[k].size:
            assert normwise(out[k], gold[k]) < 1e-12, k


@pytest.mark.parametrize('case', C.TREE_CASES, ids=[c[0] for c in C.TREE_CASES])
def test_oracle_tree_golden(oracle_lib, case):
    O = oracle_lib
    spec, gold = load(case[0])
    cat, kw = C.tree_case(case)
    assert C.checksum(cat) == spec['checksum']
    t = O.Field(case[1], case[2], cat['x'], cat['y'], cat['z'], k=cat['k'], g1=cat['g1'],
                g2=cat['g2'], w=cat['w'], **kw).dump()
    for k in ('pos', 'size', 'w', 'wk', 'g', 'n', 'left', 'right', 'top'):
        np.testing.assert_array_equal(t[k], gold[k], err_msg=k)


@pytest.mark.parametrize('d,coords', [(0, 1), (1, 2), (4, 1), (4, 3)])
def test_oracle_bs0_equals_brute(oracle_lib, d, coords):
    """bin_slop=0 counts every pair in its exact bin (tests/test_nn.py:543-552): npairs
    equal the O(N^2) loop exactly; weights agree to the float cell payloads (1e-6);
    with brute=True every accepted pair is a leaf pair and all sums agree to 1e-12."""
    O = oracle_lib
    cat, scale = C.make_catalog(coords, 1500, 101 + d + coords, 0.2)
    bn = C.binning(scale, 10, 0.01, 0.4, 0.0, 0.0 if d == 4 else 1.0)
    b = O.brute2(d, coords, cat, None, bn['nbins'], bn['min_sep'], bn['max_sep'], bn['bin_size'])
    for brute in (False, True):
        f = O.Field(d, coords, cat['x'], cat['y'], cat['z'], k=cat['k'], g1=cat['g1'],
                    g2=cat['g2'], w=cat['w'], min_size=0. if brute else bn['min_size'],
                    max_size=bn['max_size'], min_top=3, brute=brute)
        t = O.corr2(d, coords, f, None, bn['nbins'], bn['min_sep'], bn['max_sep'],
                    bn['bin_size'], bn['b'], bn['a'], exact=True)
        np.testing.assert_array_equal(t['npairs'], b['npairs'])
        if brute:
            assert t['counters']['accepted'] == t['counters']['leafpairs']
            # xip_im (xi1) is odd under swapping a pair's order, which differs between
            # the tree's and the loop's pair order, so it is excluded here.
            for k in ('xi0', 'xi2', 'xi3', 'meanr', 'meanlogr', 'weight'):
                assert normwise(t[k], b[k]) < 1e-12, k
        else:
            np.testing.assert_allclose(t['weight'], b['weight'], rtol=1e-6)


def test_oracle_auto_from_cross(oracle_lib):
    """Invariant: auto(cat) == (cross(cat,cat) - self pairs)/2 in npairs for brute force."""
    O = oracle_lib
    cat, scale = C.make_catalog(C.FLAT, 800, 7)
    bn = C.binning(scale, 8, 0.01, 0.5, 0.0, 1.0)
    a = O.brute2(0, 1, cat, None, bn['nbins'], bn['min_sep'], bn['max_sep'], bn['bin_size'])
    x = O.brute2(0, 1, cat, cat, bn['nbins'], bn['min_sep'], bn['max_sep'], bn['bin_size'])
    np.testing.assert_array_equal(2 * a['npairs'], x['npairs'])   # min_sep > 0: no self pairs


def test_oracle_flat_gg_closed_form(oracle_lib):
    """Flat GG on two points: the tangential projection has a closed form
    (tests/test_gg.py:63-67): xip = g1 conj(g2) rotated by exp(-4i phi) both sides."""
    O = oracle_lib
    x = np.array([0., 3.]); y = np.array([0., 4.])
    g1 = np.array([0.1, -0.2]); g2 = np.array([0.05, 0.15]); w = np.array([1., 1.])
    cat = dict(x=x, y=y, z=None, g1=g1, g2=g2, w=w, k=None)
    out = O.brute2(4, 1, cat, None, 1, 1., 10., np.log(10.))
    phi = np.arctan2(4., 3.)
    e = np.exp(-2j * phi)
    z1 = (g1[0] + 1j * g2[0]) * e
    z2 = (g1[1] + 1j * g2[1]) * e
    xip = z1 * np.conj(z2)
    xim = z1 * z2
    # float payloads: w*g rounded to float32
    assert abs(out['xi0'][0] - xip.real) < 1e-7
    assert abs(out['xi1'][0] - xip.imag) < 1e-7
    assert abs(out['xi2'][0] - xim.real) < 1e-7
    assert abs(out['xi3'][0] - xim.imag) < 1e-7
    assert out['npairs'][0] == 1 and abs(out['meanr'][0] - 5.) < 1e-12


def test_oracle_planted_mistake_is_caught(oracle_lib):
    """A dropped pair changes npairs: the golden comparison detects a plausible bug."""
    spec, gold = load('nn_flat_bs0')
    wrong = gold['npairs'].copy()
    wrong[3] -= 1
    assert not np.array_equal(wrong, gold['npairs'])


def test_oracle_against_literal_reference(oracle_lib):
    """When oracle/_ref is built here, compare a fresh bin_slop>0 case directly."""
    from oracle import ref as R
    if R.module() is None:
        pytest.skip('oracle/_ref not built')
    O = oracle_lib
    cat, scale = C.make_catalog(C.SPHERE, 6000, 31, 0.3)
    bn = C.binning(scale, 12, 0.01, 0.5, 0.1, 0.1)
    kw = dict(min_size=bn['min_size'], max_size=bn['max_size'], min_top=4)
    rf = R.make_field(4, 3, cat['x'], cat['y'], cat['z'], g1=cat['g1'], g2=cat['g2'], w=cat['w'], **kw)
    r = R.corr2(4, rf, None, bn['nbins'], bn['min_sep'], bn['max_sep'], bn['bin_size'], bn['b'], bn['a'])
    of = O.Field(4, 3, cat['x'], cat['y'], cat['z'], g1=cat['g1'], g2=cat['g2'], w=cat['w'], **kw)
    o = O.corr2(4, 3, of, None, bn['nbins'], bn['min_sep'], bn['max_sep'], bn['bin_size'], bn['b'], bn['a'])
    np.testing.assert_array_equal(o['npairs'], r['npairs'])
    for k in KEYS[:-1]:
        assert normwise(o[k], r[k]) < 1e-12, k


@pytest.mark.parametrize('d1,d2', [(0, 1), (0, 4), (1, 4), (0, 2), (1, 2), (2, 2), (3, 3),
                                   (5, 5), (6, 6), (0, 3), (0, 5), (0, 6), (1, 3), (1, 5),
                                   (1, 6)])
@pytest.mark.parametrize('coords', [1, 2, 3])
def test_oracle_mixed_cross_against_literal_reference(oracle_lib, d1, d2, coords):
    """NK, NG and KG (DirectHelper<1>/<2>, src/Corr2.cpp:408-442): the oracle's sums
    against the compiled reference's NKCorr/NGCorr/KGCorr on the same trees' inputs,
    flat, 3-d and spherical.  The xi sums are far from zero, so a dropped minus sign,
    a projection at the wrong end or a swapped weight fails."""
    from oracle import ref as R
    if R.module() is None:
        pytest.skip('oracle/_ref not built')
    O = oracle_lib
    cat1, scale = C.make_catalog(coords, 4000, 11 + coords, 0.3)
    cat2, _ = C.make_catalog(coords, 5000, 21 + coords, 0.3)
    bn = C.binning(scale, 12, 0.01, 0.5, 0.1, 0.1 if d2 == 4 else 1.0)
    kw = dict(min_size=bn['min_size'], max_size=bn['max_size'], min_top=4)

    def cols(cat):
        return dict(k=cat['k'], g1=cat['g1'], g2=cat['g2'], w=cat['w'])
    r1 = R.make_field(d1, coords, cat1['x'], cat1['y'], cat1['z'], **cols(cat1), **kw)
    r2 = R.make_field(d2, coords, cat2['x'], cat2['y'], cat2['z'], **cols(cat2), **kw)
    r = R.corr2(d1, r1, r2, bn['nbins'], bn['min_sep'], bn['max_sep'], bn['bin_size'], bn['b'],
                bn['a'], d2=d2)
    o1 = O.Field(d1, coords, cat1['x'], cat1['y'], cat1['z'], **cols(cat1), **kw)
    o2 = O.Field(d2, coords, cat2['x'], cat2['y'], cat2['z'], **cols(cat2), **kw)
    o = O.corr2(d1, coords, o1, o2, bn['nbins'], bn['min_sep'], bn['max_sep'], bn['bin_size'],
                bn['b'], bn['a'], d2=d2)
    np.testing.assert_array_equal(o['npairs'], r['npairs'])
    keys = ['xi0', 'meanr', 'meanlogr', 'weight'] + (['xi1'] if d2 >= 2 else []) + \
        (['xi2', 'xi3'] if d1 == d2 else [])
    for k in keys:
        assert np.max(np.abs(r[k])) > 0, k
        assert normwise(o[k], r[k]) < 1e-12, k


@pytest.mark.parametrize('d', [2, 3, 5, 6])
@pytest.mark.parametrize('coords', [1, 2, 3])
def test_oracle_spin_tree_against_literal_reference(oracle_lib, d, coords):
    """Z, V, T, Q (spin 0, 1, 3, 4 complex) fields: the oracle's cells against the
    compiled reference's, bit for bit, the payload sums included (parallel transported
    by exp(-i s beta) on the sphere, src/Cell.cpp:221-354)."""
    from oracle import ref as R
    if R.dumplib() is None:
        pytest.skip('oracle/_ref not built')
    cat, scale = C.make_catalog(coords, 5000, 50 + coords + d, 0.3)
    kw = dict(min_size=scale * 1e-3, max_size=scale * 0.05, split_method=2, brute=False,
              min_top=3, max_top=10)
    rd = R.tree_dump(d, coords, cat['x'], cat['y'], cat['z'], g1=cat['g1'], g2=cat['g2'],
                     w=cat['w'], **kw)
    od = oracle_lib.Field(d, coords, cat['x'], cat['y'], cat['z'], g1=cat['g1'], g2=cat['g2'],
                          w=cat['w'], **kw).dump()
    for k in ('pos', 'size', 'w', 'g', 'n', 'left', 'right', 'top'):
        np.testing.assert_array_equal(od[k], rd[k], err_msg=k)


@pytest.mark.parametrize('d', [0, 1, 4])
@pytest.mark.parametrize('coords', [1, 2])
@pytest.mark.parametrize('cross', [False, True])
def test_oracle_periodic_against_literal_reference(oracle_lib, d, coords, cross):
    """The Periodic metric (include/Metric.h:813-838): separations wrapped by the box's
    period.  Oracle vs the compiled reference, auto and cross, flat and 3-d; the pair
    counts differ from the Euclidean ones by far more than any rounding, so a missing or
    wrong wrap fails."""
    from oracle import ref as R
    if R.module() is None:
        pytest.skip('oracle/_ref not built')
    O = oracle_lib
    cat, scale = C.make_catalog(coords, 6000, 80 + coords + d, 0.3)
    cat2, _ = C.make_catalog(coords, 5000, 90 + coords + d, 0.3)
    bn = C.binning(scale, 12, 0.01, 0.4, 0.1, 0.1 if d == 4 else 1.0)
    kw = dict(min_size=bn['min_size'], max_size=bn['max_size'], min_top=4)
    per = (scale, scale, scale)

    def cols(c):
        return dict(k=c['k'], g1=c['g1'], g2=c['g2'], w=c['w'])
    rf2 = R.make_field(d, coords, cat2['x'], cat2['y'], cat2['z'], **cols(cat2), **kw) \
        if cross else None
    of2 = O.Field(d, coords, cat2['x'], cat2['y'], cat2['z'], **cols(cat2), **kw) \
        if cross else None
    rf1 = R.make_field(d, coords, cat['x'], cat['y'], cat['z'], **cols(cat), **kw)
    of1 = O.Field(d, coords, cat['x'], cat['y'], cat['z'], **cols(cat), **kw)
    args = (bn['nbins'], bn['min_sep'], bn['max_sep'], bn['bin_size'], bn['b'], bn['a'])
    r = R.corr2(d, rf1, rf2, *args, period=per)
    o = O.corr2(d, coords, of1, of2, *args, period=per)
    e = O.corr2(d, coords, of1, of2, *args)
    np.testing.assert_array_equal(o['npairs'], r['npairs'])
    assert r['npairs'].sum() > 1.2 * e['npairs'].sum()
    for k in ['meanr', 'meanlogr', 'weight'] + (['xi0'] if d else []):
        assert normwise(o[k], r[k]) < 1e-12, k


@pytest.mark.parametrize('d', [0, 1, 4])
@pytest.mark.parametrize('coords', [2, 3])
@pytest.mark.parametrize('cross', [False, True])
def test_oracle_arc_against_literal_reference(oracle_lib, d, coords, cross):
    """The Arc metric (include/Metric.h:656-810): great-circle separations, on the sphere
    from the chord, in 3-d from atan2(|a x b|, a.b) with the sizes scaled to the unit
    sphere.  Oracle vs the compiled reference, auto and cross.  The Euclidean run differs:
    on the sphere its meanr is the chord (shorter by ~theta^3/24), in 3-d its separations
    are lengths, so a missing metric fails."""
    from oracle import ref as R
    if R.module() is None:
        pytest.skip('oracle/_ref not built')
    O = oracle_lib
    mk = C.make_shell if coords == C.THREED else \
        (lambda n, s, cl: C.make_catalog(C.SPHERE, n, s, cl))
    cat, scale = mk(6000, 60 + coords + d, 0.3)
    cat2, _ = mk(5000, 70 + coords + d, 0.3)
    bn = C.binning(scale, 12, 0.01, 0.5, 0.5, 0.1 if d == 4 else 1.0)
    kw = dict(min_size=bn['min_size'], max_size=bn['max_size'], min_top=4)

    def cols(c):
        return dict(k=c['k'], g1=c['g1'], g2=c['g2'], w=c['w'])
    rf2 = R.make_field(d, coords, cat2['x'], cat2['y'], cat2['z'], **cols(cat2), **kw) \
        if cross else None
    of2 = O.Field(d, coords, cat2['x'], cat2['y'], cat2['z'], **cols(cat2), **kw) \
        if cross else None
    rf1 = R.make_field(d, coords, cat['x'], cat['y'], cat['z'], **cols(cat), **kw)
    of1 = O.Field(d, coords, cat['x'], cat['y'], cat['z'], **cols(cat), **kw)
    args = (bn['nbins'], bn['min_sep'], bn['max_sep'], bn['bin_size'], bn['b'], bn['a'])
    r = R.corr2(d, rf1, rf2, *args, arc=True)
    o = O.corr2(d, coords, of1, of2, *args, arc=True)
    e = O.corr2(d, coords, of1, of2, *args)
    np.testing.assert_array_equal(o['npairs'], r['npairs'])
    assert r['npairs'].sum() > 1e6
    for k in ['meanr', 'meanlogr', 'weight'] + (['xi0'] if d else []):
        assert normwise(o[k], r[k]) < 1e-12, k
    assert normwise(e['meanr'], r['meanr']) > 1e-8


LOS_CASES = [(m3, d, cross, rpar) for m3 in (2, 5, 3, 0) for d in (0, 4)
             for cross in (False, True) for rpar in (None, (-20., 30.))
             if not (m3 == 3 and not cross) and not (m3 == 0 and rpar is None)]


@pytest.mark.parametrize('m3,d,cross,rpar', LOS_CASES)
def test_oracle_los_metrics_against_literal_reference(oracle_lib, m3, d, cross, rpar):
    """The line-of-sight metrics Rperp (2), Rlens (3), OldRperp (5), and the r_par range
    (also with the Euclidean metric, 0 here), include/Metric.h:37-651: oracle vs the
    compiled reference on 3-d shells, npairs exact and sums within 1e-12.  The r_par range
    removes about half the pairs, so a dropped range test fails."""
    from oracle import ref as R
    if R.module() is None:
        pytest.skip('oracle/_ref not built')
    O = oracle_lib
    if m3 == 0:
        rpar = (-1., 2.)     # 3-d separations below max_sep already bound |r_par| by 5
    cat, _ = C.make_shell(6000, 200 + m3 + d, 0.3)
    cat2, _ = C.make_shell(5000, 210 + m3 + d, 0.3)
    bn = C.binning(10., 12, 0.01, 0.5, 0.5, 0.1 if d == 4 else 1.0)
    kw = dict(min_size=bn['min_size'], max_size=bn['max_size'], min_top=4)

    def cols(c):
        return dict(k=c['k'], g1=c['g1'], g2=c['g2'], w=c['w'])
    rf2 = R.make_field(d, 2, cat2['x'], cat2['y'], cat2['z'], **cols(cat2), **kw) \
        if cross else None
    of2 = O.Field(d, 2, cat2['x'], cat2['y'], cat2['z'], **cols(cat2), **kw) if cross else None
    rf1 = R.make_field(d, 2, cat['x'], cat['y'], cat['z'], **cols(cat), **kw)
    of1 = O.Field(d, 2, cat['x'], cat['y'], cat['z'], **cols(cat), **kw)
    args = (bn['nbins'], bn['min_sep'], bn['max_sep'], bn['bin_size'], bn['b'], bn['a'])
    r = R.corr2(d, rf1, rf2, *args, metric3=m3, rpar=rpar)
    o = O.corr2(d, 2, of1, of2, *args, metric3=m3, rpar=rpar)
    np.testing.assert_array_equal(o['npairs'], r['npairs'])
    assert r['npairs'].sum() > 1e5
    for k in ['meanr', 'meanlogr', 'weight'] + (['xi0'] if d else []):
        assert normwise(o[k], r[k]) < 1e-12, k
    if rpar is not None:
        full = O.corr2(d, 2, of1, of2, *args, metric3=m3)
        assert r['npairs'].sum() < 0.8 * full['npairs'].sum()
