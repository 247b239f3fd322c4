"""GPU 2-point correlations vs the literal reference's golden vectors and vs the oracle."""
import ctypes
import json
import math
import os

import numpy as np
import pytest

import cases as C

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')
CNAME = {1: 'flat', 2: '3d', 3: 'spherical'}
KEYS = ('xi0', 'xi1', 'xi2', 'xi3', 'meanr', 'meanlogr', 'weight', 'npairs')
# Tolerance for the FP64 sums (xi, meanr, meanlogr, weight), normwise relative: the
# device accumulates with compensation, the reference naively; 1e-12 is the bound.
RTOL = 1e-12


def normwise(a, b):
    den = np.max(np.abs(b))
    return 0. if den == 0 else np.max(np.abs(a - b)) / den


def raw_corr2(tca, d, coords, cat1, cat2, bn, brute=False, min_top=3, quick=False):
    """Run the GPU engine through the C ABI and return the raw (unfinalized) sums."""
    import capi
    kw = dict(min_size=0. if brute else bn['min_size'], max_size=bn['max_size'], brute=brute,
              min_top=min_top)
    f1 = capi.Field(d, coords, cat1, **kw)
    f2 = capi.Field(d, coords, cat2, **kw) if cat2 is not None else None
    return capi.corr2(d, f1, f2, bn['nbins'], bn['min_sep'], bn['max_sep'], bn['bin_size'],
                      bn['b'], bn['a'], quick=quick)


@pytest.mark.gpu
@pytest.mark.parametrize('case', C.CORR2_CASES, ids=[c[0] for c in C.CORR2_CASES])
def test_gpu_corr2_matches_reference_golden(gpu, case):
    z = np.load(os.path.join(GOLD, case[0] + '.npz'), allow_pickle=False)
    spec = json.loads(str(z['spec']))
    gold = {k: z[k] for k in z.files if k != 'spec'}
    cat1, cat2, bn, extra = C.corr2_case(case)
    assert C.checksum(cat1) == spec['checksum1']
    out = raw_corr2(gpu, case[1], case[2], cat1, cat2, bn, brute=extra['brute'],
                    min_top=extra['min_top'])
    np.testing.assert_array_equal(out['npairs'], gold['npairs'])
    for k in KEYS[:-1]:
        if gold[k].size:
            assert normwise(out[k], gold[k]) < RTOL, k


@pytest.mark.gpu
@pytest.mark.parametrize('d,coords,bs', [(0, 1, 0.0), (1, 2, 0.1), (4, 3, 0.1), (4, 1, 1.0),
                                         (0, 3, 0.5), (4, 2, 0.2)])
def test_gpu_corr2_matches_oracle_large(gpu, oracle_lib, d, coords, bs):
    O = oracle_lib
    cat, scale = C.make_catalog(coords, 60000, 70 + d + coords, 0.3)
    bn = C.binning(scale, 20, 0.002, 0.3, bs, 0.1 if d == 4 else 1.0)
    out = raw_corr2(gpu, d, coords, cat, None, bn, min_top=5)
    kw = dict(min_size=bn['min_size'], max_size=bn['max_size'], min_top=5)
    f = O.Field(d, coords, cat['x'], cat['y'], cat['z'], k=cat['k'], g1=cat['g1'], g2=cat['g2'],
                w=cat['w'], **kw)
    ref = O.corr2(d, coords, f, None, bn['nbins'], bn['min_sep'], bn['max_sep'], bn['bin_size'],
                  bn['b'], bn['a'], exact=True)
    np.testing.assert_array_equal(out['npairs'], ref['npairs'])
    for k in KEYS[:-1]:
        if out[k].size:
            assert normwise(out[k], ref[k]) < RTOL, k
    for k in ('visits', 'accepted', 'leafpairs', 'refined'):
        assert out['counters'][k] == ref['counters'][k], k


@pytest.mark.gpu
def test_gpu_corr2_cross_and_quick(gpu, oracle_lib):
    O = oracle_lib
    cat1, scale = C.make_catalog(C.SPHERE, 20000, 91, 0.2)
    cat2, _ = C.make_catalog(C.SPHERE, 15000, 92, 0.2)
    bn = C.binning(scale, 12, 0.005, 0.4, 0.3, 0.1)
    out = raw_corr2(gpu, 4, 3, cat1, cat2, bn, min_top=4, quick=True)
    kw = dict(min_size=bn['min_size'], max_size=bn['max_size'], min_top=4)
    f1 = O.Field(4, 3, cat1['x'], cat1['y'], cat1['z'], g1=cat1['g1'], g2=cat1['g2'], w=cat1['w'], **kw)
    f2 = O.Field(4, 3, cat2['x'], cat2['y'], cat2['z'], g1=cat2['g1'], g2=cat2['g2'], w=cat2['w'], **kw)
    ref = O.corr2(4, 3, f1, f2, bn['nbins'], bn['min_sep'], bn['max_sep'], bn['bin_size'],
                  bn['b'], bn['a'], quick=True, exact=True)
    for k in ('xi0', 'xi1', 'xi2', 'xi3', 'weight'):
        assert normwise(out[k], ref[k]) < RTOL, k
    assert not np.any(out['npairs']) and not np.any(out['meanr'])


@pytest.mark.gpu
def test_gpu_nn_bs0_exact_vs_brute(gpu, oracle_lib):
    """NN flat at bin_slop=0 is exact vs the O(N^2) pair loop (the 100k-point workload's
    check, here at a size the CPU loop finishes quickly)."""
    O = oracle_lib
    rng = np.random.default_rng(2026)
    n = 20000
    cat = dict(x=rng.uniform(0, 1000, n), y=rng.uniform(0, 1000, n), z=None, w=np.ones(n),
               k=None, g1=None, g2=None)
    # bin_slop = 0: b = 0, so min_size = max_size = 0 (corr2base.py:1595-1625)
    bn = dict(nbins=20, min_sep=1., max_sep=100., bin_size=math.log(100.) / 20, b=0., a=1.,
              min_size=0., max_size=0.)
    out = raw_corr2(gpu, 0, 1, cat, None, bn, min_top=3)
    b = O.brute2(0, 1, cat, None, 20, 1., 100., math.log(100.) / 20)
    np.testing.assert_array_equal(out['npairs'], b['npairs'])
    np.testing.assert_allclose(out['weight'], b['weight'], rtol=1e-12)


@pytest.mark.gpu
def test_gpu_finalized_api(gpu, oracle_lib):
    """The public process()/finalize path: xip/xim/meanr/varxi as TreeCorr defines them."""
    cat, scale = C.make_catalog(C.SPHERE, 30000, 95, 0.3)
    c = gpu.Catalog(ra=np.arctan2(cat['y'], cat['x']), dec=np.arcsin(cat['z']), ra_units='rad',
                    dec_units='rad', g1=cat['g1'], g2=cat['g2'], w=cat['w'])
    gg = gpu.GGCorrelation(nbins=10, min_sep=1., max_sep=100., sep_units='arcmin',
                           bin_slop=0.2, min_top=4)
    gg.process(c)
    assert np.all(np.isfinite(gg.xip)) and np.all(gg.npairs > 0)
    assert np.all((gg.meanr > gg.left_edges * 0.9) & (gg.meanr < gg.right_edges * 1.1))
    varg = c.varg
    np.testing.assert_allclose(gg.varxip, 2 * varg ** 2 / gg.weight, rtol=1e-12)


@pytest.mark.gpu
@pytest.mark.parametrize('chunk', [3000, 40000])
def test_gpu_small_chunks_same_result(gpu, oracle_lib, chunk):
    """The traversal's work stack across many launches (children read in place, the
    excess moved onto the stack) gives the same sums as the oracle."""
    from treecorr_amd import _lib
    O = oracle_lib
    cat, scale = C.make_catalog(C.SPHERE, 8000, 97, 0.3)
    bn = C.binning(scale, 15, 0.005, 0.3, 0.1, 0.1)
    L = _lib.lib()
    _lib.check(L.tc_set_traversal_chunk(chunk))
    _lib.check(L.tc_set_traversal_mode(0))
    try:
        out = raw_corr2(gpu, 4, 3, cat, None, bn, min_top=5)
    finally:
        _lib.check(L.tc_set_traversal_chunk(0))
        _lib.check(L.tc_set_traversal_mode(1))
    kw = dict(min_size=bn['min_size'], max_size=bn['max_size'], min_top=5)
    f = O.Field(4, 3, cat['x'], cat['y'], cat['z'], g1=cat['g1'], g2=cat['g2'], w=cat['w'], **kw)
    ref = O.corr2(4, 3, f, None, bn['nbins'], bn['min_sep'], bn['max_sep'], bn['bin_size'],
                  bn['b'], bn['a'], exact=True)
    np.testing.assert_array_equal(out['npairs'], ref['npairs'])
    for k in KEYS[:-1]:
        assert normwise(out[k], ref[k]) < RTOL, k
    for k in ('visits', 'accepted', 'leafpairs', 'refined'):
        assert out['counters'][k] == ref['counters'][k], k


@pytest.mark.gpu
@pytest.mark.parametrize('d1,d2', [(0, 1), (0, 4), (1, 4), (0, 2), (1, 2), (2, 2), (3, 3),
                                   (5, 5), (6, 6), (0, 3), (0, 5), (0, 6), (1, 3), (1, 5),
                                   (1, 6)])
@pytest.mark.parametrize('coords', [1, 2, 3])
def test_gpu_mixed_cross_matches_oracle(gpu, oracle_lib, d1, d2, coords):
    """NK, NG and KG cross correlations (src/Corr2.cpp DirectHelper<1>/<2>) against the
    oracle, itself pinned to the compiled reference: npairs and the node-pair counters
    exact, the sums to RTOL."""
    import capi
    O = oracle_lib
    cat1, scale = C.make_catalog(coords, 30000, 300 + 10 * d1 + d2 + coords, 0.3)
    cat2, _ = C.make_catalog(coords, 40000, 400 + 10 * d1 + d2 + coords, 0.3)
    bn = C.binning(scale, 15, 0.003, 0.3, 0.1, 0.1 if d2 == 4 else 1.0)
    kw = dict(min_size=bn['min_size'], max_size=bn['max_size'], min_top=4)
    f1 = capi.Field(d1, coords, cat1, brute=False, **kw)
    f2 = capi.Field(d2, coords, cat2, brute=False, **kw)
    out = capi.corr2(d1, f1, f2, bn['nbins'], bn['min_sep'], bn['max_sep'], bn['bin_size'],
                     bn['b'], bn['a'], d2=d2)

    def ofield(d, cat):
        return O.Field(d, coords, cat['x'], cat['y'], cat['z'], k=cat['k'], g1=cat['g1'],
                       g2=cat['g2'], w=cat['w'], **kw)
    ref = O.corr2(d1, coords, ofield(d1, cat1), ofield(d2, cat2), bn['nbins'], bn['min_sep'],
                  bn['max_sep'], bn['bin_size'], bn['b'], bn['a'], exact=True, d2=d2)
    np.testing.assert_array_equal(out['npairs'], ref['npairs'])
    keys = ['xi0', 'meanr', 'meanlogr', 'weight'] + (['xi1'] if d2 >= 2 else []) + \
        (['xi2', 'xi3'] if d1 == d2 else [])
    for k in keys:
        assert np.max(np.abs(ref[k])) > 0, k
        assert normwise(out[k], ref[k]) < RTOL, k
    for k in ('visits', 'accepted', 'leafpairs', 'refined'):
        assert out['counters'][k] == ref['counters'][k], k


@pytest.mark.gpu
def test_gpu_mixed_rejects_auto_and_wrong_sides(gpu):
    """A mixed correlation is cross only, with the d1 field first."""
    import ctypes
    import capi
    from treecorr_amd import _lib
    cat, scale = C.make_catalog(C.FLAT, 500, 5, 0.)
    kw = dict(min_size=0., max_size=1., min_top=2, brute=False)
    fn = capi.Field(0, 1, cat, **kw)
    fg = capi.Field(4, 1, cat, **kw)
    L = _lib.lib()
    xi0, xi1, m, ml, w, npr = (np.zeros(5) for _ in range(6))
    h = ctypes.c_void_p()
    big = 1.7976931348623157e308     # no r_par range
    _lib.check(L.tc_corr2_create(0, 4, 0, 1., 10., 5, np.log(10.) / 5, 0.1, 1., -big, big,
                                 0., 0., 0., _lib.dptr(xi0), _lib.dptr(xi1), None, None,
                                 _lib.dptr(m), _lib.dptr(ml), _lib.dptr(w), _lib.dptr(npr),
                                 ctypes.byref(h)))
    try:
        assert L.tc_corr2_process_auto(h, fn.h, 0, 0, 1) != 0
        assert L.tc_corr2_process_cross(h, fg.h, fn.h, 0, 0, 1) != 0
        _lib.check(L.tc_corr2_process_cross(h, fn.h, fg.h, 0, 0, 1))
    finally:
        L.tc_corr2_destroy(h)


@pytest.mark.gpu
def test_gpu_mixed_finalized_api(gpu):
    """NG, NK and KG through the public classes: xi is the raw sum over the weight,
    the shot-noise variance is var/weight (ngcorrelation.py, nkcorrelation.py,
    kzcorrelation.py finalize), and calculateXi(rg=...) subtracts the random-point
    correlation and adds its variance."""
    def sph(cat):
        return dict(ra=np.arctan2(cat['y'], cat['x']), dec=np.arcsin(cat['z']), ra_units='rad',
                    dec_units='rad')
    lc, _ = C.make_catalog(C.SPHERE, 8000, 601, 0.3)
    sc, _ = C.make_catalog(C.SPHERE, 20000, 602, 0.3)
    rc, _ = C.make_catalog(C.SPHERE, 16000, 603, 0.0)
    lens = gpu.Catalog(**sph(lc), w=lc['w'], k=lc['k'])
    src = gpu.Catalog(**sph(sc), g1=sc['g1'], g2=sc['g2'], k=sc['k'], w=sc['w'])
    rand = gpu.Catalog(**sph(rc))
    kw = dict(nbins=10, min_sep=1., max_sep=100., sep_units='arcmin', bin_slop=0.1)
    ng = gpu.NGCorrelation(**kw)
    ng.process(lens, src)
    assert np.all(ng.npairs > 0) and np.all(np.isfinite(ng.xi)) and np.all(np.isfinite(ng.xi_im))
    np.testing.assert_allclose(ng.varxi, src.varg / ng.weight, rtol=1e-12)
    rg = gpu.NGCorrelation(**kw)
    rg.process(rand, src)
    raw = ng.xi.copy()
    xi, xi_im, varxi = ng.calculateXi(rg=rg)
    np.testing.assert_allclose(xi, raw - rg.xi, rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(varxi, ng.raw_varxi + rg.varxi, rtol=1e-12)
    np.testing.assert_array_equal(ng.raw_xi, raw)

    nk = gpu.NKCorrelation(**kw)
    nk.process(lens, src)
    np.testing.assert_allclose(nk.varxi, src.vark / nk.weight, rtol=1e-12)
    kg = gpu.KGCorrelation(**kw)
    kg.process(lens, src)
    np.testing.assert_allclose(kg.varxi, lens.vark * src.varg / kg.weight, rtol=1e-12)
    with pytest.raises(TypeError):
        kg.process(lens)


@pytest.mark.gpu
def test_gpu_gg_2m_matches_compiled_reference(gpu):
    """At the benchmark's size (2,000,000 galaxies on the sphere, GG, bin_slop 0.1) the
    GPU's raw sums against TreeCorr's own C++ engine compiled from its sources: npairs
    identical in every bin, the FP64 sums to RTOL."""
    import capi
    from oracle import ref as R
    if R.module() is None:
        pytest.skip('oracle/_ref not built')
    cat, scale = C.make_catalog(C.SPHERE, 2000000, 2027, 0.3)
    bn = C.binning(scale, 25, 0.003, 0.6, 0.1, 0.1)
    kw = dict(min_size=bn['min_size'], max_size=bn['max_size'], min_top=4, max_top=10)
    out = raw_corr2(gpu, 4, 3, cat, None, bn, min_top=4)
    rf = R.make_field(4, 3, cat['x'], cat['y'], cat['z'], g1=cat['g1'], g2=cat['g2'],
                      w=cat['w'], **kw)
    ref = R.corr2(4, rf, None, bn['nbins'], bn['min_sep'], bn['max_sep'], bn['bin_size'],
                  bn['b'], bn['a'])
    assert ref['npairs'].sum() > 1e10
    np.testing.assert_array_equal(out['npairs'], ref['npairs'])
    for k in KEYS[:-1]:
        assert normwise(out[k], ref[k]) < RTOL, k


@pytest.mark.gpu
@pytest.mark.parametrize('letter,d', [('Z', 2), ('V', 3), ('T', 5), ('Q', 6)])
def test_gpu_spin_finalized_api(gpu, oracle_lib, letter, d):
    """ZZ/VV/TT/QQ, NZ/NV/NT/NQ and KZ/KV/KT/KQ through the public classes: the auto
    correlation's raw sums equal the oracle's (the catalog's two components as the
    field's complex values), xi is raw/weight, the shot variance is 2 var^2/weight
    (auto), var/weight (N), vark var/weight (K); the random-point compensation takes
    its keyword after the field's letter."""
    O = oracle_lib
    c1, scale = C.make_catalog(C.THREED, 12000, 611 + d, 0.3)
    c2, _ = C.make_catalog(C.THREED, 15000, 612 + d, 0.3)
    lo = letter.lower()
    kw2 = {lo + '1': c2['g1'], lo + '2': c2['g2']}
    zc = gpu.Catalog(x=c2['x'], y=c2['y'], z=c2['z'], w=c2['w'], **kw2)
    lc = gpu.Catalog(x=c1['x'], y=c1['y'], z=c1['z'], w=c1['w'], k=c1['k'])
    r3, _ = C.make_catalog(C.THREED, 8000, 613 + d, 0.0)
    rc = gpu.Catalog(x=r3['x'], y=r3['y'], z=r3['z'])
    var = getattr(zc, 'var' + lo)
    kw = dict(nbins=10, min_sep=scale * 0.01, max_sep=scale * 0.4, bin_slop=0.1, min_top=4)
    zz = getattr(gpu, letter * 2 + 'Correlation')(**kw)
    zz.process(zc)
    np.testing.assert_allclose(zz.varxip, 2 * var ** 2 / zz.weight, rtol=1e-12)
    min_size, max_size = zz._get_minmax_size()
    of = O.Field(d, C.THREED, c2['x'], c2['y'], c2['z'], g1=c2['g1'], g2=c2['g2'], w=c2['w'],
                 min_size=min_size, max_size=max_size, min_top=4)
    ref = O.corr2(d, C.THREED, of, None, zz.nbins, zz._min_sep, zz._max_sep, zz._bin_size,
                  zz.b, zz.angle_slop, exact=True)
    np.testing.assert_array_equal(zz.npairs, ref['npairs'])
    assert normwise(zz.xip * zz.weight, ref['xi0']) < RTOL
    assert normwise(zz.xim * zz.weight, ref['xi2']) < RTOL

    nz = getattr(gpu, 'N' + letter + 'Correlation')(**kw)
    nz.process(lc, zc)
    np.testing.assert_allclose(nz.varxi, var / nz.weight, rtol=1e-12)
    rz = getattr(gpu, 'N' + letter + 'Correlation')(**kw)
    rz.process(rc, zc)
    raw = nz.xi.copy()
    xi, xi_im, varxi = nz.calculateXi(**{'r' + lo: rz})
    np.testing.assert_allclose(xi, raw - rz.xi, rtol=1e-12, atol=1e-15)
    kz = getattr(gpu, 'K' + letter + 'Correlation')(**kw)
    kz.process(lc, zc)
    np.testing.assert_allclose(kz.varxi, lc.vark * var / kz.weight, rtol=1e-12)


@pytest.mark.gpu
@pytest.mark.parametrize('d,coords,cross', [(0, 1, False), (4, 1, True), (1, 2, False),
                                            (0, 2, True), (4, 2, False)])
def test_gpu_periodic_matches_oracle(gpu, oracle_lib, d, coords, cross):
    """The Periodic metric on the GPU against the oracle (itself pinned to the compiled
    reference): npairs and the node-pair counters exact, the sums to RTOL."""
    import capi
    O = oracle_lib
    cat, scale = C.make_catalog(coords, 30000, 120 + d + coords, 0.3)
    cat2, _ = C.make_catalog(coords, 25000, 130 + d + coords, 0.3) if cross else (None, 0)
    bn = C.binning(scale, 15, 0.003, 0.35, 0.1, 0.1 if d == 4 else 1.0)
    kw = dict(min_size=bn['min_size'], max_size=bn['max_size'], min_top=4)
    per = (scale, scale, scale)
    f1 = capi.Field(d, coords, cat, brute=False, **kw)
    f2 = capi.Field(d, coords, cat2, brute=False, **kw) if cross else None
    out = capi.corr2(d, f1, f2, bn['nbins'], bn['min_sep'], bn['max_sep'], bn['bin_size'],
                     bn['b'], bn['a'], period=per)

    def ofield(c):
        return O.Field(d, coords, c['x'], c['y'], c['z'], k=c['k'], g1=c['g1'], g2=c['g2'],
                       w=c['w'], **kw)
    ref = O.corr2(d, coords, ofield(cat), ofield(cat2) if cross else None, bn['nbins'],
                  bn['min_sep'], bn['max_sep'], bn['bin_size'], bn['b'], bn['a'], exact=True,
                  period=per)
    np.testing.assert_array_equal(out['npairs'], ref['npairs'])
    for k in ['meanr', 'meanlogr', 'weight'] + (['xi0'] if d else []):
        assert normwise(out[k], ref[k]) < RTOL, k
    for k in ('visits', 'accepted', 'leafpairs', 'refined'):
        assert out['counters'][k] == ref['counters'][k], k


@pytest.mark.gpu
def test_gpu_periodic_front_end(gpu, oracle_lib):
    """NNCorrelation(metric='Periodic', period=L) through the public class: trees to the
    leaves (corr2base.py:1620-1624) and the pair counts of the periodic box equal the
    oracle's."""
    O = oracle_lib
    cat, scale = C.make_catalog(C.THREED, 20000, 141, 0.3)
    tcat = gpu.Catalog(x=cat['x'], y=cat['y'], z=cat['z'])
    nn = gpu.NNCorrelation(nbins=10, min_sep=scale * 0.01, max_sep=scale * 0.3,
                           metric='Periodic', period=scale, min_top=4)
    nn.process(tcat)
    f = O.Field(0, C.THREED, cat['x'], cat['y'], cat['z'], min_size=0., max_size=0., min_top=4)
    ref = O.corr2(0, C.THREED, f, None, nn.nbins, nn._min_sep, nn._max_sep, nn._bin_size,
                  nn.b, nn.angle_slop, exact=True, period=(scale, scale, scale))
    np.testing.assert_array_equal(nn.npairs, ref['npairs'])
    with pytest.raises(ValueError):
        gpu.NNCorrelation(nbins=10, min_sep=1., max_sep=10., metric='Periodic').process(tcat)
    with pytest.raises(ValueError):
        gpu.NNCorrelation(nbins=10, min_sep=1., max_sep=10., period=5.).process(tcat)


@pytest.mark.gpu
@pytest.mark.parametrize('d,coords,cross', [(0, 3, False), (4, 3, True), (1, 3, True),
                                            (0, 2, True), (4, 2, False), (1, 2, False)])
def test_gpu_arc_matches_oracle(gpu, oracle_lib, d, coords, cross):
    """The Arc metric on the GPU against the oracle (itself pinned to the compiled
    reference): npairs and the node-pair counters exact, the sums to RTOL."""
    import capi
    O = oracle_lib
    mk = C.make_shell if coords == C.THREED else \
        (lambda n, s, cl: C.make_catalog(C.SPHERE, n, s, cl))
    cat, scale = mk(30000, 150 + d + coords, 0.3)
    cat2, _ = mk(25000, 160 + d + coords, 0.3) if cross else (None, 0)
    bn = C.binning(scale, 15, 0.003, 0.5, 0.5, 0.1 if d == 4 else 1.0)
    kw = dict(min_size=bn['min_size'], max_size=bn['max_size'], min_top=4)
    f1 = capi.Field(d, coords, cat, brute=False, **kw)
    f2 = capi.Field(d, coords, cat2, brute=False, **kw) if cross else None
    out = capi.corr2(d, f1, f2, bn['nbins'], bn['min_sep'], bn['max_sep'], bn['bin_size'],
                     bn['b'], bn['a'], arc=True)

    def ofield(c):
        return O.Field(d, coords, c['x'], c['y'], c['z'], k=c['k'], g1=c['g1'], g2=c['g2'],
                       w=c['w'], **kw)
    ref = O.corr2(d, coords, ofield(cat), ofield(cat2) if cross else None, bn['nbins'],
                  bn['min_sep'], bn['max_sep'], bn['bin_size'], bn['b'], bn['a'], exact=True,
                  arc=True)
    np.testing.assert_array_equal(out['npairs'], ref['npairs'])
    assert ref['npairs'].sum() > 1e6
    for k in ['meanr', 'meanlogr', 'weight'] + (['xi0'] if d else []):
        assert normwise(out[k], ref[k]) < RTOL, k
    for k in ('visits', 'accepted', 'leafpairs', 'refined'):
        assert out['counters'][k] == ref['counters'][k], k


@pytest.mark.gpu
def test_gpu_arc_front_end(gpu, oracle_lib):
    """NNCorrelation/GGCorrelation(metric='Arc') on 3-d positions with sep_units: the
    pair counts and xi of the public classes equal the oracle's Arc run on trees to the
    leaves, and meanr is the mean angle in the requested units."""
    O = oracle_lib
    cat, scale = C.make_shell(20000, 171, 0.3)
    tcat = gpu.Catalog(x=cat['x'], y=cat['y'], z=cat['z'], g1=cat['g1'], g2=cat['g2'],
                       w=cat['w'])
    deg = np.pi / 180.
    gg = gpu.GGCorrelation(nbins=10, min_sep=scale * 0.01 / deg, max_sep=scale * 0.4 / deg,
                           sep_units='deg', metric='Arc', min_top=4)
    gg.process(tcat)
    f = O.Field(4, C.THREED, cat['x'], cat['y'], cat['z'], g1=cat['g1'], g2=cat['g2'],
                w=cat['w'], min_size=0., max_size=0., min_top=4)
    ref = O.corr2(4, C.THREED, f, None, gg.nbins, gg._min_sep, gg._max_sep, gg._bin_size,
                  gg.b, gg.angle_slop, exact=True, arc=True)
    np.testing.assert_array_equal(gg.npairs, ref['npairs'])
    assert normwise(gg.weight, ref['weight']) < RTOL
    m = ref['weight'] > 0
    np.testing.assert_allclose(gg.meanr[m], ref['meanr'][m] / ref['weight'][m] / deg,
                               rtol=1e-12)
    with pytest.raises(ValueError):
        fcat = gpu.Catalog(x=cat['x'], y=cat['y'])
        gpu.NNCorrelation(nbins=10, min_sep=1., max_sep=10., metric='Arc').process(fcat)


@pytest.mark.gpu
@pytest.mark.parametrize('m3,d,cross,rpar', [(2, 0, False, None), (2, 4, True, (-20., 30.)),
                                             (5, 0, True, (-20., 30.)), (5, 4, False, None),
                                             (3, 4, True, None), (3, 0, True, (-20., 30.)),
                                             (0, 4, False, (-1., 2.))])
def test_gpu_los_metrics_match_oracle(gpu, oracle_lib, m3, d, cross, rpar):
    """Rperp (2), Rlens (3), OldRperp (5) and the r_par range (0: with the Euclidean
    metric) on the GPU against the oracle (itself pinned to the compiled reference):
    npairs and the node-pair counters exact, the sums to RTOL."""
    import capi
    O = oracle_lib
    cat, _ = C.make_shell(30000, 220 + m3 + d, 0.3)
    cat2, _ = C.make_shell(25000, 230 + m3 + d, 0.3) if cross else (None, 0)
    bn = C.binning(10., 15, 0.005, 0.5, 0.5, 0.1 if d == 4 else 1.0)
    kw = dict(min_size=bn['min_size'], max_size=bn['max_size'], min_top=4)
    f1 = capi.Field(d, 2, cat, brute=False, **kw)
    f2 = capi.Field(d, 2, cat2, brute=False, **kw) if cross else None
    out = capi.corr2(d, f1, f2, bn['nbins'], bn['min_sep'], bn['max_sep'], bn['bin_size'],
                     bn['b'], bn['a'], metric3=m3, rpar=rpar)

    def ofield(c):
        return O.Field(d, 2, c['x'], c['y'], c['z'], k=c['k'], g1=c['g1'], g2=c['g2'],
                       w=c['w'], **kw)
    ref = O.corr2(d, 2, ofield(cat), ofield(cat2) if cross else None, bn['nbins'],
                  bn['min_sep'], bn['max_sep'], bn['bin_size'], bn['b'], bn['a'], exact=True,
                  metric3=m3, rpar=rpar)
    np.testing.assert_array_equal(out['npairs'], ref['npairs'])
    assert ref['npairs'].sum() > 1e5
    for k in ['meanr', 'meanlogr', 'weight'] + (['xi0'] if d else []):
        assert normwise(out[k], ref[k]) < RTOL, k
    for k in ('visits', 'accepted', 'leafpairs', 'refined'):
        assert out['counters'][k] == ref['counters'][k], k


@pytest.mark.gpu
def test_gpu_rperp_front_end(gpu, oracle_lib):
    """NNCorrelation(metric='Rperp', min_rpar, max_rpar) and NNCorrelation(metric='Rlens')
    through the public classes: pair counts equal the oracle's on trees to the leaves;
    the metrics' argument checks."""
    O = oracle_lib
    cat, _ = C.make_shell(20000, 241, 0.3)
    cat2, _ = C.make_shell(15000, 242, 0.3)
    t1 = gpu.Catalog(x=cat['x'], y=cat['y'], z=cat['z'])
    t2 = gpu.Catalog(x=cat2['x'], y=cat2['y'], z=cat2['z'])
    f1 = O.Field(0, 2, cat['x'], cat['y'], cat['z'], min_size=0., max_size=0., min_top=4)
    f2 = O.Field(0, 2, cat2['x'], cat2['y'], cat2['z'], min_size=0., max_size=0., min_top=4)
    for metric, m3, rpar, c2, o2 in (('Rperp', 2, (-10., 20.), None, None),
                                     ('OldRperp', 5, None, t2, f2), ('Rlens', 3, None, t2, f2)):
        kw = dict(min_rpar=rpar[0], max_rpar=rpar[1]) if rpar else {}
        nn = gpu.NNCorrelation(nbins=10, min_sep=0.05, max_sep=5., metric=metric, min_top=4,
                               **kw)
        nn.process(t1, c2)
        ref = O.corr2(0, 2, f1, o2, nn.nbins, nn._min_sep, nn._max_sep, nn._bin_size, nn.b,
                      nn.angle_slop, exact=True, metric3=m3, rpar=rpar)
        np.testing.assert_array_equal(nn.npairs, ref['npairs'])
        assert ref['npairs'].sum() > 1e5
    with pytest.raises(ValueError):
        gpu.NNCorrelation(nbins=10, min_sep=0.05, max_sep=5., metric='Rlens').process(t1)
    flat = gpu.Catalog(x=cat['x'], y=cat['y'])
    with pytest.raises(ValueError):
        gpu.NNCorrelation(nbins=10, min_sep=0.05, max_sep=5., metric='Rperp').process(flat)
    with pytest.raises(ValueError):
        gpu.NNCorrelation(nbins=10, min_sep=0.05, max_sep=5., min_rpar=0.).process(flat)


@pytest.mark.gpu
@pytest.mark.parametrize('mode,levels,min_jobs', [(0, 0, 0), (1, 0, 0), (1, 0, 1), (2, 0, 1),
                                                   (1, 1, 1), (2, 1, 1)])
@pytest.mark.parametrize('cross', [False, True])
def test_gpu_traversal_modes_same_result(gpu, oracle_lib, mode, levels, min_jobs, cross):
    """Chunked launches, the persistent launch with jobs by a shared counter or by a
    static equal-count partition (from the initial node pairs, min_jobs 1, or after the
    chunked split into 65536 jobs), and the persistent launch with a stack too small (the
    fallback to chunked launches): npairs and the visit counters equal the oracle's."""
    from treecorr_amd import _lib
    O = oracle_lib
    L = _lib.lib()
    cat, scale = C.make_catalog(C.SPHERE, 12000, 98, 0.3)
    cat2, _ = C.make_catalog(C.SPHERE, 9000, 99, 0.3) if cross else (None, None)
    bn = C.binning(scale, 15, 0.005, 0.3, 0.1, 0.1)
    _lib.check(L.tc_set_traversal_mode(mode))
    _lib.check(L.tc_set_persist_params(levels, min_jobs))
    fb0 = L.tc_traversal_fallbacks()
    try:
        out = raw_corr2(gpu, 4, 3, cat, cat2, bn, min_top=3)
        assert L.tc_set_traversal_mode(-1) == mode
    finally:
        _lib.check(L.tc_set_traversal_mode(1))
        _lib.check(L.tc_set_persist_params(0, 0))
    assert L.tc_traversal_fallbacks() - fb0 == (1 if levels and mode else 0)
    kw = dict(min_size=bn['min_size'], max_size=bn['max_size'], min_top=3)
    f = O.Field(4, 3, cat['x'], cat['y'], cat['z'], g1=cat['g1'], g2=cat['g2'], w=cat['w'], **kw)
    f2 = None if cat2 is None else O.Field(4, 3, cat2['x'], cat2['y'], cat2['z'], g1=cat2['g1'],
                                           g2=cat2['g2'], w=cat2['w'], **kw)
    ref = O.corr2(4, 3, f, f2, bn['nbins'], bn['min_sep'], bn['max_sep'], bn['bin_size'],
                  bn['b'], bn['a'], exact=True)
    np.testing.assert_array_equal(out['npairs'], ref['npairs'])
    for k in KEYS[:-1]:
        assert normwise(out[k], ref[k]) < RTOL, k
    for k in ('visits', 'accepted', 'leafpairs', 'refined'):
        assert out['counters'][k] == ref['counters'][k], k


@pytest.mark.gpu
def test_gpu_graph_mode_replays_match_eager(gpu, oracle_lib):
    """Graph mode: the persistent launch captured once and replayed; every replay, with
    eager launches between them, gives the eager launch's npairs and counters, and the
    captures and replays are counted."""
    import capi
    from treecorr_amd import _lib
    L = _lib.lib()
    cat, scale = C.make_catalog(C.SPHERE, 20000, 131, 0.3)
    bn = C.binning(scale, 12, 0.005, 0.3, 0.1, 0.1)
    kw = dict(min_size=bn['min_size'], max_size=bn['max_size'], brute=False, min_top=3)
    f1 = capi.Field(4, 3, cat, **kw)
    args = (bn['nbins'], bn['min_sep'], bn['max_sep'], bn['bin_size'], bn['b'], bn['a'])
    _lib.check(L.tc_set_persist_params(0, 2000))
    try:
        eager = capi.corr2(4, f1, None, *args)
        _lib.check(L.tc_set_graph_mode(1))
        assert L.tc_set_graph_mode(-1) == 1
        c0, r0 = ctypes.c_longlong(), ctypes.c_longlong()
        _lib.check(L.tc_graph_stats(ctypes.byref(c0), ctypes.byref(r0)))
        outs = []
        for rep in range(3):
            outs.append(capi.corr2(4, f1, None, *args))
            _lib.check(L.tc_set_graph_mode(0))
            between = capi.corr2(4, f1, None, *args)      # eager work between replays
            np.testing.assert_array_equal(between['npairs'], eager['npairs'])
            _lib.check(L.tc_set_graph_mode(1))
        c1, r1 = ctypes.c_longlong(), ctypes.c_longlong()
        _lib.check(L.tc_graph_stats(ctypes.byref(c1), ctypes.byref(r1)))
    finally:
        _lib.check(L.tc_set_graph_mode(0))
        _lib.check(L.tc_set_persist_params(0, 0))
    assert c1.value - c0.value >= 1 and r1.value - r0.value >= 1
    assert (c1.value - c0.value) + (r1.value - r0.value) == 3
    for out in outs:
        np.testing.assert_array_equal(out['npairs'], eager['npairs'])
        assert out['counters'] == eager['counters']
        for k in KEYS[:-1]:
            assert normwise(out[k], eager[k]) < 1e-13, k


@pytest.mark.gpu
def test_gpu_autotune_keeps_results(gpu):
    import capi
    from treecorr_amd import _lib, tune
    cat, scale = C.make_catalog(C.SPHERE, 15000, 132, 0.3)
    bn = C.binning(scale, 12, 0.005, 0.3, 0.1, 0.1)
    kw = dict(min_size=bn['min_size'], max_size=bn['max_size'], brute=False, min_top=3)
    f1 = capi.Field(4, 3, cat, **kw)
    args = (bn['nbins'], bn['min_sep'], bn['max_sep'], bn['bin_size'], bn['b'], bn['a'])
    ref = capi.corr2(4, f1, None, *args)
    outs = []
    try:
        best, times = tune.autotune(lambda: outs.append(capi.corr2(4, f1, None, *args)),
                                    min_jobs=(500, 5000), grab_div=(2, 32))
    finally:
        tune.set_params()
    assert best in times and len(times) == 4 and all(t > 0 for t in times.values())
    for out in outs:
        np.testing.assert_array_equal(out['npairs'], ref['npairs'])
        assert out['counters'] == ref['counters']
