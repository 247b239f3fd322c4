"""GPU multipole (NNN, GGG) vs the literal reference's golden vectors and the oracle.

Each test prints the case it runs (name, type, coordinates, size) before the device
call, so that a failure that ends the process still names its case."""
import json
import os

import numpy as np
import pytest

import cases as C

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')
RTOL = 1e-12   # normwise relative tolerance for the FP64 sums; ntri is exact


def normwise(a, b):
    den = np.max(np.abs(b))
    return 0. if den == 0 else np.max(np.abs(a - b)) / den


@pytest.mark.gpu
@pytest.mark.parametrize('case', C.CORR3_CASES, ids=[c[0] for c in C.CORR3_CASES])
def test_gpu_multipole_matches_reference_golden(gpu, case):
    import capi
    z = np.load(os.path.join(GOLD, case[0] + '.npz'), allow_pickle=False)
    spec = json.loads(str(z['spec']))
    gold = {k: z[k] for k in z.files if k != 'spec'}
    cat, bn, extra = C.corr3_case(case)
    assert C.checksum(cat) == spec['checksum']
    d, coords = case[1], case[2]
    print('multipole golden case', case[0], 'd', d, 'coords', coords, 'n', len(cat['x']), flush=True)
    f = capi.Field(d, coords, cat, min_size=bn['min_size'], max_size=bn['max_size'],
                   min_top=extra['min_top'])
    assert f.ntop == spec['ntop']
    out = capi.corr3(d, f, bn['nbins'], bn['min_sep'], bn['max_sep'], bn['bin_size'], bn['b'],
                     bn['a'], bn['max_n'])
    np.testing.assert_array_equal(out['ntri'], gold['ntri'])
    for k, v in gold.items():
        if k != 'ntri':
            assert normwise(out[k], v) < RTOL, k


@pytest.mark.gpu
@pytest.mark.parametrize('d,coords,n,max_n', [(4, 3, 40000, 30), (0, 1, 40000, 20),
                                              (4, 1, 30000, 16)])
def test_gpu_multipole_matches_oracle(gpu, oracle_lib, d, coords, n, max_n):
    import capi
    O = oracle_lib
    cat, scale = C.make_catalog(coords, n, 60 + d + coords, 0.3)
    bn = capi.multipole_binning(scale, 12, 0.01, 0.2, max_n)
    print('multipole oracle case d', d, 'coords', coords, 'n', n, 'max_n', max_n, flush=True)
    kw = dict(min_size=bn['min_size'], max_size=bn['max_size'], min_top=4)
    f = capi.Field(d, coords, cat, **kw)
    out = capi.corr3(d, f, bn['nbins'], bn['min_sep'], bn['max_sep'], bn['bin_size'], bn['b'],
                     bn['a'], bn['max_n'])
    of = O.Field(d, coords, cat['x'], cat['y'], cat['z'], g1=cat['g1'], g2=cat['g2'], w=cat['w'],
                 **kw)
    ref = O.corr3_multipole(d, coords, of, bn['nbins'], bn['min_sep'], bn['max_sep'],
                            bn['bin_size'], bn['b'], bn['a'], bn['max_n'])
    np.testing.assert_array_equal(out['ntri'], ref['ntri'])
    keys = ['weight', 'weight_im', 'meand2', 'meanlogd2', 'meand3', 'meanlogd3']
    if d == 4:
        keys += ['zeta%d' % i for i in range(8)]
    for k in keys:
        assert normwise(out[k], ref[k]) < RTOL, k
    for k in ('gn_calls', 'zeta_calls', 'zeta_rows', 'stages'):
        assert out['counters'][k] == ref['counters'][k], k


@pytest.mark.gpu
def test_gpu_ggg_public_api(gpu):
    """GGGCorrelation(bin_type='LogMultipole').process(cat): finalized arrays."""
    rng = np.random.default_rng(12)
    n = 20000
    cat = gpu.Catalog(x=rng.uniform(0, 300, n), y=rng.uniform(0, 300, n), x_units='arcmin',
                      y_units='arcmin', g1=rng.normal(0, 0.2, n), g2=rng.normal(0, 0.2, n),
                      w=rng.uniform(0.5, 1.5, n))
    ggg = gpu.GGGCorrelation(min_sep=5., max_sep=100., nbins=10, max_n=10, sep_units='arcmin',
                             bin_type='LogMultipole', min_top=4)
    ggg.process(cat)
    assert ggg.gam0.shape == (10, 10, 21)
    # ntri/meand broadcast from the n = max_n slice (corr3base.py:2386-2396)
    assert np.all(ggg.ntri[:, :, 0] == ggg.ntri[:, :, 10])
    good = ggg.weightr[:, :, 10] > 0
    assert np.all(ggg.meand2[:, :, 10][good] > 4.) and np.all(ggg.meand2[:, :, 10][good] < 110.)
    # W_{-n} = conj(W_n) symmetry of the weight
    np.testing.assert_allclose(ggg.weightr[:, :, ::-1], ggg.weightr.transpose(1, 0, 2), rtol=1e-12,
                               atol=1e-9 * np.max(np.abs(ggg.weightr)))


@pytest.mark.gpu
@pytest.mark.parametrize('mode', [1, 2], ids=['fp32_mfma', 'bf16x3_mfma'])
@pytest.mark.parametrize('d,coords,n,max_n', [(4, 1, 30000, 16), (0, 1, 30000, 12),
                                              (4, 3, 20000, 10)])
def test_gpu_multipole_zeta_on_matrix_cores(gpu, mode, d, coords, n, max_n):
    """The zeta contraction on the matrix cores (FP32, or BF16 split in three) against
    the FP64 contraction of the same run: for every multipole n, each complex output
    (weight + i weight_im, and gam0..3 = zeta0 + i zeta1, ...) within 1e-6 of its largest
    modulus over (k2, k3) (normwise); ntri and the mean distances, which stay FP64,
    unchanged to 1e-12."""
    import capi
    from treecorr_amd import _lib
    L = _lib.lib()
    cat, scale = C.make_catalog(coords, n, 90 + d + coords, 0.3)
    bn = capi.multipole_binning(scale, 12, 0.01, 0.2, max_n)
    kw = dict(min_size=bn['min_size'], max_size=bn['max_size'], min_top=4)
    print('zeta mode', mode, 'd', d, 'coords', coords, 'n', n, 'max_n', max_n, flush=True)
    f = capi.Field(d, coords, cat, **kw)
    args = (d, f, bn['nbins'], bn['min_sep'], bn['max_sep'], bn['bin_size'], bn['b'], bn['a'],
            bn['max_n'])
    ref = capi.corr3(*args)
    assert L.tc_set_zeta_mode(mode) == 0
    try:
        got = capi.corr3(*args)
    finally:
        L.tc_set_zeta_mode(0)
    np.testing.assert_array_equal(got['ntri'], ref['ntri'])
    for k in ('meand2', 'meanlogd2', 'meand3', 'meanlogd3'):
        assert normwise(got[k], ref[k]) < RTOL, k
    pairs = [('weight', 'weight_im')] + ([('zeta%d' % (2 * i), 'zeta%d' % (2 * i + 1))
                                           for i in range(4)] if d == 4 else [])
    for kr, ki in pairs:
        zr = ref[kr] + 1j * ref[ki]
        zg = got[kr] + 1j * got[ki]
        for i in range(2 * max_n + 1):
            den = np.max(np.abs(zr[:, :, i]))
            if den == 0:
                continue
            err = np.max(np.abs(zg[:, :, i] - zr[:, :, i])) / den
            assert err < 1e-6, (kr, i - max_n, err)


CROSS_KINDS = [('cross12', 0), ('cross12', 1), ('cross21', 0), ('cross21', 3), ('cross', 0),
               ('cross', 1), ('cross', 2), ('cross', 3), ('cross', 4)]


@pytest.mark.gpu
@pytest.mark.parametrize('case', C.CORR3_CROSS_CASES, ids=[c[0] for c in C.CORR3_CROSS_CASES])
def test_gpu_multipole_cross_matches_oracle(gpu, oracle_lib, case):
    """Multipole cross-correlations through the C-ABI (processCross12 0/1, processCross21
    0/3, processCross 0-4) against the oracle's passes on the same three catalogs: ntri
    exact, the other sums within 1e-12 normwise, the work counters equal."""
    import capi
    O = oracle_lib
    d, coords = case[1], case[2]
    cats, bn, extra = C.corr3_cross_case(case, 3)
    kw = dict(min_size=bn['min_size'], max_size=bn['max_size'], min_top=extra['min_top'])
    args = (bn['nbins'], bn['min_sep'], bn['max_sep'], bn['bin_size'], bn['b'], bn['a'],
            bn['max_n'])
    gf = [capi.Field(d, coords, c, **kw) for c in cats]
    of = [O.Field(d, coords, c['x'], c['y'], c['z'], g1=c['g1'], g2=c['g2'], w=c['w'], **kw)
          for c in cats]
    keys = ['weight', 'weight_im', 'meand2', 'meanlogd2', 'meand3', 'meanlogd3']
    if d == 4:
        keys += ['zeta%d' % i for i in range(8)]
    for kind, ordered in CROSS_KINDS:
        nf = 3 if kind == 'cross' else 2
        print('multipole cross case', case[0], kind, ordered, flush=True)
        out = capi.corr3_cross(d, kind, ordered, gf[:nf], *args)
        ref = O.corr3_multipole_cross(d, coords, kind, ordered, of[:nf], *args)
        assert np.sum(ref['ntri']) > 0
        np.testing.assert_array_equal(out['ntri'], ref['ntri'], err_msg=f'{kind} {ordered}')
        for k in keys:
            assert normwise(out[k], ref[k]) < RTOL, (kind, ordered, k)
        for k in ('gn_calls', 'zeta_calls', 'zeta_rows', 'stages'):
            assert out['counters'][k] == ref['counters'][k], (kind, ordered, k)


@pytest.mark.gpu
@pytest.mark.parametrize('d,coords,n', [(4, 3, 20000), (0, 1, 25000)])
def test_gpu_multipole_cross_larger(gpu, oracle_lib, d, coords, n):
    """Cross-correlations of catalogs deep enough for many finish levels and scratch
    copies (cross12 ordered 0 and the full cross ordered 0, every pass of each) against
    the oracle."""
    import capi
    O = oracle_lib
    cats = [C.make_catalog(coords, n, 200 + 7 * i + d + coords, 0.3)[0] for i in range(3)]
    scale = C.make_catalog(coords, 10, 200, 0.)[1]
    bn = capi.multipole_binning(scale, 10, 0.01, 0.15, 12)
    kw = dict(min_size=bn['min_size'], max_size=bn['max_size'], min_top=4)
    args = (bn['nbins'], bn['min_sep'], bn['max_sep'], bn['bin_size'], bn['b'], bn['a'],
            bn['max_n'])
    gf = [capi.Field(d, coords, c, **kw) for c in cats]
    of = [O.Field(d, coords, c['x'], c['y'], c['z'], g1=c['g1'], g2=c['g2'], w=c['w'], **kw)
          for c in cats]
    for kind, ordered, nf in (('cross12', 0, 2), ('cross', 0, 3)):
        print('multipole cross larger d', d, 'coords', coords, 'n', n, kind, ordered, flush=True)
        out = capi.corr3_cross(d, kind, ordered, gf[:nf], *args)
        ref = O.corr3_multipole_cross(d, coords, kind, ordered, of[:nf], *args)
        np.testing.assert_array_equal(out['ntri'], ref['ntri'])
        for k in ('weight', 'weight_im', 'meand2', 'meand3') + (('zeta0', 'zeta3', 'zeta6')
                                                                if d == 4 else ()):
            assert normwise(out[k], ref[k]) < RTOL, (kind, k)
        assert out['counters'] == ref['counters']
