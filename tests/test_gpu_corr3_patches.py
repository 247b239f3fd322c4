"""3-point correlations with patches: the multipole algorithm's local rule.

Each patch i contributes the triangles whose first vertex is in the patch and whose other
two vertices are among the objects of the (other) catalogs near it, stored as
results[(i, i, i)] (corr3base.py:1287-1351, 1388-1421, 1533-1563, 1760-1818).  With
brute=True every object is its own leaf, so the nearby objects give exactly the triangles
of the whole catalog: each results[(i, i, i)] must equal the oracle's processCross12
(ordered=1) of patch i against the whole catalog, and the totals the whole catalog's
auto-correlation (ntri exactly, the FP64 sums to 1e-12 normwise).  Unordered cross
correlations, which the local rule runs as several ordered passes, must give the
unpatched result.  Each test prints its case before the device calls."""
import numpy as np
import pytest

import cases as C

RTOL = 1e-12
FLAT, SPHERE = 1, 3


def normwise(a, b):
    den = np.max(np.abs(b))
    return 0. if den == 0 else np.max(np.abs(a - b)) / den


def labels_for(cat, coords, scale, ng):
    if coords == SPHERE:
        ra = np.arctan2(cat['y'], cat['x'])
        dec = np.arcsin(cat['z'])
        u, v = ra / scale, (dec + scale / 2) / scale
    else:
        u, v = cat['x'] / scale, cat['y'] / scale
    iu = np.clip((u * ng).astype(int), 0, ng - 1)
    iv = np.clip((v * ng).astype(int), 0, ng - 1)
    return iu * ng + iv


def public_catalog(tc, cat, coords, labels=None, d=0):
    kw = dict(w=cat['w'])
    if coords == SPHERE:
        ra = np.arctan2(cat['y'], cat['x'])
        dec = np.arcsin(cat['z'])
        kw.update(ra=ra, dec=dec, ra_units='rad', dec_units='rad')
    else:
        kw.update(x=cat['x'], y=cat['y'])
    if d == 4:
        kw.update(g1=cat['g1'], g2=cat['g2'])
    if labels is not None:
        kw['patch'] = labels
    return tc.Catalog(**kw)


def multipole(tc, d, scale, **kw):
    cls = tc.GGGCorrelation if d == 4 else tc.NNNCorrelation
    return cls(min_sep=scale * 0.02, max_sep=scale * 0.2, nbins=6, max_n=8,
               bin_type='LogMultipole', brute=True, **kw)


def raw(corr):
    out = dict(ntri=corr.ntri, weight=corr.weightr, weight_im=corr.weighti,
               meand2=corr.meand2, meand3=corr.meand3)
    if corr._d == 4:
        for i in range(8):
            out['zeta%d' % i] = corr._z[i]
    return out


def sub(cat, sel):
    return {k: (None if v is None else np.ascontiguousarray(v[sel])) for k, v in cat.items()}


@pytest.mark.gpu
@pytest.mark.parametrize('d,coords', [(0, FLAT), (4, FLAT), (4, SPHERE)])
def test_gpu_local_patches_auto(gpu, oracle_lib, d, coords):
    O = oracle_lib
    ng = 2
    cat, scale = C.make_catalog(coords, 1500, 71 + d + coords, 0.2)
    lab = labels_for(cat, coords, scale, ng)
    print('3pt local patches auto d', d, 'coords', coords, flush=True)
    corr = multipole(gpu, d, scale)
    pc = public_catalog(gpu, cat, coords, lab, d)
    if coords == SPHERE:
        # the oracle sees the positions the catalog made from ra, dec
        pc._generate_xyz()
        cat = dict(cat, x=pc._x, y=pc._y, z=pc._z)
    corr.process(pc, finalize=False)
    assert sorted(corr.results) == [(i, i, i) for i in range(ng * ng)]
    assert corr.npatch1 == corr.npatch2 == corr.npatch3 == ng * ng
    # brute fields have min_size 0 (field.py:569): every object its own leaf
    kw = dict(min_size=0., max_size=corr._get_minmax_size()[1], brute=True)

    def fld(c):
        return O.Field(d, coords, c['x'], c['y'], c['z'], g1=c['g1'], g2=c['g2'], w=c['w'],
                       **kw)
    whole = fld(cat)
    args = (corr.nbins, corr._min_sep, corr._max_sep, corr._bin_size, corr.b,
            corr.angle_slop, corr.max_n)
    keys = ['weight', 'weight_im', 'meand2', 'meand3'] + \
        (['zeta%d' % i for i in range(8)] if d == 4 else [])
    for i in range(ng * ng):
        ref = O.corr3_multipole_cross(d, coords, 'cross12', 1, [fld(sub(cat, lab == i)), whole],
                                      *args)
        got = raw(corr.results[(i, i, i)])
        assert np.sum(ref['ntri']) > 0
        np.testing.assert_array_equal(got['ntri'], ref['ntri'], err_msg='patch %d' % i)
        for k in keys:
            assert normwise(got[k], ref[k]) < RTOL, (i, k)
    # the totals are the sums of the patches, and the whole catalog's auto-correlation
    tot = sum(corr.results[(i, i, i)].ntri for i in range(ng * ng))
    np.testing.assert_array_equal(corr.ntri, tot)
    auto = O.corr3_multipole(d, coords, whole, *args)
    np.testing.assert_array_equal(corr.ntri, auto['ntri'])
    for k in keys:
        assert normwise(raw(corr)[k], auto[k]) < RTOL, k
    if d == 0:
        # each patch adds 1/2 w_i W^2 (its expanded patch carries the whole weight W)
        W = np.sum(cat['w'])
        assert corr.tot == pytest.approx(0.5 * W ** 3, rel=1e-12)


@pytest.mark.gpu
def test_gpu_local_patches_sas_results(gpu):
    """LogSAS through toSAS keeps each patch's result as raw sums (corr3base.py:
    2626-2640, 2665-2675): their weights add up to the total weight, their zetas to the
    total zeta times the weight, and the finalized object equals the unpatched one."""
    cat, scale = C.make_catalog(FLAT, 1200, 91, 0.2)
    lab = labels_for(cat, FLAT, scale, 2)
    kw = dict(min_sep=scale * 0.02, max_sep=scale * 0.2, nbins=5, nphi_bins=8, max_n=8,
              brute=True)
    print('3pt local patches sas', flush=True)
    gp = gpu.GGGCorrelation(**kw)
    gp.process(public_catalog(gpu, cat, FLAT, lab, 4))
    g1 = gpu.GGGCorrelation(**kw)
    g1.process(public_catalog(gpu, cat, FLAT, None, 4))
    assert len(gp.results) == 4
    np.testing.assert_allclose(sum(r.weightr for r in gp.results.values()), gp.weightr,
                               rtol=1e-11, atol=1e-11 * np.max(np.abs(gp.weightr)))
    for name in ('gam0', 'gam1', 'gam2', 'gam3'):
        tot = sum(getattr(r, name) for r in gp.results.values())
        want = getattr(gp, name) * gp.weightr
        assert normwise(tot, want) < 1e-11, name
        assert normwise(getattr(gp, name), getattr(g1, name)) < 1e-11, name
    np.testing.assert_allclose(gp.ntri, g1.ntri, rtol=1e-11, atol=1e-9)
    assert gp._var_num == pytest.approx(g1._var_num, rel=1e-12)


@pytest.mark.gpu
@pytest.mark.parametrize('kind,ordered', [('cross12', False), ('cross12', True),
                                          ('cross', False), ('cross', 3), ('cross', True)])
def test_gpu_local_patches_cross(gpu, kind, ordered):
    """Patched cross-correlations (the unordered ones as extra ordered passes with another
    catalog first) against the same correlation without patches."""
    cats = [C.make_catalog(FLAT, 900, 120 + i, 0.2)[0] for i in range(3)]
    scale = 100.
    nf = 2 if kind == 'cross12' else 3
    print('3pt local patches', kind, ordered, flush=True)
    patched = [public_catalog(gpu, c, FLAT, labels_for(c, FLAT, scale, 2)) for c in cats[:nf]]
    plain = [public_catalog(gpu, c, FLAT) for c in cats[:nf]]
    cp = multipole(gpu, 0, scale)
    cp.process(*patched, ordered=ordered, finalize=False)
    c1 = multipole(gpu, 0, scale)
    c1.process(*plain, ordered=ordered, finalize=False)
    assert np.sum(c1.ntri) > 0
    np.testing.assert_array_equal(cp.ntri, c1.ntri)
    for k in ('weightr', 'weighti', 'meand2', 'meand3'):
        assert normwise(getattr(cp, k), getattr(c1, k)) < RTOL, k
    assert sorted(cp.results) == [(i, i, i) for i in range(4)]


def _ggg_stat(r):
    return np.concatenate([(r['zeta%d' % (2 * i)] + 1j * r['zeta%d' % (2 * i + 1)]).ravel()
                           for i in range(4)])


def _gram(v, w=None):
    vc = v - np.mean(v, axis=0)
    return (vc.conj().T * (1. if w is None else w)).dot(vc)


@pytest.mark.gpu
def test_gpu_3pt_covariances_from_oracle_rows(gpu, oracle_lib):
    """Covariances of a LogMultipole GGG with patches (corr3base.py:2678-2790, the triple
    lists :2997-3330).  With brute=True, jackknife row m ('simple') holds the triangles
    whose first vertex is outside patch m: the oracle's processCross12 (ordered=1) of the
    catalog without patch m against the whole catalog; a 'sample' row is patch m's own
    (the oracle's cross12 of the patch against the catalog).  The estimates must be the
    textbook sums over those independently computed rows (corr2base.py:2423-2538); the
    marked and plain bootstrap rows of local results are the same draws' sums, so equal
    for one seed; shot is var_num / |weight|."""
    O = oracle_lib
    ng, n = 2, 4
    cat, scale = C.make_catalog(FLAT, 1400, 301, 0.2)
    lab = labels_for(cat, FLAT, scale, ng)
    print('3pt covariances: jackknife, sample, bootstrap', flush=True)
    g = gpu.GGGCorrelation(min_sep=scale * 0.02, max_sep=scale * 0.2, nbins=4, max_n=4,
                           bin_type='LogMultipole', brute=True, rng=np.random.default_rng(8))
    g.process(public_catalog(gpu, cat, FLAT, lab, 4))
    kw = dict(min_size=0., max_size=g._get_minmax_size()[1], brute=True)

    def fld(c):
        return O.Field(4, FLAT, c['x'], c['y'], c['z'], g1=c['g1'], g2=c['g2'], w=c['w'], **kw)
    whole = fld(cat)
    args = (g.nbins, g._min_sep, g._max_sep, g._bin_size, g.b, g.angle_slop, g.max_n)
    jk, own, wown = [], [], []
    for m in range(n):
        r = O.corr3_multipole_cross(4, FLAT, 'cross12', 1, [fld(sub(cat, lab != m)), whole],
                                    *args)
        jk.append(_ggg_stat(r))
        r = O.corr3_multipole_cross(4, FLAT, 'cross12', 1, [fld(sub(cat, lab == m)), whole],
                                    *args)
        own.append(_ggg_stat(r))
        wown.append(4 * np.sum(np.abs(r['weight'] + 1j * r['weight_im'])))
    jk, own, wown = np.array(jk), np.array(own), np.array(wown)
    want = (1 - 1 / n) * _gram(jk)
    got = g.estimate_cov('jackknife', cross_patch_weight='simple')
    assert normwise(got, want) < 1e-9
    w = wown / np.sum(wown)
    want = _gram(own, w) / (n - 1)
    got = g.estimate_cov('sample')
    assert normwise(got, want) < 1e-9
    # the diagonal is real and the matrix Hermitian
    np.testing.assert_allclose(got, got.conj().T, rtol=0, atol=1e-12 * np.max(np.abs(got)))
    # bootstrap draws: rng.choice(n, n) per row; local results make both kinds equal
    g._rng = np.random.default_rng(21)
    cb = g.estimate_cov('bootstrap', num_bootstrap=40)
    draws = np.random.default_rng(21)
    rows = []
    for _ in range(40):
        idx = draws.choice(np.arange(n), size=n, replace=True)
        rows.append(np.sum([own[i] for i in idx], axis=0))
    want = _gram(np.array(rows)) / 39
    assert normwise(cb, want) < 1e-9
    g._rng = np.random.default_rng(21)
    cm = g.estimate_cov('marked_bootstrap', num_bootstrap=40)
    assert normwise(cm, cb) < 1e-12
    shot = g.estimate_cov('shot')
    wt = g.getWeight()
    np.testing.assert_allclose(shot[wt != 0], g._var_num / wt[wt != 0], rtol=1e-14)
    np.testing.assert_allclose(g.vargam0.ravel(), shot[:wt.size // 4].real, rtol=1e-14)


@pytest.mark.gpu
def test_gpu_3pt_logsas_jackknife_rows(gpu):
    """LogSAS GGG with patches: jackknife rows built on the GPU (R = W S, then the Gram
    product) against numpy sums of the stored patch results, normalised by their summed
    weight (corr3base.py:2312-2320, 2451-2481); with cross_patch_weight 'mean' a row keeps
    every triple with weight (number of its patches other than m) / 3."""
    cat, scale = C.make_catalog(FLAT, 1200, 302, 0.2)
    lab = labels_for(cat, FLAT, scale, 2)
    print('3pt logsas jackknife', flush=True)
    g = gpu.GGGCorrelation(min_sep=scale * 0.02, max_sep=scale * 0.2, nbins=4, nphi_bins=6,
                           max_n=6, brute=True)
    g.process(public_catalog(gpu, cat, FLAT, lab, 4))
    res = [g.results[(i, i, i)] for i in range(4)]
    for cpw in ('simple', 'mean'):
        rows = []
        for m in range(4):
            keep = [r for i, r in enumerate(res) if i != m]
            W = np.sum([r.weightr for r in keep], axis=0)
            mask = W != 0
            stat = []
            for name in ('gam0', 'gam1', 'gam2', 'gam3'):
                z = np.sum([getattr(r, name) for r in keep], axis=0)
                z[mask] /= W[mask]
                stat.append(z.ravel())
            rows.append(np.concatenate(stat))
        want = 0.75 * _gram(np.array(rows))
        got = g.estimate_cov('jackknife', cross_patch_weight=cpw)
        assert normwise(got, want) < 1e-9, cpw


@pytest.mark.gpu
def test_gpu_nnn_calculate_zeta(gpu):
    """NNN's estimator (nnncorrelation.py:228-460).  Randoms that are the data themselves
    give zeta = 0 in every bin with triangles, for (DDD-RRR)/RRR and for
    (DDD-RDD+DRR-RRR)/RRR: the unordered DRR and RDD hold each triangle of DDD three
    times, and the ratios of the totals (W^3/6 for an auto-correlation, W^3/2 for a
    cross12) must cancel that.
    With patches the estimate equals the one without, and the jackknife rows (the
    randoms summed over the same triples) give a positive variance."""
    cat, scale = C.make_catalog(FLAT, 900, 401, 0.3)
    rnd = C.make_catalog(FLAT, 1500, 402, 0.0)[0]
    kw = dict(min_sep=scale * 0.03, max_sep=scale * 0.25, nbins=4, nphi_bins=6, max_n=6,
              brute=True)
    print('nnn calculateZeta', flush=True)

    def nnn(*cats):
        # DRR and RDD unordered, as the estimator expects (a random in any vertex)
        c = gpu.NNNCorrelation(**kw)
        c.process(*cats, ordered=len(cats) == 1)
        return c
    d = public_catalog(gpu, cat, FLAT)
    ddd = nnn(d)
    same = nnn(d)
    drr, rdd = nnn(d, d), nnn(d, d)
    z, vz = ddd.calculateZeta(rrr=same)
    mask = same.weight != 0
    assert np.sum(mask) > 20
    assert np.max(np.abs(z[mask])) < 1e-12
    z, vz = ddd.calculateZeta(rrr=same, drr=drr, rdd=rdd)
    assert np.max(np.abs(z[mask])) < 1e-12
    # with randoms, patched against plain
    r = public_catalog(gpu, rnd, FLAT)
    ld, lr = labels_for(cat, FLAT, scale, 2), labels_for(rnd, FLAT, scale, 2)
    dp, rp = public_catalog(gpu, cat, FLAT, ld), public_catalog(gpu, rnd, FLAT, lr)
    z0, v0 = nnn(d).calculateZeta(rrr=nnn(r), drr=nnn(d, r), rdd=nnn(r, d))
    dddp = nnn(dp)
    z1, v1 = dddp.calculateZeta(rrr=nnn(rp), drr=nnn(dp, rp), rdd=nnn(rp, dp))
    assert normwise(z1, z0) < 1e-10
    np.testing.assert_allclose(v1, v0, rtol=1e-10)
    cov = dddp.estimate_cov('jackknife')
    assert np.all(np.isfinite(cov)) and np.all(np.diag(cov)[dddp.getWeight() != 0] >= 0)
    assert np.sum(np.diag(cov) > 0) > 20


@pytest.mark.gpu
def test_gpu_3pt_write_patch_results_round_trip(gpu, tmp_path):
    """write(write_patch_results=True) then Corr3.from_file: the object read back gives the
    same jackknife covariance as the one processed (corr3base.py:3403-3570)."""
    cat, scale = C.make_catalog(FLAT, 1000, 303, 0.2)
    lab = labels_for(cat, FLAT, scale, 2)
    print('3pt write_patch_results round trip', flush=True)
    g = gpu.GGGCorrelation(min_sep=scale * 0.02, max_sep=scale * 0.2, nbins=4, nphi_bins=6,
                           max_n=6, brute=True)
    g.process(public_catalog(gpu, cat, FLAT, lab, 4))
    f = str(tmp_path / 'ggg.out')
    g.write(f, write_patch_results=True)
    g2 = gpu.Corr3.from_file(f)
    c1 = g.estimate_cov('jackknife')
    assert normwise(g2.estimate_cov('jackknife'), c1) < 1e-10
