"""Patch-based 2-point correlations on the GPU and their covariance estimates.

Every patch pair of a patch-based process() runs in one device traversal
(tc_corr2_process_pairs).  Checked here:
  - each results[(i,j)] against the oracle run on the two patch catalogs (npairs
    exactly, the FP64 sums to 1e-12 normwise), the set of stored pairs, NN's tot;
  - with bin_slop = 0 (and angle_slop = 0 for shear, unit weights for counts: every
    pair binned exactly, so sums do not depend on the tree), each jackknife row
    against the oracle's correlation of the catalog with that patch removed, and the 'mean'/'match' rows against the combination of
    whole-catalog runs those weights imply.  The covariance is then the textbook
    jackknife sum over those independently computed rows.
"""
import math

import numpy as np
import pytest

import cases as C

RTOL = 1e-12
FLAT, THREED = 1, 2


def normwise(a, b):
    den = np.max(np.abs(b))
    return 0. if den == 0 else np.max(np.abs(a - b)) / den


def grid_labels(x, y, extent, ng):
    ix = np.minimum((x / extent * ng).astype(int), ng - 1)
    iy = np.minimum((y / extent * ng).astype(int), ng - 1)
    return ix * ng + iy


def sub(cat, sel):
    return {k: (None if v is None else np.ascontiguousarray(v[sel])) for k, v in cat.items()}


def oracle_raw(O, d, coords, c1, c2, bn, min_top):
    kw = dict(min_size=bn['min_size'], max_size=bn['max_size'], min_top=min_top)

    def fld(c):
        return O.Field(d, coords, c['x'], c['y'], c['z'], k=c.get('k'), g1=c.get('g1'),
                       g2=c.get('g2'), w=c['w'], **kw)
    f1 = fld(c1)
    f2 = fld(c2) if c2 is not None else None
    return O.corr2(d, coords, f1, f2, bn['nbins'], bn['min_sep'], bn['max_sep'],
                   bn['bin_size'], bn['b'], bn['a'], exact=True)


def front_catalog(tc, cat, coords, labels):
    kw = dict(x=cat['x'], y=cat['y'], w=cat['w'], patch=labels)
    if coords != FLAT:
        kw['z'] = cat['z']
    for k in ('k', 'g1', 'g2'):
        if cat.get(k) is not None:
            kw[k] = cat[k]
    return tc.Catalog(**kw)


def corr_obj(tc, d, bn, bin_slop, angle_slop, min_top, **kw):
    cls = {0: tc.NNCorrelation, 1: tc.KKCorrelation, 4: tc.GGCorrelation}[d]
    return cls(nbins=bn['nbins'], min_sep=bn['min_sep'], max_sep=bn['max_sep'],
               bin_slop=bin_slop, angle_slop=angle_slop, min_top=min_top, **kw)


def check_pair(res, ref, nxi):
    np.testing.assert_array_equal(res.npairs, ref['npairs'])
    for q in range(nxi):
        assert normwise(res._xi[q], ref['xi%d' % q]) < RTOL, q
    for k in ('meanr', 'meanlogr', 'weight'):
        assert normwise(getattr(res, k), ref[k]) < RTOL, k


@pytest.mark.gpu
@pytest.mark.parametrize('d,coords,bs', [(4, THREED, 0.1), (1, FLAT, 0.5), (0, THREED, 0.0)])
def test_gpu_patch_auto_pairs_match_oracle(gpu, oracle_lib, d, coords, bs):
    O = oracle_lib
    ng, min_top = 4, 3
    cat, scale = C.make_catalog(coords, 6000, 300 + d + coords, 0.2)
    labels = grid_labels(cat['x'], cat['y'], scale, ng)
    bn = C.binning(scale, 12, 0.01, 0.3, bs, 0.1 if d == 4 else 1.0)
    corr = corr_obj(gpu, d, bn, bs, bn['a'], min_top)
    corr.process(front_catalog(gpu, cat, coords, labels), finalize=False)
    npatch = ng * ng
    assert corr.npatch1 == corr.npatch2 == npatch
    nxi = {0: 0, 1: 1, 4: 4}[d]
    tot_w = np.zeros(bn['nbins'])
    tot_np = np.zeros(bn['nbins'])
    nskip = 0
    for i in range(npatch):
        for j in range(i, npatch):
            ci = sub(cat, labels == i)
            cj = sub(cat, labels == j)
            if i == j:
                ref = oracle_raw(O, d, coords, ci, None, bn, min_top)
            else:
                ref = oracle_raw(O, d, coords, ci, cj, bn, min_top)
            if (i, j) in corr.results:
                check_pair(corr.results[(i, j)], ref, nxi)
            else:
                nskip += 1
                assert i != j and not np.any(ref['weight']), (i, j)
            tot_w += ref['weight']
            tot_np += ref['npairs']
    # far patch pairs are skipped before the traversal (NN keeps them with their tot)
    if d == 0:
        assert nskip == 0 and len(corr.results) == npatch * (npatch + 1) // 2
        w = np.array([np.sum(cat['w'][labels == p]) for p in range(npatch)])
        assert math.isclose(corr.tot, 0.5 * np.sum(w) ** 2, rel_tol=1e-12)
    else:
        assert nskip > 0
    np.testing.assert_array_equal(corr.npairs, tot_np)
    assert normwise(corr.weight, tot_w) < RTOL


@pytest.mark.gpu
def test_gpu_patch_cross_pairs_match_oracle(gpu, oracle_lib):
    O = oracle_lib
    ng, min_top = 3, 3
    c1, scale = C.make_catalog(THREED, 4000, 321, 0.2)
    c2, _ = C.make_catalog(THREED, 3000, 322, 0.0)
    l1 = grid_labels(c1['x'], c1['y'], scale, ng)
    l2 = grid_labels(c2['x'], c2['y'], scale, ng)
    bn = C.binning(scale, 10, 0.01, 0.3, 0.2, 1.0)
    kk = corr_obj(gpu, 1, bn, 0.2, 1.0, min_top)
    kk.process(front_catalog(gpu, c1, THREED, l1), front_catalog(gpu, c2, THREED, l2),
               finalize=False)
    npatch = ng * ng
    for i in range(npatch):
        for j in range(npatch):
            ref = oracle_raw(O, 1, THREED, sub(c1, l1 == i), sub(c2, l2 == j), bn, min_top)
            if (i, j) in kk.results:
                check_pair(kk.results[(i, j)], ref, 1)
            else:
                assert i != j and not np.any(ref['weight'])
    for i in range(npatch):     # the i == j pairs are always stored
        assert (i, i) in kk.results


@pytest.mark.gpu
@pytest.mark.parametrize('mode,levels,min_jobs,chunk',
                         [(0, 0, 0, 3000), (1, 0, 2, 0), (1, 0, 2000, 3000), (2, 0, 2, 0),
                          (1, 1, 2, 3000)],
                         ids=['chunked', 'persistent', 'persistent-split', 'static',
                              'overflow'])
def test_gpu_patch_pairs_traversal_modes(gpu, oracle_lib, mode, levels, min_jobs, chunk):
    """Batched patch pairs across many chunked launches, in one persistent launch (from
    the initial node pairs or after a chunked split; every workgroup meets many patch
    pairs and recycles its LDS histograms), from static shares, and through the fallback
    of a persistent launch whose stacks are too small."""
    from treecorr_amd import _lib
    O = oracle_lib
    L = _lib.lib()
    ng, min_top = 3, 4
    cat, scale = C.make_catalog(THREED, 5000, 331, 0.3)
    labels = grid_labels(cat['x'], cat['y'], scale, ng)
    bn = C.binning(scale, 10, 0.01, 0.4, 0.1, 0.1)
    gg = corr_obj(gpu, 4, bn, 0.1, 0.1, min_top)
    _lib.check(L.tc_set_traversal_chunk(chunk))
    _lib.check(L.tc_set_traversal_mode(mode))
    _lib.check(L.tc_set_persist_params(levels, min_jobs))
    fb0 = L.tc_traversal_fallbacks()
    try:
        gg.process(front_catalog(gpu, cat, THREED, labels), finalize=False)
    finally:
        _lib.check(L.tc_set_traversal_chunk(0))
        _lib.check(L.tc_set_traversal_mode(1))
        _lib.check(L.tc_set_persist_params(0, 0))
    assert L.tc_traversal_fallbacks() - fb0 == (1 if levels else 0)
    for (i, j), res in gg.results.items():
        ref = oracle_raw(O, 4, THREED, sub(cat, labels == i),
                         None if i == j else sub(cat, labels == j), bn, min_top)
        check_pair(res, ref, 4)


def _xi_gg(raw):
    w = raw['weight']
    m = w != 0
    xip = np.zeros_like(w)
    xim = np.zeros_like(w)
    xip[m] = raw['xi0'][m] / w[m]
    xim[m] = raw['xi2'][m] / w[m]
    return np.concatenate([xip, xim])


@pytest.mark.gpu
def test_gpu_jackknife_gg_rows_are_catalogs_without_patch(gpu, oracle_lib):
    O = oracle_lib
    ng, min_top = 3, 3
    cat, scale = C.make_catalog(THREED, 3000, 341, 0.2)
    labels = grid_labels(cat['x'], cat['y'], scale, ng)
    # bin_slop = angle_slop = 0: every pair is binned and projected on its own, so the
    # sums do not depend on the tree
    bn = C.binning(scale, 8, 0.02, 0.3, 0.0, 0.0)
    gg = corr_obj(gpu, 4, bn, 0.0, 0.0, min_top, var_method='jackknife')
    gg.process(front_catalog(gpu, cat, THREED, labels))
    npatch = ng * ng
    rows = np.array([_xi_gg(oracle_raw(O, 4, THREED, sub(cat, labels != k), None, bn, min_top))
                     for k in range(npatch)])
    v, w = gg.build_cov_design_matrix('jackknife', cross_patch_weight='simple')
    assert v.shape == rows.shape
    for k in range(npatch):
        assert normwise(v[k], rows[k]) < 1e-11, k
    dv = rows - rows.mean(axis=0)
    cov = (1. - 1. / npatch) * dv.T.dot(dv)
    got = gg.estimate_cov('jackknife', cross_patch_weight='simple')
    assert normwise(got, cov) < 1e-8
    # varxip/varxim are the jackknife diagonal
    np.testing.assert_allclose(gg.varxip, np.diag(got)[:bn['nbins']], rtol=1e-14)

    # 'mean' and 'match' weight the pairs with one point in patch k by wc:
    # S = (1-wc) S(catalog without k) + wc (S(all) - S(patch k alone)).
    full = oracle_raw(O, 4, THREED, cat, None, bn, min_top)
    for cpw in ('mean', 'match'):
        wc = 0.5 if cpw == 'mean' else 1 - npatch / (2 + 2 ** 0.5 * (npatch - 1))
        v, _ = gg.build_cov_design_matrix('jackknife', cross_patch_weight=cpw)
        for k in range(npatch):
            wo = oracle_raw(O, 4, THREED, sub(cat, labels != k), None, bn, min_top)
            kk_ = oracle_raw(O, 4, THREED, sub(cat, labels == k), None, bn, min_top)
            comb = {key: (1 - wc) * wo[key] + wc * (full[key] - kk_[key])
                    for key in ('xi0', 'xi2', 'weight')}
            assert normwise(v[k], _xi_gg(comb)) < 1e-11, (cpw, k)


@pytest.mark.gpu
def test_gpu_jackknife_nn_landy_szalay(gpu, oracle_lib):
    """DD, RR, DR with patches; jackknife rows are Landy-Szalay on the catalogs with
    one patch removed."""
    O = oracle_lib
    ng, min_top = 3, 3
    dat, scale = C.make_catalog(FLAT, 2500, 351, 0.3)
    ran, _ = C.make_catalog(FLAT, 5000, 352, 0.0)
    # unit weights: a cell's weight (a float) is then exact, so with bin_slop = 0 the
    # sums do not depend on the tree
    dat['w'][:] = 1.
    ran['w'][:] = 1.
    ld = grid_labels(dat['x'], dat['y'], scale, ng)
    lr = grid_labels(ran['x'], ran['y'], scale, ng)
    bn = C.binning(scale, 8, 0.01, 0.25, 0.0, 1.0)
    D = front_catalog(gpu, dat, FLAT, ld)
    R = front_catalog(gpu, ran, FLAT, lr)
    kw = dict(var_method='jackknife', cross_patch_weight='simple')
    dd = corr_obj(gpu, 0, bn, 0.0, 1.0, min_top, **kw)
    rr = corr_obj(gpu, 0, bn, 0.0, 1.0, min_top, **kw)
    dr = corr_obj(gpu, 0, bn, 0.0, 1.0, min_top, **kw)
    dd.process(D)
    rr.process(R)
    dr.process(D, R)
    xi, varxi = dd.calculateXi(rr=rr, dr=dr)
    npatch = ng * ng

    def ls(dsel, rsel):
        wd, wr = np.sum(dat['w'][dsel]), np.sum(ran['w'][rsel])
        ddk = oracle_raw(O, 0, FLAT, sub(dat, dsel), None, bn, min_top)['weight']
        rrk = oracle_raw(O, 0, FLAT, sub(ran, rsel), None, bn, min_top)['weight']
        drk = oracle_raw(O, 0, FLAT, sub(dat, dsel), sub(ran, rsel), bn, min_top)['weight']
        dd_tot, rr_tot, dr_tot = 0.5 * wd ** 2, 0.5 * wr ** 2, wd * wr
        rrf, drf = dd_tot / rr_tot, dd_tot / dr_tot
        den = rrk * rrf
        out = ddk - 2. * drk * drf + den
        den[den == 0] = 1.
        return out / den

    np.testing.assert_allclose(xi, ls(np.ones(len(ld), bool), np.ones(len(lr), bool)),
                               rtol=1e-10, atol=1e-13)
    rows = np.array([ls(ld != k, lr != k) for k in range(npatch)])
    v, _ = dd.build_cov_design_matrix('jackknife')
    for k in range(npatch):
        assert normwise(v[k], rows[k]) < 1e-10, k
    dv = rows - rows.mean(axis=0)
    cov = (1. - 1. / npatch) * dv.T.dot(dv)
    assert normwise(dd.cov, cov) < 1e-7
    np.testing.assert_allclose(varxi, np.diag(dd.cov), rtol=1e-14)


@pytest.mark.gpu
def test_gpu_sample_cov_rows(gpu, oracle_lib):
    """Sample covariance row k sums the stored pairs (k, j) (cross_patch_weight simple)."""
    O = oracle_lib
    ng, min_top = 3, 3
    cat, scale = C.make_catalog(THREED, 3000, 361, 0.2)
    labels = grid_labels(cat['x'], cat['y'], scale, ng)
    bn = C.binning(scale, 8, 0.02, 0.3, 0.0, 1.0)
    kk = corr_obj(gpu, 1, bn, 0.0, 1.0, min_top)
    kk.process(front_catalog(gpu, cat, THREED, labels))
    v, w = kk.build_cov_design_matrix('sample', cross_patch_weight='simple')
    for k in range(ng * ng):
        s = np.zeros(bn['nbins'])
        ws = np.zeros(bn['nbins'])
        for j in range(k, ng * ng):
            if (k, j) in kk.results:
                ref = oracle_raw(O, 1, THREED, sub(cat, labels == k),
                                 None if j == k else sub(cat, labels == j), bn, min_top)
                s += ref['xi0']
                ws += ref['weight']
        m = ws != 0
        row = np.zeros_like(s)
        row[m] = s[m] / ws[m]
        assert normwise(v[k], row) < 1e-11, k
        assert math.isclose(w[k], np.sum(ws), rel_tol=1e-12)


@pytest.mark.gpu
@pytest.mark.parametrize('dim', [2, 3])
def test_gpu_quick_assign_matches_oracle(gpu, oracle_lib, dim):
    """Catalog(patch_centers=...) labels every object with its nearest centre on the
    GPU, exactly as the oracle's QuickAssign: more centres than one LDS batch, a
    duplicated centre (the lower index wins the tie) and objects sitting on a centre."""
    O = oracle_lib
    coords = THREED if dim == 3 else FLAT
    cat, scale = C.make_catalog(coords, 100_000, 371 + dim, 0.3)
    rng = np.random.default_rng(dim)
    npatch = 1500
    cen = rng.uniform(0, scale, (npatch, dim))
    cen[700] = cen[3]                       # a tie: patch 3 must win
    for q in range(5):                      # objects exactly on centres
        cat['x'][q], cat['y'][q] = cen[3 * q, 0], cen[3 * q, 1]
        if dim == 3:
            cat['z'][q] = cen[3 * q, 2]
    kw = dict(x=cat['x'], y=cat['y'], w=cat['w'], patch_centers=cen)
    if dim == 3:
        kw['z'] = cat['z']
    c = gpu.Catalog(**kw)
    ref = O.quick_assign(cen, cat['x'], cat['y'], cat['z'] if dim == 3 else None)
    np.testing.assert_array_equal(c.patch, ref)
    assert c.npatch == npatch and not np.any(c.patch == 700)


@pytest.mark.gpu
def test_gpu_jackknife_ng_rows_are_catalogs_without_patch(gpu, oracle_lib):
    """NG with patches (every patch pair in one traversal, the N fields on one side and
    the G fields on the other): with bin_slop = angle_slop = 0, jackknife row k is
    the oracle's NG of the lens and source catalogs without patch k."""
    O = oracle_lib
    ngrid, min_top = 3, 3
    lc, scale = C.make_catalog(THREED, 2000, 351, 0.2)
    sc, _ = C.make_catalog(THREED, 3000, 352, 0.2)
    ll = grid_labels(lc['x'], lc['y'], scale, ngrid)
    sl = grid_labels(sc['x'], sc['y'], scale, ngrid)
    bn = C.binning(scale, 8, 0.02, 0.3, 0.0, 0.0)
    ng = gpu.NGCorrelation(nbins=bn['nbins'], min_sep=bn['min_sep'], max_sep=bn['max_sep'],
                           bin_slop=0., angle_slop=0., min_top=min_top, var_method='jackknife')
    lens = {k: lc[k] for k in ('x', 'y', 'z', 'w')}
    ng.process(front_catalog(gpu, lens, THREED, ll), front_catalog(gpu, sc, THREED, sl))
    kw = dict(min_size=bn['min_size'], max_size=bn['max_size'], min_top=min_top)

    def raw(c1, c2):
        f1 = O.Field(0, THREED, c1['x'], c1['y'], c1['z'], w=c1['w'], **kw)
        f2 = O.Field(4, THREED, c2['x'], c2['y'], c2['z'], g1=c2['g1'], g2=c2['g2'], w=c2['w'],
                     **kw)
        return O.corr2(0, THREED, f1, f2, bn['nbins'], bn['min_sep'], bn['max_sep'],
                       bn['bin_size'], bn['b'], bn['a'], exact=True, d2=4)
    full = raw(lc, sc)
    np.testing.assert_array_equal(ng.npairs, full['npairs'])
    assert normwise(ng.raw_xi * ng.weight, full['xi0']) < 1e-11
    npatch = ngrid * ngrid
    rows = []
    for k in range(npatch):
        r = raw(sub(lc, ll != k), sub(sc, sl != k))
        rows.append(np.where(r['weight'] != 0, r['xi0'] / np.where(r['weight'] != 0,
                                                                   r['weight'], 1.), 0.))
    rows = np.array(rows)
    v, w = ng.build_cov_design_matrix('jackknife', cross_patch_weight='simple')
    assert v.shape == rows.shape
    for k in range(npatch):
        assert normwise(v[k], rows[k]) < 1e-11, k


@pytest.mark.gpu
@pytest.mark.parametrize('alt', [False, True], ids=['lloyd', 'alt'])
@pytest.mark.parametrize('coords,npatch', [(3, 20), (1, 12), (2, 16), (3, 90)])
def test_gpu_kmeans_matches_compiled_reference(gpu, coords, npatch, alt):
    """Field.run_kmeans's steps against the compiled reference from the same initial
    centres (which are the reference's own for the same rng, see test_kmeans_init.py):
    the reference iterates over tree cells with float weights, the GPU over the objects,
    so the converged centres agree to ~1e-7 of the field size and the labels almost
    everywhere.  alt=True: the patch inertia comes from the reference's own cell
    decomposition on the GPU (a centre farther than that can move the result by far
    more than the tolerance, since the inertia terms shift every boundary)."""
    from oracle import ref as R
    if R.module() is None:
        pytest.skip('oracle/_ref not built')
    cat, scale = C.make_catalog(coords, 20000, 700 + npatch, 0.3)
    if coords == 3:
        tcat = gpu.Catalog(ra=np.arctan2(cat['y'], cat['x']), dec=np.arcsin(cat['z']),
                           ra_units='rad', dec_units='rad', w=cat['w'])
    else:
        kw = dict(x=cat['x'], y=cat['y'], w=cat['w'])
        if coords == 2:
            kw['z'] = cat['z']
        tcat = gpu.Catalog(**kw)
    max_top = int.bit_length(npatch) - 1
    fld = tcat.getNField(max_top=max_top)
    c0 = fld.kmeans_initialize_centers(npatch, rng=np.random.default_rng(5))
    # the alternate iteration does not settle on these clumped catalogs (in the reference
    # as here): compare 30 steps of it
    it, tol = (30, 1e-14) if alt else (200, 1e-5)
    ours = c0.copy()
    print('kmeans coords', coords, 'npatch', npatch, 'alt', alt, flush=True)
    fld.kmeans_refine_centers(ours, alt=alt, max_iter=it, tol=tol)
    labels = fld.kmeans_assign_patches(ours)

    zc = None if coords == 1 else tcat.z
    rf = R.make_field(0, coords, tcat.x, tcat.y, zc, w=tcat.w, min_size=0., max_size=np.inf,
                      min_top=0, max_top=max_top)
    ref = c0.copy()
    rf.KMeansRun(ref, npatch, it, tol, bool(alt))
    rlab = np.zeros(tcat.ntot, dtype=int)
    rf.KMeansAssign(ref, npatch, rlab)
    assert np.max(np.abs(ours - ref)) < 1e-5 * fld.size
    assert np.mean(labels == rlab) > 0.995
    assert len(np.unique(labels)) == npatch
    if alt:
        # the inertia moves the result: plain Lloyd from the same start ends elsewhere
        plain = c0.copy()
        fld.kmeans_refine_centers(plain, max_iter=it, tol=tol)
        assert np.max(np.abs(plain - ref)) > 1e-3 * fld.size


@pytest.mark.gpu
def test_gpu_catalog_npatch_uses_kmeans(gpu):
    """Catalog(npatch=N): k-means patches (catalog.py:1473-1484); every label is the
    nearest centre and every patch holds objects."""
    cat, scale = C.make_catalog(3, 30000, 808, 0.3)
    tcat = gpu.Catalog(ra=np.arctan2(cat['y'], cat['x']), dec=np.arcsin(cat['z']),
                       ra_units='rad', dec_units='rad', w=cat['w'], npatch=25,
                       rng=np.random.default_rng(3))
    assert tcat.npatch == 25
    p = np.asarray(tcat.patch)
    assert np.array_equal(np.unique(p), np.arange(25))
    cen = tcat.patch_centers
    pos = np.column_stack([tcat.x, tcat.y, tcat.z])
    d = ((pos[:, None, :] - cen[None, :, :]) ** 2).sum(axis=2)
    assert np.mean(np.argmin(d, axis=1) == p) > 0.999
    assert len(tcat.get_patches()) == 25


@pytest.mark.gpu
@pytest.mark.parametrize('dim', [2, 3])
def test_gpu_select_patch_matches_oracle(gpu, oracle_lib, dim):
    """SelectPatch on the GPU against the oracle's restatement (itself pinned to the
    compiled reference), with two equal centres; and a single-patch Catalog built from
    patch_centers keeps exactly those objects."""
    import ctypes
    from treecorr_amd import _lib
    rng = np.random.default_rng(31 + dim)
    n, npatch = 60_000, 29
    x, y, z = rng.random(n), rng.random(n), rng.random(n)
    cen = rng.random((npatch, dim))
    cen[11] = cen[3]
    for patch in (0, 3, 11, 28):
        use = np.empty(n, dtype=np.int32)
        _lib.check(_lib.lib().tc_select_patch(
            patch, _lib.dptr(np.ascontiguousarray(cen.ravel())), npatch, _lib.dptr(x),
            _lib.dptr(y), _lib.dptr(z) if dim == 3 else None, n,
            use.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))
        ref = oracle_lib.select_patch(patch, cen, x, y, z if dim == 3 else None)
        np.testing.assert_array_equal(use, ref)
    kw = dict(x=x, y=y) if dim == 2 else dict(x=x, y=y, z=z)
    cat = gpu.Catalog(patch=3, patch_centers=cen, **kw)
    sel = oracle_lib.select_patch(3, cen, x, y, z if dim == 3 else None).astype(bool)
    np.testing.assert_array_equal(cat.x, x[sel])
    assert cat.patch == 3 and cat.npatch == npatch


@pytest.mark.gpu
@pytest.mark.parametrize('cross', [False, True])
def test_gpu_local_patch_method(gpu, oracle_lib, cross):
    """patch_method='local' (corr2base.py:782-834, 918-926, 1016-1024): each patch of the
    first catalog against the objects of the other near it.  With exact binning every
    pair within max_sep of a patch's objects is among them, so results[(i, i)] equals the
    oracle's cross correlation of patch i with the whole other catalog, and the totals
    are the sums over the patches."""
    O = oracle_lib
    ng, min_top = 4, 3
    c1, scale = C.make_catalog(FLAT, 5000, 351, 0.3)
    c2 = C.make_catalog(FLAT, 4000, 352, 0.3)[0] if cross else c1
    l1 = grid_labels(c1['x'], c1['y'], scale, ng)
    l2 = grid_labels(c2['x'], c2['y'], scale, ng)
    bn = C.binning(scale, 10, 0.01, 0.15, 0.0, 0.0)
    kk = corr_obj(gpu, 1, bn, 0.0, 0.0, min_top)
    a = front_catalog(gpu, c1, FLAT, l1)
    if cross:
        kk.process(a, front_catalog(gpu, c2, FLAT, l2), patch_method='local', finalize=False)
    else:
        kk.process(a, patch_method='local', finalize=False)
    assert sorted(kk.results) == [(i, i) for i in range(ng * ng)]
    tot = 0
    for i in range(ng * ng):
        ref = oracle_raw(O, 1, FLAT, sub(c1, l1 == i), c2, bn, min_top)
        check_pair(kk.results[(i, i)], ref, 1)
        tot += ref['npairs']
    np.testing.assert_array_equal(kk.npairs, tot)


@pytest.mark.gpu
def test_gpu_write_patch_results_round_trip(gpu, tmp_path):
    """write(write_patch_results=True) then from_file: the object read back gives the
    same jackknife and bootstrap covariances as the one processed (corr2base.py:
    1959-2050), for KK and for NN with randoms (its rr written alongside)."""
    rng = np.random.default_rng(44)
    n, nr = 20000, 40000
    x, y = rng.uniform(0, 100, n), rng.uniform(0, 100, n)
    k = rng.normal(0, 1, n)
    cat = gpu.Catalog(x=x, y=y, k=k, npatch=8, rng=np.random.default_rng(1))
    rcat = gpu.Catalog(x=rng.uniform(0, 100, nr), y=rng.uniform(0, 100, nr),
                       patch_centers=cat.patch_centers)
    print('write_patch_results round trip n', n, flush=True)
    kk = gpu.KKCorrelation(min_sep=1., max_sep=20., nbins=8, var_method='jackknife')
    kk.process(cat)
    f = str(tmp_path / 'kk.out')
    kk.write(f, write_patch_results=True)
    kk2 = gpu.KKCorrelation.from_file(f)
    np.testing.assert_allclose(kk2.estimate_cov('jackknife'), kk.estimate_cov('jackknife'),
                               rtol=1e-13, atol=0)
    dd = gpu.NNCorrelation(min_sep=1., max_sep=20., nbins=8)
    rr = gpu.NNCorrelation(min_sep=1., max_sep=20., nbins=8)
    dd.process(cat)
    rr.process(rcat)
    dd.calculateXi(rr=rr)
    f = str(tmp_path / 'nn.out')
    dd.write(f, rr=rr, write_patch_results=True, precision=16)
    dd2 = gpu.NNCorrelation.from_file(f)
    assert dd2._rr is not None and dd2.tot == dd.tot
    xi2, _ = dd2.calculateXi(rr=dd2._rr)
    np.testing.assert_allclose(xi2, dd.xi, rtol=1e-10)
    np.testing.assert_allclose(dd2.estimate_cov('jackknife'), dd.estimate_cov('jackknife'),
                               rtol=1e-12)
