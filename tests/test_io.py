"""Output files: TreeCorr's ASCII layout, and write/read round trips of the correlation
classes with their patch results (no GPU: the arrays are filled by hand)."""
import os

import numpy as np
import pytest


def _filled(cls, seed, nbins=6, npatch=3, **kw):
    from treecorr_amd.corr2 import _PairSums
    rng = np.random.default_rng(seed)
    c = cls(min_sep=1., max_sep=50., nbins=nbins, **kw)
    c.coords, c.metric = 'flat', 'Euclidean'
    for a in [c.meanr, c.meanlogr, c.weight, c.npairs] + [x for x in c._xi if x is not None]:
        a[:] = rng.normal(size=nbins)
    c.weight[:] = np.abs(c.weight)
    c._var_num = 0.3
    c.npatch1 = c.npatch2 = npatch
    for i in range(npatch):
        for j in range(i, npatch):
            ps = _PairSums(nbins, c._nxi, rng.normal(size=(8, nbins)))
            if (i, j) == (0, 2):
                ps.weight[:] = 0.       # a pair without weight
            ps.tot = float(rng.uniform(1, 2))
            c.results[(i, j)] = ps
    if hasattr(c, 'tot'):
        c.tot = 12.5
    return c


def test_ascii_layout_matches_treecorr_writer(tmp_path):
    """writer.py:51-77: '## ext', '## {params}', names centred in precision + 8 columns
    (the first one narrower by the '#'), rows in %{p+8}.{p}e."""
    from treecorr_amd.io import make_writer
    f = str(tmp_path / 'x.dat')
    with make_writer(f, 3) as w:
        w.write(['r_nom', 'xi'], [np.array([1.5, 20.]), np.array([-0.25, 3e-7])],
                params={'corr': 'KK', 'tot': np.float64(2.)}, ext='main')
    lines = open(f).read().splitlines()
    assert lines[0] == '## main'
    assert lines[1] == "## {'corr': 'KK', 'tot': 2.0}"
    assert lines[2] == '#' + '{:^10}'.format('r_nom') + ' ' + '{:^11}'.format('xi')
    assert lines[3] == '%11.3e %11.3e' % (1.5, -0.25)
    assert lines[4] == '%11.3e %11.3e' % (20., 3e-7)


@pytest.mark.parametrize('name', ['KK', 'GG', 'NG', 'NK', 'KQ', 'NN'])
def test_write_read_round_trip_with_patch_results(tmp_path, name):
    import treecorr_amd as tca
    cls = getattr(tca, name + 'Correlation')
    c = _filled(cls, 3)
    f = str(tmp_path / ('c_%s.out' % name))
    c.write(f, write_patch_results=True, precision=16)
    d = cls.from_file(f)
    assert d.nbins == c.nbins and d.npatch1 == 3 and d.coords == 'flat'
    for nm in ('meanr', 'meanlogr', 'weight', 'npairs'):
        np.testing.assert_array_equal(getattr(d, nm), getattr(c, nm))
    for a, b in zip(c._xi, d._xi):
        if a is not None:
            np.testing.assert_array_equal(a, b)
    assert sorted(d.results) == sorted(c.results)
    for k, ps in c.results.items():
        q = d.results[k]
        if name == 'NN' and not ps._nonzero:
            assert q.tot == ps.tot and not np.any(q.weight)      # stored as a zero pair
            continue
        for nm in ('meanr', 'meanlogr', 'weight', 'npairs'):
            np.testing.assert_array_equal(getattr(q, nm), getattr(ps, nm))
        for a, b in zip(ps._xi, q._xi):
            if a is not None:
                np.testing.assert_array_equal(a, b)
    if name == 'NN':
        assert d.tot == c.tot
    # read() into an object of the same configuration
    e = cls(min_sep=1., max_sep=50., nbins=6)
    e.read(f)
    np.testing.assert_array_equal(e.weight, c.weight)


def test_default_precision_is_four_digits(tmp_path):
    import treecorr_amd as tca
    c = _filled(tca.KKCorrelation, 4)
    f = str(tmp_path / 'kk.out')
    c.write(f)
    d = tca.KKCorrelation.from_file(f)
    np.testing.assert_allclose(d.xi, c.xi, rtol=1e-4)
    assert not np.array_equal(d.xi, c.xi)
    assert d.results == {}


def test_fits_and_hdf_are_refused(tmp_path):
    import treecorr_amd as tca
    c = _filled(tca.KKCorrelation, 5)
    for ext in ('fits', 'hdf5'):
        with pytest.raises(NotImplementedError):
            c.write(str(tmp_path / ('kk.' + ext)))


@pytest.mark.parametrize('kind,bin_type', [('GGG', 'LogSAS'), ('GGG', 'LogMultipole'),
                                           ('NNN', 'LogSAS'), ('NNN', 'LogMultipole')])
def test_corr3_write_read_round_trip(tmp_path, kind, bin_type):
    """3-point files (corr3base.py:3352-3662, gggcorrelation.py:224-275,
    nnncorrelation.py:459-770): the column layout of the reference, and every array,
    patch result, NNN tot and the variances back from Corr3.from_file."""
    import treecorr_amd as tc
    rng = np.random.default_rng(7)
    kw = dict(min_sep=1., max_sep=10., nbins=3, bin_type=bin_type, sep_units='arcmin')
    kw.update(max_n=4) if bin_type == 'LogMultipole' else kw.update(nphi_bins=5)
    c = getattr(tc, kind + 'Correlation')(**kw)
    c.coords, c.metric = 'flat', 'Euclidean'

    def fill(o):
        for a in o._sums():
            # meanv belongs to LogRUV binning, meanu (phi) to LogSAS
            if a is not o.meanv and (a is not o.meanu or bin_type == 'LogSAS'):
                a[...] = rng.uniform(0.5, 2., a.shape)
        if kind == 'NNN':
            o.tot = float(rng.uniform(1, 9))
    fill(c)
    c._var_num = 0.25
    c.npatch1 = c.npatch2 = c.npatch3 = 3
    for i in range(3):
        r = c.copy()
        r.results = {}
        fill(r)
        c.results[(i, i, i)] = r
    if kind == 'NNN':
        c.results[(1, 1, 1)]._clear()       # an empty result keeps only its tot
    f = str(tmp_path / ('%s.out' % kind))
    c.write(f, write_patch_results=True)
    header = open(f).readlines()[2].split()
    first = ['#', 'd2_nom', 'd3_nom', 'n' if bin_type == 'LogMultipole' else 'phi_nom']
    assert header[:4] == first
    c2 = tc.Corr3.from_file(f)
    assert type(c2) is type(c) and c2.bin_type == bin_type and c2.npatch3 == 3
    for a, b in zip(c2._sums(), c._sums()):
        np.testing.assert_allclose(a, b, rtol=1e-4)     # the main table keeps 4 digits
    assert sorted(c2.results) == sorted(c.results)
    for key, r in c.results.items():
        for a, b in zip(c2.results[key]._sums(), r._sums()):
            np.testing.assert_allclose(a, b, rtol=1e-15)
        if kind == 'NNN':
            assert c2.results[key].tot == r.tot
    if kind == 'GGG':
        np.testing.assert_allclose(c2.vargam2, c.vargam2, rtol=1e-4)
    else:
        assert c2.tot == c.tot


def test_nnn_write_with_randoms_round_trip(tmp_path):
    """NNN files with the randoms (nnncorrelation.py:459-770): zeta from calculateZeta,
    the RRR, DRR, RDD columns scaled to DDD's total, and with write_patch_results the
    randoms' groups, read back into _rrr, _drr, _rdd with their patch results."""
    import treecorr_amd as tc
    rng = np.random.default_rng(9)
    kw = dict(min_sep=1., max_sep=10., nbins=3, nphi_bins=4)

    def make(tot):
        c = tc.NNNCorrelation(**kw)
        c.coords, c.metric = 'flat', 'Euclidean'
        c.npatch1 = c.npatch2 = c.npatch3 = 2
        c.weightr[...] = rng.uniform(1., 2., c.weightr.shape)
        c.ntri[...] = c.weightr * 3
        c.tot = tot
        for i in range(2):
            r = tc.Corr3.copy(c)
            r.results = {}
            r.weightr[...] = c.weightr / 2
            r.tot = tot / 2
            c.results[(i, i, i)] = r
        return c
    ddd, rrr, drr, rdd = make(10.), make(40.), make(20.), make(30.)
    f = str(tmp_path / 'nnn.out')
    ddd.write(f, rrr=rrr, drr=drr, rdd=rdd, write_patch_results=True)
    header = open(f).readlines()[2].split()
    assert header[-7:] == ['zeta', 'sigma_zeta', 'DDD', 'RRR', 'DRR', 'RDD', 'ntri']
    want = (ddd.weightr - rdd.weightr / 3 + drr.weightr / 2 - rrr.weightr / 4) / \
        (rrr.weightr / 4)
    np.testing.assert_allclose(ddd.zeta, want, rtol=1e-14)
    d2 = tc.Corr3.from_file(f)
    np.testing.assert_allclose(d2.zeta, ddd.zeta, rtol=1e-4)
    assert d2.tot == ddd.tot
    for nm, o in (('_rrr', rrr), ('_drr', drr), ('_rdd', rdd)):
        r2 = getattr(d2, nm)
        assert r2.tot == o.tot and sorted(r2.results) == sorted(o.results)
        np.testing.assert_allclose(r2.results[(1, 1, 1)].weightr, o.results[(1, 1, 1)].weightr,
                                   rtol=1e-15)
