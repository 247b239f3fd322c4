"""LogSAS binning through toSAS, and the 3-point public API for cross-correlations.

The SAS weight is pinned by its definition: with brute=True every multipole term is a
single object, so the converted weight at phi is the sum over triangles of
w1 w2 w3 D_M(phi - Delta) dphi / 2 pi, D_M the Dirichlet kernel of order max_n and Delta
the triangle's opening angle from the d2 side to the d3 side -- computed here from the
positions alone.  The GGG conversion (Fourier sum, division by the weight and the turn to
the centroid projection) is pinned on single triangles whose opening angle is a phi bin
centre: there it must give the product of the three shears projected onto the directions
from each vertex to the centroid.  Each test prints its case before the device call."""
import math

import numpy as np
import pytest


def _flat_cat(gpu, x, y, w, g1=None, g2=None):
    kw = dict(x=np.asarray(x, float), y=np.asarray(y, float), w=np.asarray(w, float))
    if g1 is not None:
        kw.update(g1=np.asarray(g1, float), g2=np.asarray(g2, float))
    return gpu.Catalog(**kw)


@pytest.mark.gpu
def test_gpu_nnn_sas_weight_is_the_dirichlet_sum(gpu):
    rng = np.random.default_rng(17)
    n = 45
    x = rng.uniform(0, 10, n)
    y = rng.uniform(0, 10, n)
    w = rng.uniform(0.5, 1.5, n)
    nbins, min_sep, max_sep, max_n, nphi = 4, 1., 8., 9, 12
    print('nnn sas dirichlet n', n, flush=True)
    cat = _flat_cat(gpu, x, y, w)
    sas = gpu.NNNCorrelation(min_sep=min_sep, max_sep=max_sep, nbins=nbins, nphi_bins=nphi,
                             max_n=max_n, brute=True, bin_slop=0)
    sas.process(cat)
    bin_size = math.log(max_sep / min_sep) / nbins
    w32 = w.astype(np.float32).astype(float)
    phi_c = sas.phi1d
    dphi = sas.phi_bin_size
    nn = np.arange(-max_n, max_n + 1)
    expect = np.zeros((nbins, nbins, nphi))
    flipped = np.zeros((nbins, nbins, nphi))     # the opposite sense of phi
    for i in range(n):
        d = np.hypot(x - x[i], y - y[i])
        ang = np.arctan2(y - y[i], x - x[i])
        ok = (d >= min_sep) & (d < max_sep)
        k = np.floor(np.log(np.where(ok, d, 1.) / min_sep) / bin_size).astype(int)
        idx = np.where(ok)[0]
        for a in idx:          # the vertex across from d2 ... at distance d2 = d[a]
            for b in idx:      # ... and at distance d3 = d[b]
                if a == b:
                    continue
                delta = ang[b] - ang[a]
                www = w32[i] * w32[a] * w32[b] * dphi / (2 * np.pi)
                expect[k[a], k[b]] += www * np.cos(nn[:, None] * (phi_c[None, :] - delta)).sum(0)
                flipped[k[a], k[b]] += www * np.cos(nn[:, None] * (phi_c[None, :] + delta)).sum(0)
    den = np.max(np.abs(expect))
    assert np.max(np.abs(sas.weight - expect)) / den < 1e-10
    assert np.max(np.abs(sas.weight - flipped)) / den > 1e-3


def _centroid_product(p, g):
    """g1 g2 g3 projected onto the directions from each vertex to the centroid, and the
    products with one factor conjugated (gam0..gam3)."""
    c = (p[0] + p[1] + p[2]) / 3.
    gp = []
    for i in range(3):
        q = c - p[i]
        e = np.conj(q / abs(q)) ** 2
        gp.append(g[i] * e)
    g1, g2, g3 = gp
    return [g1 * g2 * g3, np.conj(g1) * g2 * g3, g1 * np.conj(g2) * g3, g1 * g2 * np.conj(g3)]


@pytest.mark.gpu
@pytest.mark.parametrize('j,d2,d3', [(7, 5.5, 2.3), (10, 1.3, 3.9), (8, 6.2, 1.4),
                                     (1, 7.0, 1.1)])
def test_gpu_ggg_sas_single_triangle_is_centroid_projected(gpu, j, d2, d3):
    nbins, min_sep, max_sep, max_n, nphi = 5, 1., 10., 14, 12
    sas = gpu.GGGCorrelation(min_sep=min_sep, max_sep=max_sep, nbins=nbins, nphi_bins=nphi,
                             max_n=max_n, brute=True, bin_slop=0)
    phi0 = sas.phi1d[j]
    p = [0j, d3 * np.exp(1j * phi0), complex(d2, 0.)]      # p1, p2 (d3 side), p3 (d2 side)
    d1 = abs(p[1] - p[2])
    bin_size = math.log(max_sep / min_sep) / nbins
    kb = [int(math.floor(math.log(d / min_sep) / bin_size)) for d in (d1, d2, d3)]
    assert len(set(kb)) == 3 and d1 < max_sep      # the cell (k(d2), k(d3)) has one triangle
    g = [0.21 - 0.13j, -0.08 + 0.3j, 0.17 + 0.05j]
    print('ggg sas triangle phi bin', j, 'd2', d2, 'd3', d3, flush=True)
    cat = _flat_cat(gpu, [q.real for q in p], [q.imag for q in p], [1., 1., 1.],
                    [q.real for q in g], [q.imag for q in g])
    sas.process(cat)
    # the tree keeps w g in single precision (include/Cell.h:131-169)
    g32 = [complex(np.float32(q.real), np.float32(q.imag)) for q in g]
    want = _centroid_product(p, g32)
    k2, k3 = kb[1], kb[2]
    got = [sas.gam0, sas.gam1, sas.gam2, sas.gam3]
    for mu in range(4):
        assert abs(got[mu][k2, k3, j] - want[mu]) < 1e-9 * max(1e-3, abs(want[mu])), mu
    # the wrong projection (no turn to the centroid) is far off
    assert abs(got[0][k2, k3, j] - np.prod(g)) > 1e-4


@pytest.mark.gpu
def test_gpu_sas_process_equals_multipole_then_tosas(gpu):
    rng = np.random.default_rng(5)
    n = 6000
    cat = gpu.Catalog(x=rng.uniform(0, 100, n), y=rng.uniform(0, 100, n),
                      g1=rng.normal(0, 0.2, n), g2=rng.normal(0, 0.2, n), w=rng.uniform(0.5, 1.5, n))
    kw = dict(min_sep=2., max_sep=20., nbins=6, min_top=3)
    print('sas process vs multipole+toSAS n', n, flush=True)
    sas = gpu.GGGCorrelation(nphi_bins=10, **kw)
    sas.process(cat)
    mp = gpu.GGGCorrelation(bin_type='LogMultipole', max_n=int(2 * np.pi / sas.phi_bin_size), **kw)
    mp.process(cat)
    sas2 = mp.toSAS(nphi_bins=10)
    for a, b in ((sas.weight, sas2.weight), (sas.gam0, sas2.gam0), (sas.gam3, sas2.gam3),
                 (sas.ntri, sas2.ntri), (sas.meand1, sas2.meand1), (sas.meanphi, sas2.meanphi)):
        # two runs: the moments' FP64 atomic adds may round differently in the last bit
        np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-12 * np.max(np.abs(b)))
    assert sas.weight.shape == (6, 6, 10)
    assert np.all(sas.meanphi == sas.phi)


@pytest.mark.gpu
def test_gpu_nnn_cross_public_api(gpu, oracle_lib):
    """NNNCorrelation.process(cat1, cat2[, cat3]) (LogMultipole): the raw sums of the
    oracle's passes, finalized by the class; tot as the reference accumulates it."""
    O = oracle_lib
    rng = np.random.default_rng(8)
    cats, raw = [], []
    for i in range(3):
        n = 3000 + 500 * i
        x, y, w = rng.uniform(0, 50, n), rng.uniform(0, 50, n), rng.uniform(0.5, 1.5, n)
        cats.append(gpu.Catalog(x=x, y=y, w=w))
        raw.append((x, y, w))
    kw = dict(min_sep=1., max_sep=10., nbins=5, max_n=6, bin_type='LogMultipole', min_top=3)
    nnn = gpu.NNNCorrelation(**kw)
    print('nnn cross public api', flush=True)
    nnn.process(cats[0], cats[1], cats[2], ordered=False)
    b1 = min(nnn.angle_slop, nnn.b)
    fk = dict(min_size=2 * nnn._min_sep * b1 / (2 + 3 * b1), max_size=nnn._max_sep * b1,
              min_top=3)
    of = [O.Field(0, 1, x, y, None, w=w, **fk) for x, y, w in raw]
    ref = O.corr3_multipole_cross(0, 1, 'cross', 0, of, nnn.nbins, nnn._min_sep, nnn._max_sep,
                                  nnn.bin_size, nnn.b, nnn.angle_slop, nnn.max_n)
    M = nnn.max_n
    np.testing.assert_array_equal(nnn.ntri[:, :, 0], ref['ntri'][:, :, M])
    np.testing.assert_allclose(nnn.weightr, ref['weight'], rtol=1e-12,
                               atol=1e-12 * np.abs(ref['weight']).max())
    assert nnn.tot == pytest.approx(cats[0].sumw * cats[1].sumw * cats[2].sumw, rel=1e-15)
    nnn.process(cats[0], cats[1])
    assert nnn.tot == pytest.approx(0.5 * cats[0].sumw * cats[1].sumw ** 2, rel=1e-15)
