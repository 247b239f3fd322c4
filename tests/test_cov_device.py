"""Patch-based covariances: the bootstrap and marked-bootstrap pair lists (against
hand-worked cases of the reference's iterators, treecorr/corr2base.py:1861-1928), the
GPU's design-matrix row sums and Gram product against numpy, and every estimator's
design matrix and covariance on the GPU against the per-row host sums that the
reference's _make_cov_design_matrix_core (:2321-2368) makes, with the same draws."""
import numpy as np
import pytest

import cases as C
from test_gpu_patches import THREED, FLAT, corr_obj, front_catalog, grid_labels


class _Stub:
    """The pair-list methods of a Corr2 over a given set of stored patch pairs."""

    def __init__(self, keys, n):
        import treecorr_amd as tc
        self.obj = tc.KKCorrelation(nbins=3, min_sep=1., max_sep=10.)
        self.obj.npatch1 = self.obj.npatch2 = n
        self.obj.results = {k: None for k in keys}


def test_bootstrap_pairs_hand_worked():
    # 3 patches; stored pairs: the autos (0,0), (2,2), and (0,1), (0,2), (1,2), (2,0)
    s = _Stub([(0, 0), (2, 2), (0, 1), (0, 2), (1, 2), (2, 0)], 3).obj
    index = np.array([0, 2, 0])           # patch 0 twice, patch 2 once
    got = sorted(s._bootstrap_pairs(index, 'simple'))
    # autos at their multiplicity; crosses among drawn patches with w1*w2
    assert got == sorted([(0, 0, 2), (2, 2, 1), (0, 2, 2), (2, 0, 2)])
    got = sorted(s._bootstrap_pairs(index, 'geom'))
    assert got == sorted([(0, 0, 2), (2, 2, 1), (0, 2, 2 ** 0.5), (2, 0, 2 ** 0.5)])
    got = sorted(s._bootstrap_pairs(index, 'mean'))
    # mean: every stored cross pair, (m_i + m_j) / 2 with m the multiplicity (0 if undrawn)
    assert got == sorted([(0, 0, 2), (2, 2, 1), (0, 1, 1.), (0, 2, 1.5), (1, 2, 0.5),
                          (2, 0, 1.5)])
    with pytest.raises(ValueError):
        s._bootstrap_pairs(index, 'match')


def test_marked_pairs_hand_worked():
    s = _Stub([(0, 0), (2, 2), (0, 1), (0, 2), (1, 2), (2, 0)], 3).obj
    index = np.array([0, 2, 0])
    got = sorted(s._marked_pairs(index, 'simple'))
    # every stored pair whose first patch was drawn, at that patch's multiplicity
    assert got == sorted([(0, 0, 2), (0, 1, 2), (0, 2, 2), (2, 0, 1), (2, 2, 1)])
    got = sorted(s._marked_pairs(index, 'mean'))
    # mean: pairs with the drawn patch on either side, half weight off the diagonal,
    # stored in whichever order they are kept
    assert got == sorted([(0, 0, 2), (0, 1, 1.), (0, 2, 1.), (2, 0, 0.5), (1, 2, 0.5),
                          (2, 2, 1)])
    # one patch on the other side: (i, 0) pairs at each draw
    s1 = _Stub([(0, 0), (1, 0), (2, 0)], 3).obj
    s1.npatch2 = 1
    assert s1._marked_pairs(np.array([1, 1, 2]), 'simple') == [(1, 0, 1), (1, 0, 1), (2, 0, 1)]


@pytest.mark.gpu
@pytest.mark.parametrize('nrow,npair,nq', [(1, 1, 1), (37, 250, 24), (130, 1000, 240),
                                           (500, 65, 300)])
def test_gpu_rowsums_and_gram_match_numpy(gpu, nrow, npair, nq):
    from treecorr_amd import _lib
    L = _lib.lib()
    rng = np.random.default_rng(nrow + npair)
    W = rng.uniform(0, 2, (nrow, npair)) * (rng.random((nrow, npair)) < 0.6)
    S = rng.standard_normal((npair, nq))
    R = np.zeros((nrow, nq))
    _lib.check(L.tc_cov_rowsums(nrow, npair, nq, _lib.dptr(W), _lib.dptr(S), _lib.dptr(R)))
    ref = W @ S
    assert np.max(np.abs(R - ref)) <= 1e-13 * max(1., np.max(np.abs(ref)))
    if nrow > 1:
        V = rng.standard_normal((nrow, nq)) + 3.
        w = rng.uniform(0.1, 1., nrow)
        for wt in (None, w):
            Cm = np.zeros((nq, nq))
            _lib.check(L.tc_cov_gram(nrow, nq, _lib.dptr(V),
                                     None if wt is None else _lib.dptr(wt), _lib.dptr(Cm)))
            vc = V - V.mean(axis=0)
            Cr = (vc.T * (1. if wt is None else wt)).dot(vc)
            assert np.max(np.abs(Cm - Cr)) <= 1e-12 * np.max(np.abs(Cr))


def _host_design(corrs, plist, func):
    """The per-row host sums (Corr2._sum without the GPU), as the reference forms them."""
    corrs = [c.copy() for c in corrs]
    if func is None:
        func = lambda cs: np.concatenate([c.getStat() for c in cs])
    v1 = func(corrs)
    v = np.zeros((len(plist), len(v1)), dtype=v1.dtype)
    w = np.zeros(len(plist))
    for row, pairs in enumerate(plist):
        for c, cpairs in zip(corrs, pairs):
            cpairs = c._keep_ok(cpairs)
            if len(cpairs) == 0:
                c._clear()
            else:
                c._calculate_xi_from_pairs(cpairs, corr_only=c._corr_only)
        v[row] = func(corrs)
        w[row] = np.sum([np.sum(c.getWeight()) for c in corrs])
    return v, w


@pytest.mark.gpu
@pytest.mark.parametrize('method,cpw', [('jackknife', 'simple'), ('jackknife', 'match'),
                                        ('sample', 'simple'), ('bootstrap', 'simple'),
                                        ('bootstrap', 'geom'), ('bootstrap', 'mean'),
                                        ('marked_bootstrap', 'simple'),
                                        ('marked_bootstrap', 'mean')])
def test_gpu_design_matrix_and_cov_match_host_rows(gpu, method, cpw):
    from treecorr_amd import cov as COV
    ng, min_top = 3, 3
    cat, scale = C.make_catalog(THREED, 6000, 341, 0.3)
    labels = grid_labels(cat['x'], cat['y'], scale, ng)
    bn = C.binning(scale, 8, 0.01, 0.4, 0.1, 0.1)
    kk = corr_obj(gpu, 1, bn, 0.1, 1.0, min_top, num_bootstrap=60,
                  rng=np.random.default_rng(5))
    gg = corr_obj(gpu, 4, bn, 0.1, 0.1, min_top, num_bootstrap=60,
                  rng=np.random.default_rng(5))
    c = front_catalog(gpu, cat, THREED, labels)
    kk.process(c)
    gg.process(c)
    corrs = [kk, gg]
    v, w = COV.build_multi_cov_design_matrix(corrs, method, cross_patch_weight=cpw)
    # the same draws on the host: a fresh Generator in the same state
    kk._rng = np.random.default_rng(5)
    npatch = ng * ng
    if method == 'jackknife':
        plist = list(zip(*[c._jackknife_pairs(cpw) for c in corrs]))
    elif method == 'sample':
        plist = list(zip(*[c._sample_pairs(cpw) for c in corrs]))
    else:
        of = '_bootstrap_pairs' if method == 'bootstrap' else '_marked_pairs'
        plist = []
        for _ in range(60):
            index = kk.rng.choice(np.arange(npatch), size=npatch, replace=True)
            plist.append([getattr(c, of)(index, cpw) for c in corrs])
    vh, wh = _host_design(corrs, plist, None)
    assert v.shape == vh.shape and v.shape[0] == (60 if 'boot' in method else npatch)
    assert np.max(np.abs(v - vh)) <= 1e-12 * np.max(np.abs(vh))
    np.testing.assert_allclose(w, wh, rtol=1e-13)
    # the covariance from the GPU's Gram product against numpy's on the host rows
    kk._rng = np.random.default_rng(5)
    cov = COV.estimate_multi_cov(corrs, method, cross_patch_weight=cpw)
    vc = vh - vh.mean(axis=0)
    if method == 'jackknife':
        ref = (1 - 1. / npatch) * vc.T.dot(vc)
    elif method == 'sample':
        ww = wh / wh.sum()
        ref = 1. / (npatch - 1) * (ww * vc.T).dot(vc)
    else:
        ref = 1. / 59 * vc.T.dot(vc)
    assert np.max(np.abs(cov - ref)) <= 1e-11 * np.max(np.abs(ref))
