"""The learned cost model of the persistent traversal: the GPU's forward, backward and
Adam kernels against the float64 numpy restatement (tests/mlp_ref.py, whose gradient is
pinned by finite differences here), the job logs against the traversal's counters, the
learned modes' results against the oracle, and what the trained network gains over the
static equal-count partition it replaces."""
import numpy as np
import pytest

import cases as C
import mlp_ref as R


def _random_params(rng, scale=0.3):
    return (rng.standard_normal(R.NP) * scale).astype(np.float32)


def _data(rng, n):
    X = rng.standard_normal((n, R.NF)).astype(np.float32)
    y = (np.sin(X[:, 0]) + 0.5 * X[:, 1] * X[:, 2] + np.abs(X[:, 3])).astype(np.float32)
    return X, y


def test_reference_gradient_matches_finite_differences():
    """The numpy backward pass against central differences of its loss (float64)."""
    rng = np.random.default_rng(5)
    p = rng.standard_normal(R.NP) * 0.3
    X, y = _data(rng, 40)
    _, g = R.loss_and_grad(p, X, y)
    idx = np.concatenate([rng.choice(np.arange(R.O_W1, R.O_B1), 6), [R.O_B1 + 3],
                          rng.choice(np.arange(R.O_W2, R.O_B2), 6), [R.O_B2 + 7],
                          [R.O_W3 + 1, R.O_W3 + 40, R.O_B3]])
    h = 1e-6
    for k in idx:
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd = (R.loss_and_grad(pp, X, y)[0] - R.loss_and_grad(pm, X, y)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1., abs(g[k])), (k, fd, g[k])


def test_reference_batch_hash_is_uniform():
    idx = R.batch_idx(7, 3, 1000, 4000)
    assert idx.min() >= 0 and idx.max() < 1000
    counts = np.bincount(idx, minlength=1000)
    assert counts.max() < 16 and (counts > 0).mean() > 0.95


@pytest.mark.gpu
def test_gpu_forward_and_gradient_match_reference(gpu):
    from treecorr_amd import cost
    rng = np.random.default_rng(11)
    m = cost.CostModel(seed=3)
    assert (m.nfeatures, m.nparams) == (R.NF, R.NP)
    p = _random_params(rng)
    m.params = p
    np.testing.assert_array_equal(m.params, p)
    X, y = _data(rng, 700)
    yhat = m.predict(X)
    ref, _ = R.forward(p, X)
    np.testing.assert_allclose(yhat, ref, rtol=2e-5, atol=2e-5 * np.max(np.abs(ref)))
    g, loss = m.grad(X, y)
    lref, gref = R.loss_and_grad(p, X, y)
    assert abs(loss - lref) <= 1e-5 * lref
    err = np.max(np.abs(g - gref)) / np.max(np.abs(gref))
    assert err < 1e-5, err
    # every parameter block has a non-trivial gradient
    for a, b in ((R.O_W1, R.O_B1), (R.O_B1, R.O_W2), (R.O_W2, R.O_B2), (R.O_W3, R.O_B3)):
        assert np.any(np.abs(gref[a:b]) > 0)


@pytest.mark.gpu
def test_gpu_adam_steps_match_reference(gpu):
    """Three training steps on the GPU equal three numpy Adam steps on the same batches."""
    from treecorr_amd import cost
    rng = np.random.default_rng(12)
    m = cost.CostModel(seed=9)
    p = _random_params(rng, 0.2)
    m.params = p
    X, y = _data(rng, 500)
    batch, lr, seed = 256, 1e-2, 9
    m.train(X, y, steps=3, batch=batch, lr=lr, standardize=False)
    q = p.astype(np.float64)
    mm = np.zeros(R.NP)
    vv = np.zeros(R.NP)
    for t in range(3):
        idx = R.batch_idx(seed, t, len(X), batch)
        _, g = R.loss_and_grad(q, X[idx], y[idx])
        q, mm, vv = R.adam(q, g, mm, vv, t + 1, lr)
    got = m.params.astype(np.float64)
    assert np.max(np.abs(got - q)) < 2e-4 * np.max(np.abs(q)), np.max(np.abs(got - q))


@pytest.mark.gpu
def test_gpu_standardized_training_reads_raw_features(gpu):
    """Training on standardized features, folded into the first layer afterwards, gives
    the same predictions on raw features as the unfolded network on standardized ones."""
    from treecorr_amd import cost
    rng = np.random.default_rng(14)
    m = cost.CostModel(seed=4)
    X, y = _data(rng, 5000)
    X = X * np.linspace(0.5, 20., R.NF, dtype=np.float32) + np.linspace(-30., 30., R.NF,
                                                                         dtype=np.float32)
    m.train(X, y, steps=200, batch=1024, lr=3e-3)
    p = m.params.astype(np.float64)
    W1 = p[R.O_W1:R.O_B1].reshape(R.H, R.NF) * m._sd
    b1 = p[R.O_B1:R.O_W2] + p[R.O_W1:R.O_B1].reshape(R.H, R.NF) @ m._mu
    q = p.copy()
    q[R.O_W1:R.O_B1] = W1.ravel()
    q[R.O_B1:R.O_W2] = b1
    ref, _ = R.forward(q, (X - m._mu) / m._sd)
    np.testing.assert_allclose(m.predict(X), ref, rtol=1e-4, atol=1e-4 * np.max(np.abs(ref)))
    # a second call trains on from the same scaling
    m.train(X, y, steps=50, batch=1024, lr=3e-3)
    assert m._folded


@pytest.mark.gpu
def test_gpu_training_fits_a_function(gpu):
    from treecorr_amd import cost
    rng = np.random.default_rng(13)
    m = cost.CostModel(seed=1)
    X, y = _data(rng, 20000)
    base = float(np.var(y))
    loss = m.train(X, y, steps=1500, batch=2048, lr=3e-3)
    assert loss < 0.05 * base, (loss, base)


def _gg_catalogs():
    cat, scale = C.make_catalog(C.SPHERE, 40000, 401, 0.5)
    cat2, _ = C.make_catalog(C.SPHERE, 30000, 402, 0.5)
    bn = C.binning(scale, 15, 0.005, 0.3, 0.1, 0.1)
    return cat, cat2, bn


def _raw(gpu, cat, cat2, bn):
    from test_gpu_corr2 import raw_corr2
    return raw_corr2(gpu, 4, 3, cat, cat2, bn, min_top=3)


@pytest.mark.gpu
def test_gpu_job_log_counts_every_item(gpu):
    """A logged cross correlation from its initial node pairs (no chunked split): the
    logged items of all jobs add up to the traversal's visit counter."""
    from treecorr_amd import _lib, cost
    L = _lib.lib()
    cat, cat2, bn = _gg_catalogs()
    _lib.check(L.tc_set_persist_params(0, 1))
    try:
        cost.set_mode(cost.DYNAMIC)
        cost.start_log(10 ** 7)
        out = _raw(gpu, cat, cat2, bn)
        X, v = cost.stop_log()
    finally:
        _lib.check(L.tc_set_persist_params(0, 0))
    assert len(v) > 0 and X.shape == (len(v), R.NF)
    assert int(v.sum()) == out['counters']['visits']
    assert np.all(X[:, 6] == 0)        # no process2 jobs in a cross correlation
    assert np.all(X[:, 10] == 1)       # 3-d positions (the sphere)


@pytest.mark.gpu
def test_gpu_learned_modes_match_oracle_and_beat_equal_counts(gpu, oracle_lib):
    """Train on the logs of one catalog pair, then on another: the learned static and
    learned dynamic modes give the oracle's npairs and counters, and the trained
    network's shares of the jobs are better balanced than equal counts."""
    from treecorr_amd import _lib, cost
    L = _lib.lib()
    O = oracle_lib
    cat, cat2, bn = _gg_catalogs()
    m = cost.CostModel(seed=2)
    _lib.check(L.tc_set_persist_params(0, 4096))
    try:
        cost.set_mode(cost.DYNAMIC)
        cost.start_log(10 ** 6)
        _raw(gpu, cat, None, bn)
        _raw(gpu, cat, cat2, bn)
        X, v = cost.stop_log()
        y = cost.target(v)
        loss = m.train(X, y, steps=1500, batch=2048, lr=3e-3)
        assert loss < 0.5 * np.var(y), (loss, np.var(y))
        m.install()
        # held out: another catalog
        c3, scale = C.make_catalog(C.SPHERE, 50000, 403, 0.6)
        bn3 = C.binning(scale, 15, 0.005, 0.3, 0.1, 0.1)
        cost.start_log(10 ** 7)
        _raw(gpu, c3, None, bn3)
        Xh, vh = cost.stop_log()
        outs = {}
        for mode in (cost.STATIC, cost.LEARNED_STATIC, cost.LEARNED_DYNAMIC):
            cost.set_mode(mode)
            outs[mode] = _raw(gpu, c3, None, bn3)
    finally:
        cost.set_mode(cost.DYNAMIC)
        cost.clear_model()
        _lib.check(L.tc_set_persist_params(0, 0))
    kw = dict(min_size=bn3['min_size'], max_size=bn3['max_size'], min_top=3)
    f = O.Field(4, 3, c3['x'], c3['y'], c3['z'], g1=c3['g1'], g2=c3['g2'], w=c3['w'], **kw)
    ref = O.corr2(4, 3, f, None, bn3['nbins'], bn3['min_sep'], bn3['max_sep'], bn3['bin_size'],
                  bn3['b'], bn3['a'], exact=True)
    for mode, out in outs.items():
        np.testing.assert_array_equal(out['npairs'], ref['npairs'])
        assert out['counters'] == ref['counters'], mode
    # the gain: the largest share of the work when G workgroups take contiguous shares of
    # equal job counts, against shares of equal predicted work
    pred = m.predict(Xh)
    cost_pred = np.exp2(np.clip(pred, 0, 60)) - 1
    G = 256
    n = len(vh)

    def worst_share(bounds):
        c = np.concatenate([[0.], np.cumsum(vh.astype(np.float64))])
        return np.max(c[bounds[1:]] - c[bounds[:-1]]) / (c[-1] / G)

    equal = np.array([n * w // G for w in range(G + 1)])
    cp = np.cumsum(cost_pred)
    learned = np.concatenate([[0], np.searchsorted(cp, cp[-1] * np.arange(1, G) / G, side='right'), [n]])
    r = np.corrcoef(pred, cost.target(vh))[0, 1]
    assert r > 0.8, r
    assert worst_share(learned) < 0.7 * worst_share(equal), (worst_share(learned), worst_share(equal))
