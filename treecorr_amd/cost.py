"""The learned cost model of the persistent 2-point traversal (treecorr_amd/csrc/tc_cost.hip).

A job of the persistent traversal is a node pair, the root of a subtree of visits.  The
model, a perceptron of three FP32 layers (11 features, two hidden layers of 64 ReLU
units, one output), predicts y = log2(1 + items processed) of a job from its two nodes;
its forward, backward and Adam kernels are the library's own.  Traversal mode
LEARNED_STATIC gives each workgroup a contiguous share of the jobs with equal predicted
work, LEARNED_DYNAMIC hands them out in order of decreasing predicted work; the static
equal-count partition (STATIC) stays the fallback and the judge.  The reference hands
out its top-level pairs with OpenMP's schedule(dynamic) (src/Corr2.cpp:158-203); it has
no cost model.

Features of a job (k_job_features): log2(1 + s1/min_sep), log2(1 + s2/min_sep),
log2 n1, log2 n2, log2(d/min_sep), min((s1+s2)/d, 8), process2 flag, log2(1 + d/max_sep),
log2 b, log2 bin_size, 3-d flag (d the separation of the centres, floored at
1e-3 min_sep; s the sizes, n the counts).
"""
import ctypes

import numpy as np

from . import _lib

CHUNKED, DYNAMIC, STATIC, LEARNED_STATIC, LEARNED_DYNAMIC = 0, 1, 2, 3, 4

_fp = ctypes.POINTER(ctypes.c_float)


def _f32(a):
    return np.ascontiguousarray(a, np.float32)


def _ptr(a):
    return a.ctypes.data_as(_fp)


def sizes():
    """(number of features, number of parameters)."""
    nf, npar = ctypes.c_int(), ctypes.c_int()
    _lib.check(_lib.lib().tc_cost_model_sizes(ctypes.byref(nf), ctypes.byref(npar)))
    return nf.value, npar.value


class CostModel:
    """The MLP on the GPU."""

    def __init__(self, seed=0):
        self._h = ctypes.c_void_p()
        _lib.check(_lib.lib().tc_cost_model_create(seed, ctypes.byref(self._h)))
        self.nfeatures, self.nparams = sizes()

    def close(self):
        if self._h:
            _lib.lib().tc_cost_model_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def params(self):
        p = np.zeros(self.nparams, np.float32)
        _lib.check(_lib.lib().tc_cost_model_params(self._h, _ptr(p), 0))
        return p

    @params.setter
    def params(self, p):
        p = _f32(p)
        if p.shape != (self.nparams,):
            raise ValueError('expected %d parameters' % self.nparams)
        _lib.check(_lib.lib().tc_cost_model_params(self._h, _ptr(p), 1))

    def predict(self, X):
        X = _f32(X).reshape(-1, self.nfeatures)
        y = np.zeros(len(X), np.float32)
        _lib.check(_lib.lib().tc_cost_model_predict(self._h, _ptr(X), len(X), _ptr(y)))
        return y

    def grad(self, X, y):
        """Gradient of the mean squared error over all samples, and the loss."""
        X = _f32(X).reshape(-1, self.nfeatures)
        y = _f32(y)
        g = np.zeros(self.nparams, np.float32)
        loss = ctypes.c_double()
        _lib.check(_lib.lib().tc_cost_model_grad(self._h, _ptr(X), _ptr(y), len(X), _ptr(g),
                                                  ctypes.byref(loss)))
        return g, loss.value

    def train(self, X, y, steps=2000, batch=4096, lr=2e-3, standardize=True):
        """Adam on mini-batches; returns the mean squared error over all samples.

        With standardize (the default), the GPU trains on features shifted and scaled to
        mean 0 and variance 1 (the statistics of the first training set), and the scaling
        is then folded into the first layer (W1 / sd, b1 - W1 mu / sd), so that the model
        reads the traversal's raw features."""
        X = _f32(X).reshape(-1, self.nfeatures)
        y = _f32(y)
        if standardize:
            if getattr(self, '_mu', None) is None:
                self._mu = X.astype(np.float64).mean(0)
                self._sd = X.astype(np.float64).std(0) + 1e-6
            self._fold(unfold=True)
            X = _f32((X - self._mu) / self._sd)
        loss = ctypes.c_double()
        _lib.check(_lib.lib().tc_cost_model_train(self._h, _ptr(X), _ptr(y), len(X), steps, batch,
                                                   ctypes.c_float(lr), ctypes.byref(loss)))
        if standardize:
            self._fold(unfold=False)
        return loss.value

    def _fold(self, unfold):
        """Fold the feature scaling into the first layer, or take it out again."""
        if getattr(self, '_mu', None) is None or getattr(self, '_folded', False) != unfold:
            return
        nf, h = self.nfeatures, 64
        p = self.params.astype(np.float64)
        W1 = p[:h * nf].reshape(h, nf)
        b1 = p[h * nf:h * nf + h]
        if unfold:      # W = W1 sd, b = b1 + W1 mu
            b1 = b1 + W1 @ self._mu
            W1 = W1 * self._sd
        else:           # W1 = W / sd, b1 = b - W mu / sd
            W1 = W1 / self._sd
            b1 = b1 - W1 @ self._mu
        p[:h * nf] = W1.ravel()
        p[h * nf:h * nf + h] = b1
        self.params = p.astype(np.float32)
        self._folded = not unfold

    def install(self):
        """Use this model in the learned traversal modes."""
        _lib.check(_lib.lib().tc_set_cost_model(self._h))


def set_mode(mode):
    _lib.check(_lib.lib().tc_set_traversal_mode(mode))


def get_mode():
    return _lib.lib().tc_set_traversal_mode(-1)


def clear_model():
    _lib.check(_lib.lib().tc_set_cost_model(None))


def start_log(max_samples):
    """Log the jobs of the following persistent traversals (at most max_samples each)."""
    _lib.check(_lib.lib().tc_set_job_log(max_samples))


def stop_log():
    """Stop logging; returns (features, items processed) of the logged jobs."""
    L = _lib.lib()
    n = ctypes.c_longlong()
    _lib.check(L.tc_job_log_fetch(None, None, 0, ctypes.byref(n)))
    nf, _ = sizes()
    X = np.zeros((n.value, nf), np.float32)
    v = np.zeros(n.value, np.float32)
    _lib.check(L.tc_job_log_fetch(_ptr(X), _ptr(v), n.value, ctypes.byref(n)))
    _lib.check(L.tc_set_job_log(0))
    return X, v


def target(visits):
    """The model's target for a job that processed `visits` items."""
    return np.log2(1. + np.asarray(visits, np.float64)).astype(np.float32)
