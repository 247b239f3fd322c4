"""Patch-based covariance estimates of 2-point correlations: shot, jackknife, sample,
bootstrap and marked bootstrap.

Restates treecorr/corr2base.py's covariance layer: estimate_multi_cov (:2131-2258),
build_multi_cov_design_matrix (:2260-2319), _make_cov_design_matrix_core
(:2321-2368), _cov_shot (:2393-2408), _check_patch_nums (:2410-2421), the
jackknife (:2423-2445), sample (:2446-2475), marked bootstrap (:2480-2515) and
bootstrap (:2517-2538) estimators.  The per-row pair lists come from each
correlation's _jackknife_pairs/_sample_pairs/_marked_pairs/_bootstrap_pairs
(:1787-1928); the bootstrap draws are the reference's, rng.choice(npatch, npatch) from
the first correlation's numpy Generator, so the same rng gives the same rows.

The linear algebra runs on the GPU (treecorr_amd/csrc/tc_cov.hip).  Every row of a
design matrix is a weighted sum of patch-pair sums, formed by the classes' own
finalize logic (_calculate_xi_from_pairs, with NN's randoms and the compensated
corrections) through Corr2._sum.  The rows are built in two passes: the first records
which pair sums each _sum call adds with which weights; the GPU then forms all of
them at once, R = W S (tc_cov_rowsums); the second pass replays the same calls with
the results and evaluates func on the finished objects.  The covariance of the rows is
one more product on the GPU (tc_cov_gram).
"""
import ctypes

import numpy as np

from . import _lib

_METHODS = ('shot', 'jackknife', 'sample', 'bootstrap', 'marked_bootstrap')

# Set while a design matrix is formed: Corr2._sum hands its calls to it.
_SUM_HOOK = None


def _check_method(method):
    if method not in _METHODS:
        raise ValueError("Invalid method: %s" % method)


def _num_bootstrap(corrs, method, num_bootstrap):
    # corr2base.py:2241-2243: the largest of the correlations' num_bootstrap
    if num_bootstrap is None and 'boot' in method:
        num_bootstrap = int(np.max([c.num_bootstrap for c in corrs]))
    return num_bootstrap


def estimate_multi_cov(corrs, method, *, func=None, comm=None, num_bootstrap=None,
                       cross_patch_weight=None):
    """Joint covariance of several correlations (corr2base.py:2131-2258)."""
    if comm is not None:
        raise NotImplementedError("MPI is not part of the single-GPU engine")
    _check_method(method)
    num_bootstrap = _num_bootstrap(corrs, method, num_bootstrap)
    cross_patch_weight = corrs[0]._check_cpw(method, cross_patch_weight)
    if method == 'shot':
        if func is not None:
            raise ValueError("func is invalid with method='shot'")
        return _cov_shot(corrs)
    if method == 'jackknife':
        v, w = _design_jackknife(corrs, func, cross_patch_weight)
        # C = (1-1/npatch) Sum_i (v_i - v_mean) (v_i - v_mean)^T  (:2432-2445)
        return (1. - 1. / v.shape[0]) * _gram(v, None)
    if method == 'sample':
        v, w = _design_sample(corrs, func, cross_patch_weight)
        # C = 1/(npatch-1) Sum_i w_i (v_i - v_mean) (v_i - v_mean)^T  (:2464-2475)
        if np.any(w == 0):
            raise RuntimeError("Cannot compute sample variance when some patches have no data.")
        w = w / np.sum(w)
        return 1. / (v.shape[0] - 1) * _gram(v, w)
    if method == 'bootstrap':
        v, w = _design_bootstrap(corrs, func, num_bootstrap, cross_patch_weight)
    else:
        v, w = _design_marked(corrs, func, num_bootstrap, cross_patch_weight)
    # C = 1/(nboot-1) Sum_i (v_i - v_mean) (v_i - v_mean)^T  (:2507-2515, 2529-2538)
    return 1. / (num_bootstrap - 1) * _gram(v, None)


def build_multi_cov_design_matrix(corrs, method, *, func=None, comm=None, num_bootstrap=None,
                                  cross_patch_weight=None):
    """Design matrix and row weights of a covariance estimate (corr2base.py:2260-2319)."""
    if comm is not None:
        raise NotImplementedError("MPI is not part of the single-GPU engine")
    _check_method(method)
    if method == 'shot':
        raise ValueError("There is no design matrix for method='shot'")
    num_bootstrap = _num_bootstrap(corrs, method, num_bootstrap)
    cross_patch_weight = corrs[0]._check_cpw(method, cross_patch_weight)
    if method == 'jackknife':
        return _design_jackknife(corrs, func, cross_patch_weight)
    if method == 'sample':
        return _design_sample(corrs, func, cross_patch_weight)
    if method == 'bootstrap':
        return _design_bootstrap(corrs, func, num_bootstrap, cross_patch_weight)
    return _design_marked(corrs, func, num_bootstrap, cross_patch_weight)


def _gram(v, w):
    """Sum_r w_r (v_r - mean)^H (v_r - mean) (w None: unit weights), on the GPU."""
    if np.iscomplexobj(v):
        # (a - ib)^T (a + ib) = a^T a + b^T b + i (a^T b - b^T a): one real product of
        # the columns [Re v, Im v]
        n = v.shape[1]
        G = _gram(np.concatenate([v.real, v.imag], axis=1), w)
        return (G[:n, :n] + G[n:, n:]) + 1j * (G[:n, n:] - G[n:, :n])
    v = np.ascontiguousarray(v, dtype=float)
    nrow, ncol = v.shape
    C = np.zeros((ncol, ncol))
    wp = None if w is None else _lib.dptr(np.ascontiguousarray(w, dtype=float))
    _lib.check(_lib.lib().tc_cov_gram(nrow, ncol, _lib.dptr(v), wp, _lib.dptr(C)))
    return C


class _Rows:
    """The two passes of a design matrix's rows (see the module docstring)."""

    def __init__(self):
        self.calls = []          # pass 1: (corr, [(pair sums, weight)], corr_only)
        self.results = None      # pass 2: the row sums of every call, in call order
        self.next = 0

    def record(self, corr, others, corr_only):
        self.calls.append((corr, others, corr_only))

    def replay(self, corr, others, corr_only):
        call, _, _ = self.calls[self.next]
        if call is not corr:
            raise RuntimeError("design matrix rows replayed out of order")
        row = self.results[self.next]
        self.next += 1
        off = 0
        for a in corr._row_arrays(corr_only):
            a[...] = row[off:off + a.size].reshape(a.shape)
            off += a.size

    def solve(self):
        """Every recorded call's weighted sum, R = W S, on the GPU: one product per kind
        of call (class, corr_only and array sizes), as the calls of one design matrix may
        mix correlations.  A row of S is one pair sum's arrays, in the order of the
        class's _row_arrays."""
        self.results = [None] * len(self.calls)
        groups = {}
        for k, (corr, _, corr_only) in enumerate(self.calls):
            sizes = tuple(a.size for a in corr._row_arrays(corr_only))
            groups.setdefault((type(corr), bool(corr_only), sizes), []).append(k)
        for (_, corr_only, sizes), ks in groups.items():
            ncol = sum(sizes)
            pid = {}
            stack = []
            for k in ks:
                for ps, _ in self.calls[k][1]:
                    if id(ps) not in pid:
                        pid[id(ps)] = len(stack)
                        stack.append(ps)
            S = np.zeros((max(1, len(stack)), ncol))
            for r, ps in enumerate(stack):
                arrs = ps._row_arrays(corr_only)
                if tuple(a.size for a in arrs) != sizes:
                    raise ValueError("a patch result does not match its correlation's layout")
                S[r] = np.concatenate([a.ravel() for a in arrs])
            W = np.zeros((len(ks), S.shape[0]))
            for row, k in enumerate(ks):
                for ps, w in self.calls[k][1]:
                    W[row, pid[id(ps)]] += w
            R = np.zeros((len(ks), ncol))
            _lib.check(_lib.lib().tc_cov_rowsums(len(ks), S.shape[0], ncol, _lib.dptr(W),
                                                 _lib.dptr(S), _lib.dptr(R)))
            for row, k in enumerate(ks):
                self.results[k] = R[row]


def _make_cov_design_matrix(corrs, plist, func, name):
    # corr2base.py:2321-2368: each row sums the listed patch pairs of every correlation,
    # finalizes, and evaluates func on the list of correlations.
    global _SUM_HOOK
    corrs = [c.copy() for c in corrs]
    if func is None:
        func = lambda cs: np.concatenate([c.getStat() for c in cs])
    v1 = func(corrs)
    plist = [[c._keep_ok(cpairs) for c, cpairs in zip(corrs, pairs)] for pairs in plist]

    def rows(evaluate):
        v = np.zeros((len(plist), len(v1)), dtype=v1.dtype) if evaluate else None
        w = np.zeros(len(plist), dtype=float)
        for row, pairs in enumerate(plist):
            for c, cpairs in zip(corrs, pairs):
                if len(cpairs) == 0:
                    c._clear()
                else:
                    # func is never None here, so this is the object's own corr_only.
                    c._calculate_xi_from_pairs(cpairs, corr_only=c._corr_only)
            if evaluate:
                v[row] = func(corrs)
                w[row] = np.sum([np.sum(c.getWeight()) for c in corrs])
        return v, w

    R = _Rows()
    try:
        _SUM_HOOK = R.record
        rows(False)
        R.solve()
        _SUM_HOOK = R.replay
        v, w = rows(True)
    finally:
        _SUM_HOOK = None
    return v, w


def _cov_shot(corrs):
    # corr2base.py:2393-2408: var_num / weight (NN: / RR weight), the diagonal only
    vlist = []
    for c in corrs:
        v = np.array(c.getWeight(), dtype=float).copy()
        mask1 = v != 0
        v[mask1] = c._var_num / v[mask1]
        vlist.append(v)
    return np.concatenate(vlist)


def _check_patch_nums(corrs, name):
    # corr2base.py:2410-2421
    for c in corrs:
        if len(c.results) == 0:
            raise ValueError("Using %s covariance requires using patches." % name)
    npatch = corrs[0]._get_npatch()
    for c in corrs[1:]:
        if c._get_npatch() != npatch:
            raise RuntimeError("All correlations must use the same number of patches")
    return npatch


def _design_jackknife(corrs, func, cross_patch_weight):
    _check_patch_nums(corrs, 'jackknife')
    plist = list(zip(*[c._jackknife_pairs(cross_patch_weight) for c in corrs]))
    return _make_cov_design_matrix(corrs, plist, func, 'jackknife')


def _design_sample(corrs, func, cross_patch_weight):
    _check_patch_nums(corrs, 'sample')
    plist = list(zip(*[c._sample_pairs(cross_patch_weight) for c in corrs]))
    return _make_cov_design_matrix(corrs, plist, func, 'sample')


def _draws(corrs, npatch, num_bootstrap, pairs_of, cross_patch_weight):
    # corr2base.py:2484-2490, 2521-2524: npatch patches drawn with replacement per row
    plist = []
    for _ in range(num_bootstrap):
        index = corrs[0].rng.choice(np.arange(npatch), size=npatch, replace=True)
        plist.append([pairs_of(c)(index, cross_patch_weight) for c in corrs])
    return plist


def _design_marked(corrs, func, num_bootstrap, cross_patch_weight):
    npatch = _check_patch_nums(corrs, 'marked_bootstrap')
    plist = _draws(corrs, npatch, num_bootstrap, lambda c: c._marked_pairs, cross_patch_weight)
    return _make_cov_design_matrix(corrs, plist, func, 'marked_bootstrap')


def _design_bootstrap(corrs, func, num_bootstrap, cross_patch_weight):
    npatch = _check_patch_nums(corrs, 'bootstrap')
    plist = _draws(corrs, npatch, num_bootstrap, lambda c: c._bootstrap_pairs, cross_patch_weight)
    return _make_cov_design_matrix(corrs, plist, func, 'bootstrap')
