"""Patch-based covariance estimates of 2-point correlations: shot, jackknife, sample.

Restates treecorr/corr2base.py's covariance layer: estimate_multi_cov (:2131-2258),
build_multi_cov_design_matrix (:2260-2319), _make_cov_design_matrix_core
(:2321-2368), _cov_shot (:2393-2408), _check_patch_nums (:2410-2421), the
jackknife (:2423-2445) and sample (:2446-2475) estimators.  The per-row pair lists
come from each correlation's _jackknife_pairs/_sample_pairs (:1787-1850).
"""
import numpy as np

_METHODS = ('shot', 'jackknife', 'sample', 'bootstrap', 'marked_bootstrap')


def _check_method(method):
    if method not in _METHODS:
        raise ValueError("Invalid method: %s" % method)
    if 'bootstrap' in method:
        raise NotImplementedError("method=%r is not implemented by treecorr_amd" % method)


def estimate_multi_cov(corrs, method, *, func=None, comm=None, num_bootstrap=None,
                       cross_patch_weight=None):
    """Joint covariance of several correlations (corr2base.py:2131-2258)."""
    if comm is not None:
        raise NotImplementedError("MPI is not part of the single-GPU engine")
    _check_method(method)
    cross_patch_weight = corrs[0]._check_cpw(method, cross_patch_weight)
    if method == 'shot':
        if func is not None:
            raise ValueError("func is invalid with method='shot'")
        return _cov_shot(corrs)
    if method == 'jackknife':
        return _cov_jackknife(corrs, func, cross_patch_weight)
    return _cov_sample(corrs, func, cross_patch_weight)


def build_multi_cov_design_matrix(corrs, method, *, func=None, comm=None, num_bootstrap=None,
                                  cross_patch_weight=None):
    """Design matrix and row weights of a covariance estimate (corr2base.py:2260-2319)."""
    if comm is not None:
        raise NotImplementedError("MPI is not part of the single-GPU engine")
    _check_method(method)
    if method == 'shot':
        raise ValueError("There is no design matrix for method='shot'")
    cross_patch_weight = corrs[0]._check_cpw(method, cross_patch_weight)
    if method == 'jackknife':
        return _design_jackknife(corrs, func, cross_patch_weight)
    return _design_sample(corrs, func, cross_patch_weight)


def _make_cov_design_matrix(corrs, plist, func, name):
    # corr2base.py:2321-2368: each row sums the listed patch pairs of every correlation,
    # finalizes, and evaluates func on the list of correlations.
    corrs = [c.copy() for c in corrs]
    if func is None:
        func = lambda cs: np.concatenate([c.getStat() for c in cs])
    v1 = func(corrs)
    v = np.zeros((len(plist), len(v1)), dtype=v1.dtype)
    w = np.zeros(len(plist), dtype=float)
    for row, pairs in enumerate(plist):
        for c, cpairs in zip(corrs, pairs):
            cpairs = c._keep_ok(cpairs)
            if len(cpairs) == 0:
                c._clear()
            else:
                # func is never None here, so this is the object's own corr_only.
                c._calculate_xi_from_pairs(cpairs, corr_only=c._corr_only)
        v[row] = func(corrs)
        w[row] = np.sum([np.sum(c.getWeight()) for c in corrs])
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


def _cov_jackknife(corrs, func, cross_patch_weight):
    # C = (1-1/npatch) Sum_i (v_i - v_mean) (v_i - v_mean)^T  (corr2base.py:2432-2445)
    v, w = _design_jackknife(corrs, func, cross_patch_weight)
    npatch = v.shape[0]
    vmean = np.mean(v, axis=0)
    v -= vmean
    return (1. - 1. / npatch) * v.conj().T.dot(v)


def _design_sample(corrs, func, cross_patch_weight):
    _check_patch_nums(corrs, 'sample')
    plist = list(zip(*[c._sample_pairs(cross_patch_weight) for c in corrs]))
    return _make_cov_design_matrix(corrs, plist, func, 'sample')


def _cov_sample(corrs, func, cross_patch_weight):
    # C = 1/(npatch-1) Sum_i w_i (v_i - v_mean) (v_i - v_mean)^T  (corr2base.py:2446-2475)
    v, w = _design_sample(corrs, func, cross_patch_weight)
    npatch = v.shape[0]
    if np.any(w == 0):
        raise RuntimeError("Cannot compute sample variance when some patches have no data.")
    w /= np.sum(w)
    vmean = np.mean(v, axis=0)
    v -= vmean
    return 1. / (npatch - 1) * (w * v.conj().T).dot(v)
