"""Pairs placed within a few ulp of the Log binning's decision points, and the
decisions the reference takes for them, computed on the host with the host libm's log
(Python's math.log calls it, as the reference's std::log does).

Two families (NN, Flat coordinates):
  leaf pairs     two points at separation r = the edge exp(logminsep + k binsize) moved
                 by a few ulp; the pair is binned by directProcess11's
                 k = int((log r - logminsep) / binsize) (src/Corr2.cpp:492-497);
  refined test   a two-point cell {(0,-s), (0,s)} and a point (r, 0), with r moved ulp
                 by ulp across the point where singleBin's refined test
                 (include/BinType.h:128-147) turns from "fits one bin" to "split".
Each family's expected sums follow the reference's arithmetic step by step; the
correctly rounded log (by the decimal module) stands beside the libm one so that a case
the libm rounds incorrectly is known.
"""
import decimal
import math

import numpy as np

decimal.getcontext().prec = 50


def cr_log(x):
    """log x rounded correctly to double."""
    return float(decimal.Decimal(x).ln())


def ulps(x, j):
    """x moved by j ulp."""
    to = np.inf if j > 0 else -np.inf
    for _ in range(abs(j)):
        x = float(np.nextafter(x, to))
    return x


# (min_sep, bin_size, nbins): the leaf-pair layouts; the edges of each at ulp offsets
LEAF_LAYOUTS = [(1.0, 0.1, 10), (0.37, 0.0731, 24), (2.5, 0.3, 8), (0.011, 0.21, 18)]
LEAF_OFFSETS = range(-3, 4)


def layout(min_sep, bin_size, nbins):
    return dict(nbins=nbins, min_sep=min_sep, max_sep=min_sep * math.exp(nbins * bin_size),
                bin_size=bin_size, logminsep=math.log(min_sep))


def leaf_pairs(L):
    """Separations at every internal edge of layout L: the first r that the reference
    bins above the edge, moved by LEAF_OFFSETS ulp."""
    rs = []
    for k in range(1, L['nbins']):
        edge = math.exp(L['logminsep'] + k * L['bin_size'])
        lo, hi = edge * (1. - 1e-9), edge * (1. + 1e-9)
        assert bin_index(math.log(math.sqrt(lo * lo)), L) == k - 1
        assert bin_index(math.log(math.sqrt(hi * hi)), L) == k
        while True:
            mid = 0.5 * (lo + hi)
            if mid in (lo, hi):
                break
            if bin_index(math.log(math.sqrt(mid * mid)), L) >= k:
                hi = mid
            else:
                lo = mid
        rs += [ulps(hi, j) for j in LEAF_OFFSETS]
    return np.array(rs)


def leaf_catalog(L, rs):
    """Pair i at (0, y_i) and (r_i, y_i); rows 3 max_sep apart, so only the pairs meet.
    dx = r exactly and dy = 0, so the reference's rsq is r*r rounded once."""
    n = len(rs)
    y = np.arange(n) * 3. * math.ceil(L['max_sep'])
    return dict(x=np.concatenate([np.zeros(n), rs]), y=np.concatenate([y, y]), z=None,
                w=np.ones(2 * n), k=None, g1=None, g2=None)


def bin_index(logr, L):
    k = int((logr - L['logminsep']) / L['bin_size'])
    return L['nbins'] - 1 if k == L['nbins'] else k


def expect_leaf(L, rs, logf=math.log):
    """BaseCorr2::process11 -> singleBin (s1ps2 = 0) -> directProcess11 for each pair."""
    nb = L['nbins']
    out = {k: np.zeros(nb) for k in ('npairs', 'weight', 'meanr', 'meanlogr')}
    ks = []
    for r in rs:
        rsq = r * r
        if not (L['min_sep'] ** 2 <= rsq < L['max_sep'] ** 2):
            ks.append(-1)
            continue
        rr = math.sqrt(rsq)
        logr = logf(rr)
        k = bin_index(logr, L)
        ks.append(k)
        out['npairs'][k] += 1
        out['weight'][k] += 1.
        out['meanr'][k] += rr
        out['meanlogr'][k] += logr
    return out, np.array(ks)


# ---------------------------------------------------------------------------
# singleBin's refined test
# ---------------------------------------------------------------------------
REFINE = dict(nbins=10, min_sep=1.0, bin_size=0.1, b=0.05, a=1.0)
REFINE['max_sep'] = math.exp(REFINE['nbins'] * REFINE['bin_size'])
REFINE['logminsep'] = 0.0
REFINE_OFFSETS = range(-3, 4)
# cell sizes per (bin, side), as multiples of the size that puts the flip at the target
REFINE_SCALES = (0.985, 0.99, 0.995, 1.0, 1.005, 1.01)


def single_bin(rsq, s1ps2, P, logf=math.log):
    """BinTypeHelper<Log>::singleBin (include/BinType.h:93-151): (fits, ik, logr)."""
    if s1ps2 == 0.:
        return True, -1, None
    s1ps2sq = s1ps2 * s1ps2
    if s1ps2sq > (P['a'] * P['a']) * rsq:
        return False, -1, None
    if s1ps2sq <= (P['b'] * P['b']) * rsq:
        return True, -1, None
    if s1ps2sq > 0.25 * ((P['bin_size'] + P['b']) * (P['bin_size'] + P['b'])) * rsq:
        return False, -1, None
    logr = 0.5 * logf(rsq)
    kk = (logr - P['logminsep']) / P['bin_size']
    ik = int(kk)
    frackk = kk - ik
    f = min(frackk, 1. - frackk)
    t1 = f * P['bin_size'] + P['b']
    if s1ps2sq > (t1 * t1) * rsq:
        return False, ik, logr
    t2 = frackk * P['bin_size'] + P['b'] - s1ps2sq / rsq
    if s1ps2sq > (t2 * t2) * rsq:
        return False, ik, logr
    return True, ik, logr


def refine_rsq(s, r):
    """The cell {(0,-s),(0,s)} has centre (0,0) and size s; the point is at (r, 0)."""
    return r * r


def find_flip(s, r_fit, r_split, P):
    """Bisect the doubles between r_fit (fits) and r_split (split) for the last r that
    fits, by the reference's decision with the libm log."""
    lo, hi = r_fit, r_split
    assert single_bin(refine_rsq(s, lo), s, P)[0] and not single_bin(refine_rsq(s, hi), s, P)[0]
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if mid in (lo, hi):
            break
        if single_bin(refine_rsq(s, mid), s, P)[0]:
            lo = mid
        else:
            hi = mid
    return lo


def refine_cases(P=REFINE):
    """(s, r) per case: for bins 2, 5 and 8, the low side (frackk ~ 0.2, where the second
    refined comparison decides) and the high side (frackk ~ 0.85, the first), at several
    cell sizes, r swept over REFINE_OFFSETS ulp around the flip."""
    cases = []
    for k, side, sc in [(k, sd, sc) for k in (2, 5, 8) for sd in ('low', 'high')
                    for sc in REFINE_SCALES]:
        frac = 0.2 if side == 'low' else 0.85
        r0 = math.exp(P['logminsep'] + (k + frac) * P['bin_size'])
        if side == 'low':      # s/r + (s/r)^2 = frackk binsize + b
            t = frac * P['bin_size'] + P['b']
            u = (-1. + math.sqrt(1. + 4. * t)) / 2.
        else:                  # s/r = (1 - frackk) binsize + b
            u = (1. - frac) * P['bin_size'] + P['b']
        s = float(np.float32(sc * u * r0))    # node sizes are stored as float
        # r grows: u = s/r falls while the bin position moves; bracket the flip
        rs = np.linspace(r0 * 0.97, r0 * 1.03, 4001)
        fit = np.array([single_bin(refine_rsq(s, r), s, P)[0] for r in rs])
        idx = np.nonzero(fit[1:] != fit[:-1])[0]
        assert len(idx), (k, side)
        i = idx[0]
        a, b = (rs[i], rs[i + 1]) if fit[i] else (rs[i + 1], rs[i])
        rf = find_flip(s, a, b, P)
        cases += [(s, ulps(rf, j)) for j in REFINE_OFFSETS]
    return cases


def refine_catalog(s, r):
    """The three points of one case; the tree must keep its root as the top cell
    (min_top = 0), so that {P1, P2} is a cell of its own."""
    return dict(x=np.array([0., 0., r]), y=np.array([-s, s, 0.]), z=None, w=np.ones(3),
                k=None, g1=None, g2=None)


def expect_refine(s, r, P=REFINE, logf=math.log):
    """process2 of the root {P1, P2, Q}: the root splits along x into {P1, P2} and {Q};
    P1-P2 (2s < min_sep) drops; process11({P1,P2}, {Q}) bins the cell pair when it fits,
    else splits the cell into two leaf pairs at sqrt(r^2 + s^2)."""
    nb = P['nbins']
    out = {k: np.zeros(nb) for k in ('npairs', 'weight', 'meanr', 'meanlogr')}
    rsq = refine_rsq(s, r)
    fits, ik, logr = single_bin(rsq, s, P, logf)
    if fits:
        rr = math.sqrt(rsq)
        k = nb - 1 if ik == nb else ik
        out['npairs'][k] += 2
        out['weight'][k] += 2.
        out['meanr'][k] += 2. * rr
        out['meanlogr'][k] += 2. * logr
    else:
        rsq2 = r * r + s * s
        rr = math.sqrt(rsq2)
        lr = logf(rr)
        k = bin_index(lr, P)
        for _ in range(2):
            out['npairs'][k] += 1
            out['weight'][k] += 1.
            out['meanr'][k] += rr
            out['meanlogr'][k] += lr
    return out, fits
