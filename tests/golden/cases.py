"""Input generators for the golden-vector cases (shared by make_golden.py and the tests).

Every case is fully determined by its spec (numpy.random.default_rng(seed),
PCG64, stable across numpy versions); a checksum of the generated inputs is
stored with the fixture so drift is detected instead of silently compared.
"""
import hashlib
import math

import numpy as np

FLAT, THREED, SPHERE = 1, 2, 3
N_, K_, G_ = 0, 1, 4


def make_catalog(coords, n, seed, clumps=0.0, extent=None):
    """Uniform (+ optional Gaussian clumps) positions with w, k, g1, g2 columns."""
    rng = np.random.default_rng(seed)
    if coords == SPHERE:
        ext = extent or 0.1   # radians
        ra = rng.uniform(0, ext, n)
        dec = np.arcsin(rng.uniform(math.sin(-ext / 2), math.sin(ext / 2), n))
        nc = int(clumps * n)
        if nc:
            cen = rng.integers(0, n, 20)
            pick = rng.integers(0, 20, nc)
            ra[:nc] = ra[cen[pick]] + rng.normal(0, ext / 100, nc)
            dec[:nc] = dec[cen[pick]] + rng.normal(0, ext / 100, nc)
        x = np.cos(dec) * np.cos(ra)
        y = np.cos(dec) * np.sin(ra)
        z = np.sin(dec)
        scale = ext
    else:
        ext = extent or 100.
        x = rng.uniform(0, ext, n)
        y = rng.uniform(0, ext, n)
        z = rng.uniform(0, ext, n) if coords == THREED else None
        nc = int(clumps * n)
        if nc:
            cen = rng.integers(0, n, 20)
            pick = rng.integers(0, 20, nc)
            x[:nc] = x[cen[pick]] + rng.normal(0, ext / 100, nc)
            y[:nc] = y[cen[pick]] + rng.normal(0, ext / 100, nc)
            if z is not None:
                z[:nc] = z[cen[pick]] + rng.normal(0, ext / 100, nc)
        scale = ext
    w = rng.uniform(0.5, 1.5, n)
    k = rng.normal(0, 1, n)
    g1 = rng.normal(0, 0.2, n)
    g2 = rng.normal(0, 0.2, n)
    return dict(x=x, y=y, z=z, w=w, k=k, g1=g1, g2=g2), scale


def make_shell(n, seed, clumps=0.0):
    """3-d positions: a spherical catalog's directions at radii uniform in [50, 150]
    (for the Arc metric, whose separations are angles)."""
    cat, scale = make_catalog(SPHERE, n, seed, clumps)
    r = np.random.default_rng(seed + 1).uniform(50., 150., n)
    for key in ('x', 'y', 'z'):
        cat[key] = cat[key] * r
    return cat, scale


def checksum(cat):
    h = hashlib.sha256()
    for key in ('x', 'y', 'z', 'w', 'k', 'g1', 'g2'):
        if cat.get(key) is not None:
            h.update(np.ascontiguousarray(cat[key]).tobytes())
    return h.hexdigest()[:16]


def binning(scale, nbins, lo, hi, bin_slop, angle_slop):
    """Log binning + the field sizes of corr2base.py:1595-1625."""
    min_sep, max_sep = scale * lo, scale * hi
    bin_size = math.log(max_sep / min_sep) / nbins
    b = bin_size * bin_slop
    b1 = min(angle_slop, b)
    return dict(nbins=nbins, min_sep=min_sep, max_sep=max_sep, bin_size=bin_size, b=b,
                a=angle_slop, min_size=min_sep * b1 / (2 + 3 * b1), max_size=max_sep * b1)


# 2-point cases: (name, d, coords, n1, n2 (0 = auto), seed, nbins, lo, hi, bin_slop, angle_slop,
#                 brute, min_top, clumps)
CORR2_CASES = [
    ('nn_flat_bs0', N_, FLAT, 3000, 0, 11, 12, 0.005, 0.3, 0.0, 1.0, False, 3, 0.0),
    ('nn_flat_brute', N_, FLAT, 1500, 0, 12, 10, 0.01, 0.4, 0.0, 1.0, True, 3, 0.0),
    ('nn_3d_bs01', N_, THREED, 4000, 0, 13, 15, 0.01, 0.5, 0.1, 1.0, False, 3, 0.3),
    ('nn_sphere_cross', N_, SPHERE, 3000, 2500, 14, 10, 0.01, 0.5, 0.5, 1.0, False, 4, 0.0),
    ('kk_flat_bs1', K_, FLAT, 4000, 0, 15, 12, 0.005, 0.3, 1.0, 1.0, False, 3, 0.0),
    ('kk_3d_cross', K_, THREED, 3000, 3000, 16, 10, 0.01, 0.4, 0.05, 1.0, False, 3, 0.2),
    ('gg_flat_bs0', G_, FLAT, 3000, 0, 17, 12, 0.005, 0.3, 0.0, 0.1, False, 3, 0.0),
    ('gg_sphere_bs01', G_, SPHERE, 5000, 0, 18, 15, 0.01, 0.5, 0.1, 0.1, False, 4, 0.3),
    ('gg_3d_bs03', G_, THREED, 3000, 0, 19, 10, 0.01, 0.4, 0.3, 0.1, False, 3, 0.0),
    ('gg_sphere_cross', G_, SPHERE, 2500, 2500, 20, 10, 0.01, 0.5, 1.0, 0.1, False, 3, 0.2),
]

# Tree dumps: (name, d, coords, n, seed, min_size_frac, max_size_frac, split, brute, min_top)
TREE_CASES = [
    ('tree_g_flat', G_, FLAT, 1500, 21, 1e-4, 0.05, 2, False, 3),
    ('tree_g_sphere', G_, SPHERE, 1500, 22, 1e-4, 0.05, 2, False, 3),
    ('tree_k_3d', K_, THREED, 1500, 23, 1e-4, 0.05, 2, False, 3),
    ('tree_n_flat_median', N_, FLAT, 1000, 24, 1e-4, 0.05, 1, False, 2),
    ('tree_g_flat_middle', G_, FLAT, 1000, 25, 1e-4, 0.05, 0, False, 2),
    ('tree_n_sphere_brute', N_, SPHERE, 1000, 26, 0.0, 0.0, 2, True, 3),
]


def corr2_case(case):
    (name, d, coords, n1, n2, seed, nbins, lo, hi, bs, asl, brute, min_top, clumps) = case
    cat1, scale = make_catalog(coords, n1, seed, clumps)
    cat2 = make_catalog(coords, n2, seed + 1000, clumps)[0] if n2 else None
    bn = binning(scale, nbins, lo, hi, bs, asl)
    if brute:
        bn['min_size'] = 0.
    return cat1, cat2, bn, dict(brute=brute, min_top=min_top)


def tree_case(case):
    (name, d, coords, n, seed, fmin, fmax, sm, brute, min_top) = case
    cat, scale = make_catalog(coords, n, seed, 0.2)
    return cat, dict(min_size=scale * fmin, max_size=scale * fmax if fmax else np.inf,
                     split_method=sm, brute=brute, min_top=min_top, max_top=10)

# Multipole (LogMultipole) cases: (name, d, coords, n, seed, nbins, lo, hi, max_n, bin_slop,
#                                  min_top, clumps)
CORR3_CASES = [
    ('nnn_flat_mp', N_, FLAT, 2500, 31, 8, 0.02, 0.3, 8, None, 3, 0.2),
    ('ggg_flat_mp', G_, FLAT, 2500, 32, 8, 0.02, 0.3, 10, None, 3, 0.2),
    ('ggg_sphere_mp', G_, SPHERE, 2500, 33, 6, 0.02, 0.3, 12, None, 3, 0.2),
    ('ggg_3d_mp', G_, THREED, 2000, 34, 6, 0.02, 0.4, 6, 0.5, 3, 0.2),
]


# Multipole cross-correlations: the same binning rules as CORR3_CASES, with one catalog
# per field (seeds seed, seed+1, seed+2) over the same area.
CORR3_CROSS_CASES = [
    ('nnn_flat_x', N_, FLAT, 1200, 41, 6, 0.02, 0.3, 6, None, 3, 0.2),
    ('ggg_flat_x', G_, FLAT, 1200, 42, 6, 0.02, 0.3, 7, None, 3, 0.2),
    ('ggg_sphere_x', G_, SPHERE, 1200, 43, 5, 0.02, 0.3, 6, None, 3, 0.2),
    ('ggg_3d_x', G_, THREED, 1000, 44, 5, 0.02, 0.4, 5, 0.5, 3, 0.2),
]


def corr3_cross_case(case, nfields):
    cats = []
    for i in range(nfields):
        c = (case[0], case[1], case[2], case[3], case[4] + i) + tuple(case[5:])
        cat, bn, extra = corr3_case(c)
        cats.append(cat)
    return cats, bn, extra


def corr3_case(case):
    (name, d, coords, n, seed, nbins, lo, hi, max_n, bs, min_top, clumps) = case
    cat, scale = make_catalog(coords, n, seed, clumps)
    min_sep, max_sep = scale * lo, scale * hi
    bin_size = math.log(max_sep / min_sep) / nbins
    a = math.pi / (2 * max_n + 1)
    if bs is None:
        b = bin_size if bin_size <= 0.1 else 0.1
    else:
        b = bin_size * bs
    b1 = min(a, b)
    bn = dict(nbins=nbins, min_sep=min_sep, max_sep=max_sep, bin_size=bin_size, b=b, a=a,
              max_n=max_n, min_size=2 * min_sep * b1 / (2 + 3 * b1), max_size=max_sep * b1)
    return cat, bn, dict(min_top=min_top)
