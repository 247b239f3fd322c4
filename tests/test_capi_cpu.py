"""The C-ABI library builds, loads and exports every entry point include/treecorr_amd.h
declares; host-side validation works without a GPU."""
import ctypes
import math
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, 'include', 'treecorr_amd.h')).read()
    text = re.sub(r'/\*.*?\*/', '', text, flags=re.S)
    return sorted(set(re.findall(r'\b(tc_[a-z0-9_]+)\s*\(', text)))


def test_library_exports_every_declared_symbol():
    from treecorr_amd import _lib
    L = _lib.lib()
    names = declared()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # every declared entry point has a ctypes signature in the binding
    assert set(names) <= set(_lib.SIGNATURES), set(names) - set(_lib.SIGNATURES)


def test_version_and_errors_without_gpu():
    from treecorr_amd import _lib
    L = _lib.lib()
    assert L.tc_version() >= 100
    h = ctypes.c_void_p()
    x = np.zeros(3)
    # invalid data type is rejected before any device work
    rc = L.tc_field_create(7, 1, _lib.dptr(x), _lib.dptr(x), None, None, None, _lib.dptr(x), None,
                           3, 0., 1., 2, 0, 0, 3, 10, ctypes.byref(h))
    assert rc != 0
    assert len(L.tc_last_error()) > 0


def test_unsupported_corr2_reports_not_implemented():
    """KN and GN have no reference class (treecorr has NK, NG, KG only); a missing
    xi array for NG is an invalid argument."""
    from treecorr_amd import _lib
    L = _lib.lib()
    h = ctypes.c_void_p()
    a = np.zeros(5)
    for d1, d2 in ((1, 0), (4, 0), (4, 1)):
        rc = L.tc_corr2_create(d1, d2, 0, 1., 10., 5, math.log(10) / 5, 0.1, 0.1, 0., 0., 0., 0.,
                               0., _lib.dptr(a), _lib.dptr(a), None, None, *[_lib.dptr(a)] * 4,
                               ctypes.byref(h))
        assert rc == _lib.TC_ERR_UNSUPPORTED, (d1, d2)
    rc = L.tc_corr2_create(0, 4, 0, 1., 10., 5, math.log(10) / 5, 0.1, 0.1, 0., 0., 0., 0., 0.,
                           _lib.dptr(a), None, None, None, *[_lib.dptr(a)] * 4, ctypes.byref(h))
    assert rc == _lib.TC_ERR_INVALID


def test_generate_xyz_matches_sincos():
    """GenerateXYZ (src/KMeans.cpp:1013-1050): x = cos(dec)cos(ra), y = cos(dec)sin(ra), z = sin(dec)."""
    from treecorr_amd import _lib
    rng = np.random.default_rng(1)
    ra = rng.uniform(0, 2 * np.pi, 1000)
    dec = rng.uniform(-1.5, 1.5, 1000)
    x, y, z = np.empty(1000), np.empty(1000), np.empty(1000)
    _lib.check(_lib.lib().tc_generate_xyz(_lib.dptr(x), _lib.dptr(y), _lib.dptr(z), _lib.dptr(ra),
                                          _lib.dptr(dec), None, 1000))
    np.testing.assert_allclose(x, np.cos(dec) * np.cos(ra), rtol=0, atol=1e-15)
    np.testing.assert_allclose(z, np.sin(dec), rtol=0, atol=1e-15)


def test_binning_setup_matches_treecorr_rules():
    """corr2base.py:420-449 and _get_minmax_size :1595-1625."""
    import treecorr_amd as tca
    gg = tca.GGCorrelation(min_sep=1., max_sep=200., nbins=25, sep_units='arcmin', bin_slop=0.1)
    assert gg.bin_size == pytest.approx(math.log(200.) / 25, rel=1e-15)
    assert gg.b == pytest.approx(0.1 * gg.bin_size, rel=1e-15)
    assert gg.angle_slop == 0.1
    mn, mx = gg._get_minmax_size()
    b1 = min(0.1, gg.b)
    assert mn == pytest.approx(gg._min_sep * b1 / (2 + 3 * b1), rel=1e-15)
    assert mx == pytest.approx(gg._max_sep * b1, rel=1e-15)
    nn = tca.NNCorrelation(min_sep=1., max_sep=100., bin_size=0.5)
    assert nn.nbins == int(math.ceil(math.log(100.) / 0.5))
    assert nn.angle_slop == 1
    # default bin_slop: min(0.1/bin_size, 1)
    assert nn.bin_slop == pytest.approx(min(0.1 / nn.bin_size, 1.0))
    with pytest.raises(TypeError):
        tca.NNCorrelation(min_sep=1., bin_size=0.1)
    with pytest.raises(ValueError):
        tca.NNCorrelation(min_sep=10., max_sep=1., nbins=3)


def test_patch_split_is_a_stable_split_by_patch():
    """Catalog.get_patches through tc_patch_split (host threads): every patch catalog
    holds exactly the objects of np.where(patch == i), in their original order, as
    TreeCorr's get_patches builds them (catalog.py:3017-3048); empty patches are left
    out; a label outside [0, npatch) is an error."""
    import treecorr_amd as tc
    from treecorr_amd import _lib
    rng = np.random.default_rng(5)
    n = 300_000    # several host threads
    x, y, z, w, k = (rng.random(n) for _ in range(5))
    lab = rng.integers(0, 41, n)
    lab[lab == 7] = 8      # patch 7 empty
    cat = tc.Catalog(x=x, y=y, z=z, w=w, k=k, patch=lab)
    ps = cat.get_patches()
    assert len(ps) == 40 and cat.get_patches() is ps
    assert [p._single_patch for p in ps] == [i for i in range(41) if i != 7]
    for p in ps:
        sel = np.where(lab == p._single_patch)[0]
        for nm in ('x', 'y', 'z', 'w', 'k'):
            np.testing.assert_array_equal(getattr(p, nm), getattr(cat, nm)[sel])
        # Catalog.patch is the single patch number (catalog.py:1200-1205)
        assert p.npatch == 41 and p.patch == p._single_patch
        assert math.isclose(p.sumw, float(np.sum(w[sel])), rel_tol=1e-12)
    bad = np.array([0, 1, 5], dtype=np.int32)
    col = np.zeros(3)
    out = np.zeros(3)
    dp = ctypes.POINTER(ctypes.c_double)
    counts = np.zeros(3, dtype=np.int64)
    rc = _lib.lib().tc_patch_split(bad.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), 3, 3, 1,
                                   (dp * 1)(col.ctypes.data_as(dp)),
                                   (dp * 1)(out.ctypes.data_as(dp)),
                                   counts.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)))
    assert rc == _lib.TC_ERR_INVALID


def test_inverted_rpar_range_is_rejected():
    """corr2base.py:545-546: min_rpar > max_rpar raises ValueError (host-side check)."""
    import pytest
    import treecorr_amd as tca
    with pytest.raises(ValueError, match='min_rpar'):
        tca.NNCorrelation(nbins=5, min_sep=1., max_sep=10., min_rpar=5., max_rpar=-5.)


def test_corr3_trivially_zero_is_conservative():
    """TriviallyZero for 3-point patches (src/Corr3.cpp:275-299): whenever it says no
    triangle can land in a bin, points drawn anywhere inside the three cells give none
    (with the vertex order fixed for ordered=1, in any order for ordered=0); it is not
    vacuous; and with cell 2 the same as cell 3 only the 1-3 distance decides."""
    import ctypes as ct
    import treecorr_amd as tc
    from treecorr_amd import _lib
    L = _lib.lib()
    corr = tc.NNNCorrelation(min_sep=1., max_sep=4., nbins=4, max_n=5,
                             bin_type='LogMultipole')
    h = corr.corr
    rng = np.random.default_rng(5)

    def tz(cells, ordered, p13):
        r = ct.c_int()
        args = []
        for (x, y, s) in cells:
            args += [x, y, 0., s]
        assert L.tc_corr3_trivially_zero(h, 1, *args, ordered, p13, ct.byref(r)) == 0
        return r.value

    def draw(x, y, s, n=300):
        rr = s * np.sqrt(rng.uniform(0, 1, n))
        t = rng.uniform(0, 2 * np.pi, n)
        return np.stack([x + rr * np.cos(t), y + rr * np.sin(t)], axis=1)

    def ok(d):
        return (d >= 1.) & (d < 4.)

    hits = {0: 0, 1: 0}
    for trial in range(400):
        cells = [(rng.uniform(0, 8), rng.uniform(0, 8), rng.choice([0., rng.uniform(0, 1.5)]))
                 for _ in range(3)]
        for ordered in (0, 1):
            if not tz(cells, ordered, 0):
                continue
            hits[ordered] += 1
            p1, p2, p3 = [draw(*c) for c in cells]
            d1 = np.linalg.norm(p2 - p3, axis=1)
            d2 = np.linalg.norm(p1 - p3, axis=1)
            d3 = np.linalg.norm(p1 - p2, axis=1)
            if ordered:
                assert not np.any(ok(d2) & ok(d3)), cells
            else:
                # some vertex as the first with the two sides from it in range
                for a, b in ((d2, d3), (d1, d3), (d1, d2)):
                    assert not np.any(ok(a) & ok(b)), cells
    assert hits[0] > 20 and hits[1] > hits[0]
    # cells whose every distance is in range are never skipped
    assert tz([(0., 0., 0.1), (2., 0., 0.1), (0., 2., 0.1)], 1, 0) == 0
    assert tz([(0., 0., 0.1), (2., 0., 0.1), (0., 2., 0.1)], 0, 0) == 0
    # p13: only |c1 - c3| against max_sep + s1 + s3
    far = [(0., 0., 0.5), (0., 0., 0.5), (5.1, 0., 0.5)]
    assert tz(far, 1, 1) == 1
    assert tz([(0., 0., 0.5), (0., 0., 0.5), (4.9, 0., 0.5)], 1, 1) == 0


def test_corr3_trivially_zero_matches_compiled_reference():
    """tc_corr3_trivially_zero against the compiled reference's own triviallyZero
    (src/Corr3.cpp:4069-4130) on random cell triples: the same decision every time,
    flat, 3-d and spherical positions, ordered 0/1, with and without p13."""
    import ctypes as ct
    import math as _m
    import treecorr_amd as tc
    from oracle import ref as R
    if R.module() is None:
        pytest.skip('oracle/_ref not built')
    from treecorr_amd import _lib
    L = _lib.lib()
    min_sep, max_sep, nbins, max_n = 1., 4., 4, 5
    bin_size = _m.log(max_sep / min_sep) / nbins
    corr = tc.NNNCorrelation(min_sep=min_sep, max_sep=max_sep, nbins=nbins, max_n=max_n,
                             bin_type='LogMultipole')
    m, rc, keep = R._corr3_mp(0, nbins, min_sep, max_sep, bin_size, corr.b, corr.angle_slop,
                              max_n, None)
    rng = np.random.default_rng(11)
    coords = {1: m.Flat, 2: m.ThreeD, 3: m.Sphere}
    seen = set()
    for trial in range(3000):
        c = int(rng.integers(1, 4))
        cells = []
        for _ in range(3):
            if c == 3:
                # near (1, 0, 0), off the unit sphere by up to 10% (the reference
                # normalises spherical positions)
                x, y, z = 1 + rng.normal(0, 0.25), rng.normal(0, 0.25), rng.normal(0, 0.25)
                s = rng.choice([0., rng.uniform(0, 0.3)])
            else:
                x, y = rng.uniform(0, 8, 2)
                z = rng.uniform(0, 8) if c == 2 else 0.
                s = rng.choice([0., rng.uniform(0, 1.5)])
            cells += [float(x), float(y), float(z), float(s)]
        ordered, p13 = int(rng.integers(0, 2)), int(rng.integers(0, 2))
        r = ct.c_int()
        assert L.tc_corr3_trivially_zero(corr.corr, c, *cells, ordered, p13, ct.byref(r)) == 0
        want = rc.triviallyZero(m.Euclidean, coords[c], *cells, ordered, bool(p13))
        assert bool(r.value) == bool(want), (c, cells, ordered, p13)
        seen.add((c, bool(want)))
    assert len(seen) == 6        # both answers in every coordinate system
