"""The oracle's QuickAssign restatement against the compiled reference.

oracle.quick_assign restates QuickAssign (src/KMeans.cpp:890-942); the reference's
own compiled module (oracle/_ref, built from its sources by oracle/Makefile.ref)
labels the same objects.  Skipped when that build is absent.
"""
import numpy as np
import pytest


@pytest.mark.parametrize('dim', [2, 3])
def test_quick_assign_restatement_matches_compiled_reference(oracle_lib, dim):
    from oracle import ref
    m = ref.module()
    if m is None:
        pytest.skip('the compiled reference (oracle/_ref) is not built')
    rng = np.random.default_rng(11 + dim)
    n, npatch = 200_000, 77
    x, y, z = rng.random(n), rng.random(n), rng.random(n)
    x[:1000] = np.round(x[:1000], 2)       # many equal coordinates
    cen = rng.random((npatch, dim))
    cen[40] = cen[9]                       # a tie between two centres
    got = oracle_lib.quick_assign(cen, x, y, z if dim == 3 else None)
    want = np.empty(n, dtype=np.int64)
    m.QuickAssign(np.ascontiguousarray(cen.ravel()), npatch, x, y,
                  z if dim == 3 else np.array([]), want)
    np.testing.assert_array_equal(got, want)
    assert not np.any(got == 40)


@pytest.mark.parametrize('dim', [2, 3])
def test_select_patch_restatement_matches_compiled_reference(oracle_lib, dim):
    """oracle.select_patch (SelectPatch, src/KMeans.cpp:944-1011) against the compiled
    reference, including objects equidistant from the selected centre and another one."""
    from oracle import ref
    m = ref.module()
    if m is None:
        pytest.skip('the compiled reference (oracle/_ref) is not built')
    rng = np.random.default_rng(21 + dim)
    n, npatch = 100_000, 33
    x, y, z = rng.random(n), rng.random(n), rng.random(n)
    cen = rng.random((npatch, dim))
    cen[20] = cen[5]                       # two equal centres: both select the same objects
    if dim == 2:
        # objects on the bisector of centres 5 and 6 (equidistant up to rounding)
        mid = 0.5 * (cen[5] + cen[6])
        x[:500], y[:500] = mid[0], mid[1]
    for patch in (0, 5, 20, 32):
        got = oracle_lib.select_patch(patch, cen, x, y, z if dim == 3 else None)
        want = np.zeros(n, dtype=np.int64)
        m.SelectPatch(patch, np.ascontiguousarray(cen.ravel()), npatch, x, y,
                      z if dim == 3 else np.array([]), want)
        np.testing.assert_array_equal(got, want)
        assert got.sum() > 0
    a = oracle_lib.select_patch(5, cen, x, y, z if dim == 3 else None)
    b = oracle_lib.select_patch(20, cen, x, y, z if dim == 3 else None)
    np.testing.assert_array_equal(a, b)       # ties keep the object for both
