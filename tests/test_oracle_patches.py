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
