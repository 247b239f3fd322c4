"""The multipole oracle against the literal reference's golden vectors and against
an O(N^3) restatement on a tiny catalog."""
import json
import math
import os

import numpy as np
import pytest

import cases as C

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')


def normwise(a, b):
    den = np.max(np.abs(b))
    return 0. if den == 0 else np.max(np.abs(a - b)) / den


@pytest.mark.parametrize('case', C.CORR3_CASES, ids=[c[0] for c in C.CORR3_CASES])
def test_oracle_multipole_golden(oracle_lib, case):
    O = oracle_lib
    z = np.load(os.path.join(GOLD, case[0] + '.npz'), allow_pickle=False)
    spec = json.loads(str(z['spec']))
    gold = {k: z[k] for k in z.files if k != 'spec'}
    cat, bn, extra = C.corr3_case(case)
    assert C.checksum(cat) == spec['checksum']
    d, coords = case[1], case[2]
    f = O.Field(d, coords, cat['x'], cat['y'], cat['z'], g1=cat['g1'], g2=cat['g2'], w=cat['w'],
                min_size=bn['min_size'], max_size=bn['max_size'], min_top=extra['min_top'])
    out = O.corr3_multipole(d, coords, f, bn['nbins'], bn['min_sep'], bn['max_sep'],
                            bn['bin_size'], bn['b'], bn['a'], bn['max_n'])
    np.testing.assert_array_equal(out['ntri'], gold['ntri'])
    for k, v in gold.items():
        if k != 'ntri':
            assert normwise(out[k], v) < 1e-12, k


def test_oracle_nnn_multipole_vs_triple_loop(oracle_lib):
    """With brute=True every Gn term is a single object, so the n=0 weight of the
    multipole NNN equals the triple loop sum_{i, j != k} w_i w_j w_k over triangles
    whose two sides from i fall in (k2, k3) -- the restatement of
    tests/test_nnn.py:4529-4562."""
    O = oracle_lib
    rng = np.random.default_rng(4)
    n = 60
    x = rng.uniform(0, 10, n); y = rng.uniform(0, 10, n); w = rng.uniform(0.5, 1.5, n)
    nbins, min_sep, max_sep, max_n = 4, 1., 8., 3
    bin_size = math.log(max_sep / min_sep) / nbins
    f = O.Field(0, 1, x, y, None, w=w, min_size=0., max_size=np.inf, brute=True, min_top=2)
    out = O.corr3_multipole(0, 1, f, nbins, min_sep, max_sep, bin_size, 0., 0.1, max_n)
    w32 = w.astype(np.float32).astype(float)
    weight = np.zeros((nbins, nbins, 2 * max_n + 1))
    ntri = np.zeros((nbins, nbins))
    for i in range(n):
        d = np.hypot(x - x[i], y - y[i])
        phi = np.arctan2(y - y[i], x - x[i])
        k = np.floor(np.log(np.where(d > 0, d, 1.) / min_sep) / bin_size).astype(int)
        ok = (d >= min_sep) & (d < max_sep)
        for j in np.where(ok)[0]:
            for kk in np.where(ok)[0]:
                if j == kk:
                    continue
                m = np.arange(-max_n, max_n + 1)
                weight[k[j], k[kk]] += (w32[i] * w32[j] * w32[kk] *
                                        np.cos(m * (phi[j] - phi[kk])))
                ntri[k[j], k[kk]] += 1
    np.testing.assert_array_equal(out['ntri'][:, :, max_n], ntri)
    assert normwise(out['weight'], weight) < 1e-10


CROSS_KINDS = [('cross12', 0), ('cross12', 1), ('cross21', 0), ('cross21', 3), ('cross', 0),
               ('cross', 1), ('cross', 2), ('cross', 3), ('cross', 4)]


@pytest.mark.parametrize('case', C.CORR3_CROSS_CASES, ids=[c[0] for c in C.CORR3_CROSS_CASES])
def test_oracle_multipole_cross_against_literal_reference(oracle_lib, case):
    """Multipole cross-correlations, every driver and ordered code (processCross12 0/1,
    processCross21 0/3, processCross 0-4; src/Corr3.cpp:3731-3985): the oracle's passes
    over two or three trees against the compiled reference's own drivers on the same
    inputs.  ntri exact, every other sum within 1e-12 normwise; the three fields differ,
    so a pass with the fields in the wrong roles or a dropped swap term fails."""
    from oracle import ref as R
    if R.module() is None:
        pytest.skip('oracle/_ref not built')
    O = oracle_lib
    d, coords = case[1], case[2]
    cats, bn, extra = C.corr3_cross_case(case, 3)
    kw = dict(min_size=bn['min_size'], max_size=bn['max_size'], min_top=extra['min_top'])
    args = (bn['nbins'], bn['min_sep'], bn['max_sep'], bn['bin_size'], bn['b'], bn['a'],
            bn['max_n'])

    def fld(mod, cat):
        return mod.Field(d, coords, cat['x'], cat['y'], cat['z'], g1=cat['g1'], g2=cat['g2'],
                         w=cat['w'], **kw) if mod is O else \
            R.make_field(d, coords, cat['x'], cat['y'], cat['z'], g1=cat['g1'], g2=cat['g2'],
                         w=cat['w'], **kw)
    of = [fld(O, c) for c in cats]
    rf = [fld(R, c) for c in cats]
    seen = []
    for kind, ordered in CROSS_KINDS:
        nf = 3 if kind == 'cross' else 2
        o = O.corr3_multipole_cross(d, coords, kind, ordered, of[:nf], *args)
        r = R.corr3_multipole_cross(d, kind, ordered, rf[:nf], *args)
        assert np.sum(r['ntri']) > 0
        np.testing.assert_array_equal(o['ntri'], r['ntri'], err_msg=f'{kind} {ordered}')
        for k, v in r.items():
            if k in o and k != 'ntri':
                assert normwise(o[k], v) < 1e-12, (kind, ordered, k)
        seen.append(r['weight'])
    # the orderings give different sums
    assert normwise(seen[0], seen[1]) > 1e-3 and normwise(seen[4], seen[8]) > 1e-3


def test_oracle_cross12_of_one_field_is_auto(oracle_lib):
    """processCross12 with ordered=1 on (f, f) is the auto-correlation's pass: the same
    sums as the one-field multipole."""
    O = oracle_lib
    case = C.CORR3_CROSS_CASES[1]
    cats, bn, extra = C.corr3_cross_case(case, 1)
    cat = cats[0]
    f = O.Field(case[1], case[2], cat['x'], cat['y'], cat['z'], g1=cat['g1'], g2=cat['g2'],
                w=cat['w'], min_size=bn['min_size'], max_size=bn['max_size'],
                min_top=extra['min_top'])
    args = (bn['nbins'], bn['min_sep'], bn['max_sep'], bn['bin_size'], bn['b'], bn['a'],
            bn['max_n'])
    a = O.corr3_multipole(case[1], case[2], f, *args, nthreads=1)
    x = O.corr3_multipole_cross(case[1], case[2], 'cross12', 1, [f, f], *args, nthreads=1)
    for k in O.Z3_KEYS:
        np.testing.assert_array_equal(a[k], x[k])
