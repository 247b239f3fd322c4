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
