"""Writes the golden vectors under tests/golden/ from the literal reference engine.

Run in the build container (needs oracle/_ref, built by `make -f oracle/Makefile.ref`
from /root/reference/src):  python tests/golden/make_golden.py
Outputs per case: the generated-input checksum, the reference's raw sums
(xi0..xi3, meanr, meanlogr, weight, npairs; before finalize), and for tree cases
the reference's cells in pre-order.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, '..', '..'))
sys.path.insert(0, HERE)

from oracle import ref as R  # noqa: E402
import cases as C  # noqa: E402


def main():
    if R.module() is None or R.dumplib() is None:
        raise SystemExit('oracle/_ref is not built')
    for case in C.CORR2_CASES:
        name, d, coords = case[0], case[1], case[2]
        cat1, cat2, bn, extra = C.corr2_case(case)
        kw = dict(min_size=bn['min_size'], max_size=bn['max_size'], min_top=extra['min_top'],
                  brute=extra['brute'])

        def field(cat):
            return R.make_field(d, coords, cat['x'], cat['y'], cat['z'], k=cat['k'],
                                g1=cat['g1'], g2=cat['g2'], w=cat['w'], **kw)
        f1 = field(cat1)
        f2 = field(cat2) if cat2 is not None else None
        out = R.corr2(d, f1, f2, bn['nbins'], bn['min_sep'], bn['max_sep'], bn['bin_size'],
                      bn['b'], bn['a'])
        spec = dict(case=list(case), checksum1=C.checksum(cat1),
                    checksum2=C.checksum(cat2) if cat2 is not None else None,
                    ntop1=f1.getNTopLevel(), ntop2=f2.getNTopLevel() if f2 else None)
        np.savez_compressed(os.path.join(HERE, name + '.npz'), spec=json.dumps(spec), **out)
        print(name, 'sum npairs', out['npairs'].sum())
    for case in C.CORR3_CASES:
        name, d, coords = case[0], case[1], case[2]
        cat, bn, extra = C.corr3_case(case)
        f = R.make_field(d, coords, cat['x'], cat['y'], cat['z'], g1=cat['g1'], g2=cat['g2'],
                         w=cat['w'], min_size=bn['min_size'], max_size=bn['max_size'],
                         min_top=extra['min_top'])
        out = R.corr3_multipole(d, f, bn['nbins'], bn['min_sep'], bn['max_sep'],
                                bn['bin_size'], bn['b'], bn['a'], bn['max_n'])
        keep = ['weight', 'weight_im', 'ntri', 'meand2', 'meanlogd2', 'meand3', 'meanlogd3']
        if d == 4:
            keep += ['zeta%d' % i for i in range(8)]
        spec = dict(case=list(case), checksum=C.checksum(cat), ntop=f.getNTopLevel())
        np.savez_compressed(os.path.join(HERE, name + '.npz'), spec=json.dumps(spec),
                            **{k: out[k] for k in keep})
        print(name, 'sum ntri', out['ntri'][:, :, bn['max_n']].sum())
    for case in C.TREE_CASES:
        name, d, coords = case[0], case[1], case[2]
        cat, kw = C.tree_case(case)
        t = R.tree_dump(d, coords, cat['x'], cat['y'], cat['z'], k=cat['k'], g1=cat['g1'],
                        g2=cat['g2'], w=cat['w'], **kw)
        spec = dict(case=list(case), checksum=C.checksum(cat))
        np.savez_compressed(os.path.join(HERE, name + '.npz'), spec=json.dumps(spec), **t)
        print(name, 'nodes', len(t['n']), 'top', len(t['top']))


if __name__ == '__main__':
    main()
