"""The 2-point traversal (treecorr_amd/csrc/tc_corr2.hip) on the CPU, on the device
tree builder's own trees (tc_field.hip), through the HIP emulation in tests/native/hip_emu
(fibers for the threads of a workgroup, barriers, cross-lane exchange, atomics), compiled
with g++ and AddressSanitizer so that every out-of-bounds access reports its source line.

Every traversal mode runs: chunked launches driven from the host, the persistent launch
taking jobs by a shared counter or from static equal-count shares (from the initial node
pairs, or after the chunked split), and the persistent launch whose stacks are too small
(the fallback to chunked launches).  Workgroups run eight at a time with the threads of
each in a shuffled order between barriers.  npairs and the visit counters must equal the
oracle's, the FP64 sums agree to 1e-12 normwise.
"""
import os
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

import cases as C
import emu_build

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
RTOL = 1e-12
KEYS = ['xi0', 'xi1', 'xi2', 'xi3', 'meanr', 'meanlogr', 'weight']


def normwise(a, b):
    den = np.max(np.abs(b))
    return 0. if den == 0 else np.max(np.abs(a - b)) / den


@pytest.fixture(scope='module')
def emu_lib2_exe():
    if shutil.which('g++') is None:
        pytest.skip('g++ is not available')
    exe, out = emu_build.link('emu_lib2', ('tc_field.hip', 'tc_corr2.hip', 'tc_cost.hip'))
    yield exe
    shutil.rmtree(out, ignore_errors=True)


def run_emu(exe, d, coords, cat1, cat2, bn, mode, levels, min_jobs, min_top, chunk, env):
    tmp = tempfile.mkdtemp(prefix='tc_emu2_run_')
    try:
        fin, fout = os.path.join(tmp, 'in.bin'), os.path.join(tmp, 'out.bin')
        n1, n2 = len(cat1['x']), (0 if cat2 is None else len(cat2['x']))
        with open(fin, 'wb') as fh:
            np.array([d, coords, n1, n2, bn['nbins'], mode, levels, min_jobs, min_top, chunk],
                     np.int64).tofile(fh)
            np.array([bn['min_sep'], bn['max_sep'], bn['bin_size'], bn['b'], bn['a'],
                      bn['min_size'], bn['max_size']], np.float64).tofile(fh)
            for cat in (cat1, cat2):
                if cat is None:
                    continue
                for k in ('x', 'y', 'z', 'g1', 'g2', 'w'):
                    v = cat.get(k)
                    (np.zeros(len(cat['x'])) if v is None else
                     np.ascontiguousarray(v, np.float64)).tofile(fh)
        r = subprocess.run([exe, fin, fout], env=dict(os.environ, ASAN_OPTIONS='detect_leaks=0', **env),
                           capture_output=True, timeout=900)
        assert r.returncode == 0, r.stderr.decode()[-3000:]
        raw = open(fout, 'rb').read()
        nb = bn['nbins']
        sums = np.frombuffer(raw, np.float64, count=7 * nb)
        npairs = np.frombuffer(raw, np.uint64, count=nb, offset=56 * nb)
        cnt = np.frombuffer(raw, np.uint64, count=4, offset=64 * nb)
        fb = int(np.frombuffer(raw, np.int64, count=1, offset=64 * nb + 32)[0])
        out = {k: sums[i * nb:(i + 1) * nb] for i, k in enumerate(KEYS)}
        out['npairs'] = npairs
        out['counters'] = dict(zip(('visits', 'accepted', 'leafpairs', 'refined'),
                                   (int(c) for c in cnt)))
        out['fallbacks'] = fb
        return out
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


# (mode, stack levels, min_jobs, chunk): 0 chunked, 1 persistent by a shared counter,
# 2 persistent from static shares; levels 1 makes the persistent stacks overflow
MODES = [(0, 0, 0, 3000), (1, 0, 1, 0), (1, 0, 200, 4096), (2, 0, 1, 0), (1, 1, 1, 4096),
         (2, 1, 1, 4096)]


RUNS = [(m, False) for m in MODES] + [(MODES[1], True), (MODES[5], True)]
IDS = ['chunked', 'dyn', 'dyn-split', 'static', 'dyn-overflow', 'static-overflow',
       'dyn-cross', 'static-overflow-cross']


@pytest.mark.parametrize('m,cross', RUNS, ids=IDS)
def test_emulated_traversal_modes_match_oracle(emu_lib2_exe, oracle_lib, m, cross):
    O = oracle_lib
    mode, levels, min_jobs, chunk = m
    cat, scale = C.make_catalog(C.SPHERE, 1500, 311, 0.3)
    cat2 = C.make_catalog(C.SPHERE, 1200, 312, 0.3)[0] if cross else None
    bn = C.binning(scale, 12, 0.005, 0.3, 0.1, 0.1)
    env = {'EMU_WORKERS': '8', 'EMU_SHUFFLE': str(40 + mode + 3 * levels)}
    out = run_emu(emu_lib2_exe, 4, 3, cat, cat2, bn, mode, levels, min_jobs, 3, chunk, env)
    kw = dict(min_size=bn['min_size'], max_size=bn['max_size'], min_top=3)
    f = O.Field(4, 3, cat['x'], cat['y'], cat['z'], g1=cat['g1'], g2=cat['g2'], w=cat['w'], **kw)
    f2 = None if cat2 is None else O.Field(4, 3, cat2['x'], cat2['y'], cat2['z'], g1=cat2['g1'],
                                           g2=cat2['g2'], w=cat2['w'], **kw)
    ref = O.corr2(4, 3, f, f2, bn['nbins'], bn['min_sep'], bn['max_sep'], bn['bin_size'],
                  bn['b'], bn['a'], exact=True)
    np.testing.assert_array_equal(out['npairs'], ref['npairs'])
    for k in KEYS:
        assert normwise(out[k], ref[k]) < RTOL, k
    assert out['counters'] == ref['counters']
    assert out['fallbacks'] == (1 if levels and mode else 0)
