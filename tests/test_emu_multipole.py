"""The device multipole kernels (treecorr_amd/csrc/tc_corr3.hip) run on the CPU.

tests/native/hip_emu emulates the HIP subset the kernels use (fibers for the
threads of a workgroup, barriers, cross-lane exchange, atomics) and the kernels are
compiled with g++ and AddressSanitizer, so every out-of-bounds access of their
indexing is reported with its source line.  The first tests take the tree from the
oracle; the last ones build it with the device tree builder itself (tc_field.hip under
the same emulation).  The results must match the reference's golden vectors: ntri
exactly, the FP64 sums to 1e-12 normwise.

Workgroups run one at a time (EMU_WORKERS=1) and eight at a time, which exercises
the persistent unit queue of k_mp_finish with concurrent workgroups.
"""
import json
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
GOLD = os.path.join(HERE, 'golden')
RTOL = 1e-12
KEYS = ['zeta%d' % i for i in range(8)] + ['meand2', 'meanlogd2', 'meand3', 'meanlogd3',
                                           'weight', 'weight_im', 'ntri']


def normwise(a, b):
    den = np.max(np.abs(b))
    return 0. if den == 0 else np.max(np.abs(a - b)) / den


@pytest.fixture(scope='module')
def emu_exe():
    if shutil.which('g++') is None:
        pytest.skip('g++ is not available')
    out = tempfile.mkdtemp(prefix='tc_emu_')
    exe = os.path.join(out, 'emu_corr3')
    cmd = ['g++', '-std=c++17', '-O1', '-g', '-fsanitize=address', '-fno-omit-frame-pointer',
           '-ffp-contract=off', '-DEMU_CORR3_MAIN', '-I', os.path.join(HERE, 'native', 'hip_emu'),
           '-I', os.path.join(ROOT, 'treecorr_amd', 'csrc'),
           os.path.join(HERE, 'native', 'emu_corr3.cpp'), '-o', exe, '-pthread']
    subprocess.run(cmd, check=True, capture_output=True, timeout=300)
    yield exe
    shutil.rmtree(out, ignore_errors=True)


def run_emu(exe, d, coords, tree, bn, workers):
    tmp = tempfile.mkdtemp(prefix='tc_emu_run_')
    try:
        fin, fout = os.path.join(tmp, 'in.bin'), os.path.join(tmp, 'out.bin')
        nn, ntop = len(tree['n']), len(tree['top'])
        with open(fin, 'wb') as fh:
            np.array([d, coords, nn, ntop, bn['nbins'], bn['max_n'], 0], np.int64).tofile(fh)
            np.array([bn['min_sep'], bn['max_sep'], bn['bin_size'], bn['b'], bn['a']],
                     np.float64).tofile(fh)
            np.ascontiguousarray(tree['pos'], np.float64).tofile(fh)
            for k in ('size', 'w'):
                np.ascontiguousarray(tree[k], np.float32).tofile(fh)
            np.ascontiguousarray(tree['g'], np.float32).tofile(fh)
            np.ascontiguousarray(tree['n'], np.int64).tofile(fh)
            for k in ('left', 'right', 'top'):
                np.ascontiguousarray(tree[k], np.int32).tofile(fh)
        env = dict(os.environ, EMU_WORKERS=str(workers), ASAN_OPTIONS='detect_leaks=0')
        r = subprocess.run([exe, fin, fout], env=env, capture_output=True, timeout=600)
        assert r.returncode == 0, r.stderr.decode()[-3000:]
        ntot = bn['nbins'] ** 2 * (2 * bn['max_n'] + 1)
        raw = np.fromfile(fout, np.float64, count=15 * ntot)
        cnt = np.fromfile(fout, np.uint64, offset=15 * ntot * 8)
        shape = (bn['nbins'], bn['nbins'], 2 * bn['max_n'] + 1)
        out = {k: raw[i * ntot:(i + 1) * ntot].reshape(shape) for i, k in enumerate(KEYS)}
        out['counters'] = dict(zip(('gn_calls', 'zeta_calls', 'zeta_rows', 'stages'),
                                   (int(c) for c in cnt)))
        return out
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


RUNS = [(c, 8) for c in C.CORR3_CASES] + [(C.CORR3_CASES[0], 1), (C.CORR3_CASES[3], 1)]


@pytest.mark.parametrize('case,workers', RUNS, ids=['%s-w%d' % (c[0], w) for c, w in RUNS])
def test_emulated_multipole_matches_reference_golden(emu_exe, oracle_lib, case, workers):
    O = oracle_lib
    z = np.load(os.path.join(GOLD, case[0] + '.npz'), allow_pickle=False)
    spec = json.loads(str(z['spec']))
    cat, bn, extra = C.corr3_case(case)
    assert C.checksum(cat) == spec['checksum']
    d, coords = case[1], case[2]
    f = O.Field(d, coords, cat['x'], cat['y'], cat['z'], k=cat.get('k'), g1=cat.get('g1'),
                g2=cat.get('g2'), w=cat['w'], min_size=bn['min_size'], max_size=bn['max_size'],
                min_top=extra['min_top'])
    assert f.ntop == spec['ntop']
    out = run_emu(emu_exe, d, coords, f.dump(), bn, workers)
    np.testing.assert_array_equal(out['ntri'], z['ntri'])
    for k in z.files:
        if k not in ('spec', 'ntri') and z[k].size:
            assert normwise(out[k], z[k]) < RTOL, k
    ref = O.corr3_multipole(d, coords, f, bn['nbins'], bn['min_sep'], bn['max_sep'],
                            bn['bin_size'], bn['b'], bn['a'], bn['max_n'])
    assert out['counters'] == ref['counters']


# ---------------------------------------------------------------------------
# The whole library on the CPU: the device tree builder (tc_field.hip) writes the node
# array, child links, counts and top list that the multipole kernels (tc_corr3.hip) then
# read, as on the GPU.  Fibers run in a shuffled order between barriers (EMU_SHUFFLE),
# so a value one thread writes and another reads without a barrier shows up.
# ---------------------------------------------------------------------------
@pytest.fixture(scope='module')
def emu_lib_exe():
    if shutil.which('g++') is None:
        pytest.skip('g++ is not available')
    exe, out = emu_build.link('emu_lib3', ('tc_field.hip', 'tc_corr3.hip'))
    yield exe
    shutil.rmtree(out, ignore_errors=True)


def run_emu_lib(exe, d, coords, cat, bn, min_top, env):
    tmp = tempfile.mkdtemp(prefix='tc_emu_lib_run_')
    try:
        fin, fout = os.path.join(tmp, 'in.bin'), os.path.join(tmp, 'out.bin')
        n = len(cat['x'])
        with open(fin, 'wb') as fh:
            np.array([d, coords, n, bn['nbins'], bn['max_n'], 0, min_top, 10], np.int64).tofile(fh)
            np.array([bn['min_sep'], bn['max_sep'], bn['bin_size'], bn['b'], bn['a'],
                      bn['min_size'], bn['max_size']], np.float64).tofile(fh)
            for k in ('x', 'y', 'z', 'g1', 'g2', 'w'):
                v = cat.get(k)
                (np.zeros(n) if v is None else np.ascontiguousarray(v, np.float64)).tofile(fh)
        r = subprocess.run([exe, fin, fout], env=dict(os.environ, ASAN_OPTIONS='detect_leaks=0', **env),
                           capture_output=True, timeout=600)
        assert r.returncode == 0, r.stderr.decode()[-3000:]
        ntot = bn['nbins'] ** 2 * (2 * bn['max_n'] + 1)
        raw = open(fout, 'rb').read()
        sums = np.frombuffer(raw, np.float64, count=15 * ntot)
        o = 15 * ntot * 8
        cnt = np.frombuffer(raw, np.uint64, count=4, offset=o)
        nn, nt = np.frombuffer(raw, np.int64, count=2, offset=o + 32)
        from treecorr_amd.field import NODE_DTYPE
        nodes = np.frombuffer(raw, NODE_DTYPE, count=int(nn), offset=o + 48)
        top = np.frombuffer(raw, np.int32, count=int(nt), offset=o + 48 + 48 * int(nn))
        shape = (bn['nbins'], bn['nbins'], 2 * bn['max_n'] + 1)
        out = {k: sums[i * ntot:(i + 1) * ntot].reshape(shape) for i, k in enumerate(KEYS)}
        out['counters'] = dict(zip(('gn_calls', 'zeta_calls', 'zeta_rows', 'stages'),
                                   (int(c) for c in cnt)))
        return out, nodes, top
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


LIB_RUNS = [(c, {'EMU_WORKERS': '4', 'EMU_SHUFFLE': str(11 + i)}) for i, c in enumerate(C.CORR3_CASES)]
# the tree's top levels partitioned speculatively (every level above min_top), validated
# by rounding bounds, their exact pass finished before the tree is read
LIB_RUNS += [(C.CORR3_CASES[1], {'EMU_WORKERS': '4', 'EMU_SHUFFLE': '41', 'EMU_TREE_SPEC': '3'}),
             (C.CORR3_CASES[2], {'EMU_WORKERS': '4', 'EMU_SHUFFLE': '43', 'EMU_TREE_SPEC': '3'})]


@pytest.mark.parametrize('case,env', LIB_RUNS,
                         ids=[c[0] + ('-spec' if 'EMU_TREE_SPEC' in e else '') for c, e in LIB_RUNS])
def test_emulated_library_multipole_on_device_tree(emu_lib_exe, oracle_lib, case, env):
    """Device tree build + multipole, all under the emulation: the device tree equals the
    oracle's node for node, ntri equals the reference's golden vector, the sums agree to
    1e-12 and the Gn/zeta/stage counters equal the oracle's."""
    from treecorr_amd.field import preorder
    O = oracle_lib
    z = np.load(os.path.join(GOLD, case[0] + '.npz'), allow_pickle=False)
    spec = json.loads(str(z['spec']))
    cat, bn, extra = C.corr3_case(case)
    assert C.checksum(cat) == spec['checksum']
    d, coords = case[1], case[2]
    out, nodes, top = run_emu_lib(emu_lib_exe, d, coords, cat, bn, extra['min_top'], env)
    f = O.Field(d, coords, cat['x'], cat['y'], cat['z'], g1=cat.get('g1'), g2=cat.get('g2'),
                w=cat['w'], min_size=bn['min_size'], max_size=bn['max_size'],
                min_top=extra['min_top'])
    ref_tree = f.dump()
    nd = nodes[preorder(nodes, top)]
    assert len(top) == spec['ntop'] == len(ref_tree['top'])
    np.testing.assert_array_equal(nd['x'], ref_tree['pos'][:, 0])
    np.testing.assert_array_equal(nd['y'], ref_tree['pos'][:, 1])
    np.testing.assert_array_equal(nd['size'], ref_tree['size'])
    np.testing.assert_array_equal(nd['w'], ref_tree['w'])
    np.testing.assert_array_equal(nd['n'], ref_tree['n'])
    if d == 4:
        np.testing.assert_array_equal(nd['a'], ref_tree['g'][:, 0])
        np.testing.assert_array_equal(nd['b'], ref_tree['g'][:, 1])
    np.testing.assert_array_equal(out['ntri'], z['ntri'])
    for k in z.files:
        if k not in ('spec', 'ntri') and z[k].size:
            assert normwise(out[k], z[k]) < RTOL, k
    ref = O.corr3_multipole(d, coords, f, bn['nbins'], bn['min_sep'], bn['max_sep'],
                            bn['bin_size'], bn['b'], bn['a'], bn['max_n'])
    assert out['counters'] == ref['counters']


def run_emu_cross(exe, d, coords, cats, bn, min_top, passes, env):
    tmp = tempfile.mkdtemp(prefix='tc_emu_x_')
    try:
        fin, fout = os.path.join(tmp, 'in.bin'), os.path.join(tmp, 'out.bin')
        with open(fin, 'wb') as fh:
            np.array([-1, d, coords, len(cats), bn['nbins'], bn['max_n'], min_top, 10,
                      len(passes)], np.int64).tofile(fh)
            np.array([bn['min_sep'], bn['max_sep'], bn['bin_size'], bn['b'], bn['a'],
                      bn['min_size'], bn['max_size']], np.float64).tofile(fh)
            for names, order in passes:
                third = names[2] - 1 if len(names) == 3 else -1
                np.array([names[0] - 1, names[1] - 1, third, 0 if order is None else order],
                         np.int64).tofile(fh)
            for cat in cats:
                n = len(cat['x'])
                np.array([n], np.int64).tofile(fh)
                for k in ('x', 'y', 'z', 'g1', 'g2', 'w'):
                    v = cat.get(k)
                    (np.zeros(n) if v is None else np.ascontiguousarray(v, np.float64)).tofile(fh)
        r = subprocess.run([exe, fin, fout], env=dict(os.environ, ASAN_OPTIONS='detect_leaks=0',
                                                      **env), capture_output=True, timeout=900)
        assert r.returncode == 0, r.stderr.decode()[-3000:]
        ntot = bn['nbins'] ** 2 * (2 * bn['max_n'] + 1)
        raw = open(fout, 'rb').read()
        sums = np.frombuffer(raw, np.float64, count=15 * ntot)
        cnt = np.frombuffer(raw, np.uint64, count=4, offset=15 * ntot * 8)
        shape = (bn['nbins'], bn['nbins'], 2 * bn['max_n'] + 1)
        out = {k: sums[i * ntot:(i + 1) * ntot].reshape(shape) for i, k in enumerate(KEYS)}
        out['counters'] = dict(zip(('gn_calls', 'zeta_calls', 'zeta_rows', 'stages'),
                                   (int(c) for c in cnt)))
        return out
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


X_RUNS = [(C.CORR3_CROSS_CASES[0], 'cross12', 0), (C.CORR3_CROSS_CASES[2], 'cross', 2),
          (C.CORR3_CROSS_CASES[3], 'cross21', 0)]


@pytest.mark.parametrize('case,kind,ordered', X_RUNS, ids=['%s-%s%d' % (c[0], k, o)
                                                          for c, k, o in X_RUNS])
def test_emulated_library_multipole_cross(emu_lib_exe, oracle_lib, case, kind, ordered):
    """The cross-correlation passes (one list over another field's tree, and two lists
    with ordered 1 and 4) under the emulation with AddressSanitizer and shuffled thread
    order, on the device builder's trees: the same sums and counters as the oracle."""
    from oracle.oracle import MULTIPOLE_PASSES
    O = oracle_lib
    d, coords = case[1], case[2]
    cats, bn, extra = C.corr3_cross_case(case, 3)
    cats = [{k: (v[:500] if v is not None else None) for k, v in c.items()} for c in cats]
    nf = 3 if kind == 'cross' else 2
    passes = MULTIPOLE_PASSES[(kind, ordered)]
    out = run_emu_cross(emu_lib_exe, d, coords, cats[:nf], bn, extra['min_top'], passes,
                        {'EMU_WORKERS': '4', 'EMU_SHUFFLE': '29'})
    kw = dict(min_size=bn['min_size'], max_size=bn['max_size'], min_top=extra['min_top'])
    of = [O.Field(d, coords, c['x'], c['y'], c['z'], g1=c.get('g1'), g2=c.get('g2'), w=c['w'],
                  **kw) for c in cats[:nf]]
    ref = O.corr3_multipole_cross(d, coords, kind, ordered, of, bn['nbins'], bn['min_sep'],
                                  bn['max_sep'], bn['bin_size'], bn['b'], bn['a'], bn['max_n'])
    assert np.sum(ref['ntri']) > 0
    np.testing.assert_array_equal(out['ntri'], ref['ntri'])
    for k in KEYS:
        if k != 'ntri' and (d == 4 or not k.startswith('zeta')):
            assert normwise(out[k], ref[k]) < RTOL, k
    assert out['counters'] == ref['counters']
