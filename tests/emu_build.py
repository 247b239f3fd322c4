"""The engine's HIP sources built for the CPU emulation of HIP (tests/native/hip_emu):
g++ with AddressSanitizer, every source compiled once per test session, all of them
in parallel, and linked into the emulation drivers of tests/native on request."""
import atexit
import os
import shutil
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(ROOT, 'treecorr_amd', 'csrc')
SOURCES = ('tc_field.hip', 'tc_corr2.hip', 'tc_cost.hip', 'tc_corr3.hip', 'tc_patch.hip')
FLAGS = ['g++', '-std=c++17', '-O1', '-ffp-contract=off', '-fsanitize=address',
         '-fno-omit-frame-pointer']
INC = ['-I', os.path.join(HERE, 'native', 'hip_emu'), '-I', CSRC]

_objects = None


def objects():
    """Object files of every engine source, compiled on first use."""
    global _objects
    if _objects is None:
        out = tempfile.mkdtemp(prefix='tc_emu_objs_')
        atexit.register(shutil.rmtree, out, True)
        procs = []
        for src in SOURCES:
            o = os.path.join(out, src + '.o')
            procs.append((src, o, subprocess.Popen(FLAGS + ['-x', 'c++', *INC, '-c',
                                                            os.path.join(CSRC, src), '-o', o],
                                                   stderr=subprocess.PIPE)))
        for src, o, p in procs:
            _, e = p.communicate(timeout=900)
            assert p.returncode == 0, (src, e.decode()[-3000:])
        _objects = {src: o for src, o, _ in procs}
    return _objects


def link(driver, sources):
    """The emulation driver tests/native/<driver>.cpp linked with the given engine
    sources; returns (executable, its folder)."""
    objs = objects()
    out = tempfile.mkdtemp(prefix='tc_emu_%s_' % driver)
    exe = os.path.join(out, driver)
    r = subprocess.run(FLAGS + [*INC, os.path.join(HERE, 'native', driver + '.cpp'),
                                *[objs[s] for s in sources], '-o', exe, '-pthread'],
                       capture_output=True, timeout=300)
    assert r.returncode == 0, r.stderr.decode()[-3000:]
    return exe, out
