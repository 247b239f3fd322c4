"""Build recipe for the treecorr_amd HIP library (gfx950).

Compiles treecorr_amd/csrc/*.hip into treecorr_amd/libtreecorr_amd.so with hipcc.
Run as ``python -m treecorr_amd.build`` (or through __graft_entry__.build()).
"""
import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
LIB = os.path.join(HERE, 'libtreecorr_amd.so')
ARCH = os.environ.get('TC_OFFLOAD_ARCH', 'gfx950')

# -ffp-contract=off: no FMA contraction, so FP64 results follow the reference's
# (non-FMA) evaluation bit for bit where decisions depend on them.
# -packed-fp32-ops off: avoids packed FP32 instructions in generated code.
HIPFLAGS = ['-O3', '-std=c++17', '-fPIC', '-ffp-contract=off', '-fno-fast-math',
            f'--offload-arch={ARCH}', '-Xclang', '-target-feature', '-Xclang',
            '-packed-fp32-ops', '-Wall', '-Wno-unused-function', '-Wno-unused-variable',
            '-Wno-unused-command-line-argument']


def hipcc():
    for cand in (os.environ.get('HIPCC'), '/opt/rocm/bin/hipcc', shutil.which('hipcc')):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError('hipcc not found')


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith('.hip'))


def build(verbose=True, jobs=None):
    """Compile every .hip source for gfx950 and link the shared library."""
    cc = hipcc()
    srcs = sources()
    tmp = tempfile.mkdtemp(prefix='tc_build_')
    objs = []
    procs = []
    for s in srcs:
        o = os.path.join(tmp, os.path.basename(s) + '.o')
        cmd = [cc, *HIPFLAGS, '-c', s, '-o', o]
        if verbose:
            print(' '.join(cmd), flush=True)
        procs.append((subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT), s))
        objs.append(o)
    failed = []
    for p, s in procs:
        out, _ = p.communicate()
        text = out.decode(errors='replace')
        if p.returncode != 0:
            failed.append((s, text))
        elif verbose and text.strip():
            # hipcc's host pass warns about the device-only target feature; keep other output.
            lines = [ln for ln in text.splitlines() if 'packed-fp32-ops' not in ln]
            if lines:
                print('\n'.join(lines))
    if failed:
        for s, text in failed:
            print(f'--- {s} ---\n{text}', file=sys.stderr)
        raise RuntimeError('hipcc failed for: ' + ', '.join(os.path.basename(s) for s, _ in failed))
    tmp_lib = LIB + '.tmp'
    cmd = [cc, f'--offload-arch={ARCH}', '-shared', '-fPIC', *objs, '-o', tmp_lib]
    if verbose:
        print(' '.join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(tmp_lib, LIB)
    shutil.rmtree(tmp, ignore_errors=True)
    return LIB


if __name__ == '__main__':
    build()
