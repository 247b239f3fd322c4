"""Benchmark of the treecorr_amd ball-tree engine on one MI355X.

Default workload (one "step" = one full process() of the hot path, tree build
included, with the catalog columns already resident in HBM):
  GG shear-shear, spherical: N = 2,000,000 galaxies in a 10 deg x 10 deg cap,
  70% uniform + 30% from 500 Gaussian clumps (sigma = 3 arcmin), g1,g2 ~ N(0, 0.2^2)
  clipped to |g| < 1, w ~ U[0.5, 1.5]; 25 log bins 1-200 arcmin, bin_slop = 0.1.
Metric: pairs/s = sum(npairs) / wall time of the step.
Multipole leg (the second half of the headline metric), under "multipole": GGG and
NNN LogMultipole (max_n = 30, 20 log bins 5-100 arcmin) of 500,000 points, 70% uniform
+ 30% in 500 clumps (sigma = 3 arcmin) over a 5 x 5 deg flat field; triangles/s =
sum(ntri at n = 0) / wall time of the step (tree build + multipole).

Prints ONE JSON line (rank 0) with the roofline of each leg's dominant kernel (the
2-point traversal; the multipole's stage kernel) measured with HIP events inside the
timed region, the plain GPU form beside the tuned one, and CPU baselines: TreeCorr's own
C++ engine (compiled from its sources into oracle/_ref) on all host cores, the whole
job, with its sums compared to the GPU's (npairs and ntri exactly); without that build,
the oracle's OpenMP restatement on a bounded sample (GG only).
"""
import argparse
import ctypes
import json
import math
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

ARCMIN = math.pi / 180. / 60.
FP64_PEAK_TFLOPS = 78.6     # MI355X FP64 (vector and matrix) dense peak
HBM_PEAK_GBS = 8000.        # MI355X HBM3E


# ---------------------------------------------------------------------------
# Synthetic workloads
# ---------------------------------------------------------------------------
def gg_sphere_catalog(n=2_000_000, seed=2026_12_03 + 2):
    """70% uniform + 30% clumped galaxies in a 10x10 deg cap (ra, dec in radians)."""
    rng = np.random.default_rng(seed)
    deg = math.pi / 180.
    ra_lo, ra_hi = 0., 10. * deg
    sd_lo, sd_hi = math.sin(-5. * deg), math.sin(5. * deg)
    nu = int(round(0.7 * n))
    ra = np.empty(n)
    dec = np.empty(n)
    ra[:nu] = rng.uniform(ra_lo, ra_hi, nu)
    dec[:nu] = np.arcsin(rng.uniform(sd_lo, sd_hi, nu))
    ncl = 500
    cra = rng.uniform(ra_lo, ra_hi, ncl)
    cdec = np.arcsin(rng.uniform(sd_lo, sd_hi, ncl))
    sig = 3. * ARCMIN
    filled = nu
    while filled < n:
        m = n - filled
        k = rng.integers(0, ncl, m)
        ddec = rng.normal(0, sig, m)
        dra = rng.normal(0, sig, m) / np.cos(cdec[k])
        r, d = cra[k] + dra, cdec[k] + ddec
        ok = (r >= ra_lo) & (r < ra_hi) & (np.sin(d) >= sd_lo) & (np.sin(d) < sd_hi)
        r, d = r[ok], d[ok]
        ra[filled:filled + len(r)] = r
        dec[filled:filled + len(d)] = d
        filled += len(r)
    g1 = rng.normal(0, 0.2, n)
    g2 = rng.normal(0, 0.2, n)
    bad = g1 ** 2 + g2 ** 2 >= 1
    while np.any(bad):
        g1[bad] = rng.normal(0, 0.2, bad.sum())
        g2[bad] = rng.normal(0, 0.2, bad.sum())
        bad = g1 ** 2 + g2 ** 2 >= 1
    w = rng.uniform(0.5, 1.5, n)
    return ra, dec, g1, g2, w


def radec_to_xyz(ra, dec):
    """GenerateXYZ through the library (glibc sincos, src/KMeans.cpp:1013-1050)."""
    from treecorr_amd import _lib
    n = len(ra)
    x, y, z = np.empty(n), np.empty(n), np.empty(n)
    _lib.check(_lib.lib().tc_generate_xyz(_lib.dptr(x), _lib.dptr(y), _lib.dptr(z),
                                          _lib.dptr(np.ascontiguousarray(ra)),
                                          _lib.dptr(np.ascontiguousarray(dec)), None, n))
    return x, y, z


def gg_binning():
    nbins, min_sep, max_sep, bin_slop, angle_slop = 25, 1. * ARCMIN, 200. * ARCMIN, 0.1, 0.1
    bin_size = math.log(max_sep / min_sep) / nbins
    b = bin_size * bin_slop
    b1 = min(angle_slop, b)
    return dict(nbins=nbins, min_sep=min_sep, max_sep=max_sep, bin_size=bin_size, b=b,
                a=angle_slop, min_size=min_sep * b1 / (2 + 3 * b1), max_size=max_sep * b1,
                min_top=4, max_top=10)


def mp_flat_catalog(n=500_000, seed=2026_12_03 + 4):
    """Multipole workload: 70% uniform + 30% from 500 Gaussian clumps (sigma = 3 arcmin)
    over a 5 x 5 deg flat field (positions in arcmin), g1,g2 ~ N(0, 0.2^2) clipped to
    |g| < 1, w ~ U[0.5, 1.5]."""
    rng = np.random.default_rng(seed)
    side = 300.
    nu = int(round(0.7 * n))
    x = np.empty(n)
    y = np.empty(n)
    x[:nu] = rng.uniform(0, side, nu)
    y[:nu] = rng.uniform(0, side, nu)
    cx = rng.uniform(0, side, 500)
    cy = rng.uniform(0, side, 500)
    filled = nu
    while filled < n:
        m = n - filled
        k = rng.integers(0, 500, m)
        xx = cx[k] + rng.normal(0, 3., m)
        yy = cy[k] + rng.normal(0, 3., m)
        ok = (xx >= 0) & (xx < side) & (yy >= 0) & (yy < side)
        xx, yy = xx[ok], yy[ok]
        x[filled:filled + len(xx)] = xx
        y[filled:filled + len(yy)] = yy
        filled += len(xx)
    g1 = rng.normal(0, 0.2, n)
    g2 = rng.normal(0, 0.2, n)
    bad = g1 ** 2 + g2 ** 2 >= 1
    while np.any(bad):
        g1[bad] = rng.normal(0, 0.2, bad.sum())
        g2[bad] = rng.normal(0, 0.2, bad.sum())
        bad = g1 ** 2 + g2 ** 2 >= 1
    w = rng.uniform(0.5, 1.5, n)
    return x, y, g1, g2, w


def mp_binning(max_n=30):
    """LogMultipole, 20 log bins 5-100 arcmin (TreeCorr's defaults: b = min(bin_size, 0.1),
    angle_slop = pi / (2 max_n + 1)); field sizes as corr3base.py:2934-2964."""
    nbins, min_sep, max_sep = 20, 5., 100.
    bin_size = math.log(max_sep / min_sep) / nbins
    b = bin_size if bin_size <= 0.1 else 0.1
    a = math.pi / (2 * max_n + 1)
    b1 = min(a, b)
    return dict(nbins=nbins, min_sep=min_sep, max_sep=max_sep, bin_size=bin_size, b=b, a=a,
                max_n=max_n, min_size=2 * min_sep * b1 / (2 + 3 * b1), max_size=max_sep * b1,
                min_top=4, max_top=10)


# ---------------------------------------------------------------------------
# Device memory (HIP runtime through ctypes)
# ---------------------------------------------------------------------------
class Hip:
    def __init__(self):
        for p in ('/opt/rocm/lib/libamdhip64.so', 'libamdhip64.so'):
            try:
                self.L = ctypes.CDLL(p)
                break
            except OSError:
                continue
        self.L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                     ctypes.c_int]
        self.L.hipFree.argtypes = [ctypes.c_void_p]

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=float)
        p = ctypes.c_void_p()
        assert self.L.hipMalloc(ctypes.byref(p), a.nbytes) == 0
        assert self.L.hipMemcpy(p, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, 1) == 0
        return p

    def sync(self):
        assert self.L.hipDeviceSynchronize() == 0

    def free(self, p):
        self.L.hipFree(p)


# ---------------------------------------------------------------------------
# One step: device tree build + 2-point traversal
# ---------------------------------------------------------------------------
def gg_step(L, dev, n, bn, out):
    from treecorr_amd import _lib
    fh = ctypes.c_void_p()
    _lib.check(L.tc_field_create_device(4, 3, dev['x'], dev['y'], dev['z'], dev['g1'], dev['g2'],
                                        dev['w'], None, n, bn['min_size'], bn['max_size'], 2, 0,
                                        0, bn['min_top'], bn['max_top'], ctypes.byref(fh)))
    ch = ctypes.c_void_p()
    arr = [out[k] for k in ('xi0', 'xi1', 'xi2', 'xi3', 'meanr', 'meanlogr', 'weight', 'npairs')]
    for a in arr:
        a[:] = 0
    _lib.check(L.tc_corr2_create(4, 4, 0, bn['min_sep'], bn['max_sep'], bn['nbins'],
                                 bn['bin_size'], bn['b'], bn['a'], -1.7976931348623157e308,
                                 1.7976931348623157e308, 0., 0., 0.,
                                 *[_lib.dptr(a) for a in arr], ctypes.byref(ch)))
    _lib.check(L.tc_corr2_process_auto(ch, fh, 0, 0, 1))
    c = (ctypes.c_longlong * 4)()
    _lib.check(L.tc_corr2_counters(ch, c))
    L.tc_corr2_destroy(ch)
    L.tc_field_destroy(fh)
    return dict(visits=c[0], accepted=c[1], leafpairs=c[2], refined=c[3])


# Algorithmic FP64 operations per unit of the 2-point path (GG, spherical):
# a visited node pair ~25 (distance, range tests, split rule); an accepted GG
# sphere pair 98 (bin index with sqrt/log, weights, two great-circle projections,
# both_complex_prod); a refined single-bin test +25.  Bytes: 2 node records of
# 48 B per visit.
OPS_VISIT, OPS_ACCEPT_GG_SPHERE, OPS_REFINED, BYTES_VISIT = 25, 98, 25, 96


def host_cores():
    """What the host offers and what the reference uses (its OpenMP thread count)."""
    try:
        nproc = int(subprocess.run(['nproc'], capture_output=True, text=True).stdout.strip())
    except Exception:
        nproc = None
    return {"os_cpu_count": os.cpu_count(), "nproc": nproc,
            "OMP_NUM_THREADS": os.environ.get('OMP_NUM_THREADS')}


def normwise(a, b):
    den = np.max(np.abs(b))
    return 0. if den == 0 else float(np.max(np.abs(a - b)) / den)


def cpu_baseline_reference(x, y, z, g1, g2, w, bn, gpu_out, runs=3):
    """The reference's own C++ engine (TreeCorr's _treecorr module compiled from its
    sources into oracle/_ref), the whole job on all host cores: tree build plus GG
    auto-correlation of the 2M galaxies, the median of `runs` runs.  Its raw sums are
    compared with the GPU's from the last timed step: npairs exactly, the FP64 sums
    normwise.  A per-core figure comes from one thread on the galaxies of a 2.5 x 2.5 deg
    corner of the cap (a bounded sample of the same workload).  None when the module is
    not built."""
    from oracle import ref as R
    m = R.module()
    if m is None:
        return None
    nthreads = int(m.GetOMPThreads())
    times = []
    for i in range(runs):
        t0 = time.perf_counter()
        f = R.make_field(4, 3, x, y, z, g1=g1, g2=g2, w=w, min_size=bn['min_size'],
                         max_size=bn['max_size'], split_method=2, brute=False,
                         min_top=bn['min_top'], max_top=bn['max_top'])
        r = R.corr2(4, f, None, bn['nbins'], bn['min_sep'], bn['max_sep'], bn['bin_size'],
                    bn['b'], bn['a'])
        times.append(time.perf_counter() - t0)
        del f
        print(f"# CPU reference run {i}: {times[-1]:.2f} s", file=sys.stderr, flush=True)
    t_full = float(np.median(times))
    npairs = float(r['npairs'].sum())
    parity = {"npairs_equal": bool(np.array_equal(gpu_out['npairs'], r['npairs'])),
              "sums_normwise_max": max(normwise(gpu_out[k], r[k]) for k in
                                       ('xi0', 'xi1', 'xi2', 'xi3', 'meanr', 'meanlogr',
                                        'weight'))}
    # one thread, a quarter by quarter corner (1/16 of the area)
    ra = np.arctan2(y, x)
    dec = np.arcsin(np.clip(z, -1., 1.))
    deg = math.pi / 180.
    sel = (ra < 2.5 * deg) & (dec < math.asin(math.sin(-5. * deg) + 0.25 * 2 * math.sin(5. * deg)))
    m.SetOMPThreads(1)
    t0 = time.perf_counter()
    f1 = R.make_field(4, 3, x[sel], y[sel], z[sel], g1=g1[sel], g2=g2[sel], w=w[sel],
                      min_size=bn['min_size'], max_size=bn['max_size'], split_method=2,
                      brute=False, min_top=bn['min_top'], max_top=bn['max_top'])
    r1 = R.corr2(4, f1, None, bn['nbins'], bn['min_sep'], bn['max_sep'], bn['bin_size'],
                 bn['b'], bn['a'])
    t1 = time.perf_counter() - t0
    m.SetOMPThreads(nthreads)
    return {"value": npairs / t_full, "unit": "pairs/s", "cores": nthreads,
            "kind": "reference", "host": host_cores(),
            "runs_s": times,
            "sample": (f"the whole job (no sampling), median of {runs} runs: TreeCorr's C++ "
                       f"engine compiled from its sources (oracle/_ref), tree build + GG auto of "
                       f"the same 2M-galaxy catalog on {nthreads} OpenMP threads, "
                       f"{t_full:.2f} s"),
            "one_thread": {"value": float(r1['npairs'].sum()) / t1, "unit": "pairs/s",
                           "sample": f"1 thread, the {int(sel.sum())} galaxies of a 2.5 x 2.5 deg "
                                     f"corner, tree build + GG auto, {t1:.2f} s"},
            "parity_vs_gpu": parity}


def cpu_baseline(x, y, z, g1, g2, w, bn, gpu_counters, gpu_npairs, target_s=15.):
    """The oracle's restatement of the reference engine (OpenMP, all host cores) on a
    bounded sample: the full tree build plus every k-th top-level row of the traversal;
    the full-job time is extrapolated by the ratio of node-pair visits.  Used when the
    compiled reference is not there."""
    from oracle import oracle as O
    if not os.path.exists(os.path.join(ROOT, 'oracle', 'libtc_oracle.so')):
        return None
    nthreads = int(os.environ.get('OMP_NUM_THREADS', '0')) or (os.cpu_count() or 1)
    t0 = time.perf_counter()
    f = O.Field(O.GDATA, O.SPHERE, x, y, z, g1=g1, g2=g2, w=w, min_size=bn['min_size'],
                max_size=bn['max_size'], min_top=bn['min_top'], max_top=bn['max_top'])
    t_build = time.perf_counter() - t0
    stride = 64
    while True:
        t1 = time.perf_counter()
        r = O.corr2(O.GDATA, O.SPHERE, f, None, bn['nbins'], bn['min_sep'], bn['max_sep'],
                    bn['bin_size'], bn['b'], bn['a'], exact=False, nthreads=nthreads,
                    row_stride=stride)
        t_samp = time.perf_counter() - t1
        if t_samp > 2. or stride == 1:
            break
        stride = max(1, stride // 4)
    frac = r['counters']['visits'] / max(1, gpu_counters['visits'])
    t_full = t_build + t_samp / max(frac, 1e-12)
    return {"value": float(gpu_npairs / t_full), "unit": "pairs/s", "cores": nthreads,
            "kind": "port",
            "sample": (f"oracle (OpenMP restatement of TreeCorr's engine, -O2) on the same "
                       f"2M-galaxy catalog: full tree build {t_build:.2f} s + every {stride}th "
                       f"top-level row of the traversal ({t_samp:.2f} s, "
                       f"{100 * frac:.2f}% of the node-pair visits); full-job time "
                       f"extrapolated by visit count = {t_full:.1f} s")}


# ---------------------------------------------------------------------------
# Multipole leg: GGG (LogMultipole, max_n = 30) of 500k points, tree build included
# ---------------------------------------------------------------------------
KEYS3 = ('zeta0', 'zeta1', 'zeta2', 'zeta3', 'zeta4', 'zeta5', 'zeta6', 'zeta7', 'meand1',
         'meanlogd1', 'meand2', 'meanlogd2', 'meand3', 'meanlogd3', 'meanu', 'meanv', 'weight',
         'weight_im', 'ntri')


# Algorithmic FP64 work of the multipole: a Gn call (accepted (c1, c2)) 24 M + 110 ops for
# GGG, 8 M + 25 for NNN; the zeta contraction 8 flops per complex multiply-add, per zeta
# cell (k2 row, k3 >= k2) and multipole n (2M+1 of them), for gam0..3 and the weight (5
# outputs; NNN: the weight only).
def gn_ops(d, M):
    return 24 * M + 110 if d == 4 else 8 * M + 25


def zeta_flops(d, M, cells):
    return 8. * cells * (2 * M + 1) * (5 if d == 4 else 1)


def mp_step(L, dev, n, bn, d, out):
    from treecorr_amd import _lib
    fh = ctypes.c_void_p()
    g = d == 4
    _lib.check(L.tc_field_create_device(d, 1, dev['x'], dev['y'], None,
                                        dev['g1'] if g else None, dev['g2'] if g else None,
                                        dev['w'], None, n, bn['min_size'], bn['max_size'], 2, 0, 0,
                                        bn['min_top'], bn['max_top'], ctypes.byref(fh)))
    for a in out.values():
        a[:] = 0
    h = ctypes.c_void_p()
    ptrs = [_lib.dptr(out[k]) if (g or not k.startswith('zeta')) else None for k in KEYS3]
    _lib.check(L.tc_corr3_create(d, 5, bn['min_sep'], bn['max_sep'], bn['nbins'],
                                 bn['bin_size'], bn['b'], bn['a'], bn['max_n'], *ptrs,
                                 ctypes.byref(h)))
    _lib.check(L.tc_corr3_process_auto(h, fh, 0, 0, 1))
    c = (ctypes.c_longlong * 4)()
    _lib.check(L.tc_corr3_counters(h, c))
    cells = ctypes.c_longlong()
    _lib.check(L.tc_corr3_zeta_cells(h, ctypes.byref(cells)))
    L.tc_corr3_destroy(h)
    L.tc_field_destroy(fh)
    return dict(gn_calls=c[0], zeta_calls=c[1], zeta_rows=c[2], stages=c[3],
                zeta_cells=cells.value)


def kernel_ms(L, which):
    from treecorr_amd import _lib
    ms = ctypes.c_double()
    la = ctypes.c_longlong()
    _lib.check(L.tc_kernel_stats(which, ctypes.byref(ms), ctypes.byref(la)))
    return ms.value, la.value


def timed_mp(L, hip, dev, n, bn, d, out, steps, warmup):
    for _ in range(warmup):
        mp_step(L, dev, n, bn, d, out)
    L.tc_reset_kernel_stats()
    L.tc_set_timing(1)
    hip.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        cnt = mp_step(L, dev, n, bn, d, out)
    hip.sync()
    t1 = time.perf_counter()
    L.tc_set_timing(0)
    return (t1 - t0) / steps, cnt, kernel_ms(L, 2), kernel_ms(L, 3)


def cpu_baseline_mp(x, y, g1, g2, w, bn, d, gpu_out):
    """TreeCorr's C++ engine (oracle/_ref) on the same catalog, all host cores, one run of
    the whole job (tree build + multipole); ntri compared exactly, the sums normwise."""
    from oracle import ref as R
    m = R.module()
    if m is None:
        return None
    t0 = time.perf_counter()
    f = R.make_field(d, 1, x, y, None, g1=g1 if d == 4 else None, g2=g2 if d == 4 else None,
                     w=w, min_size=bn['min_size'], max_size=bn['max_size'], split_method=2,
                     brute=False, min_top=bn['min_top'], max_top=bn['max_top'])
    r = R.corr3_multipole(d, f, bn['nbins'], bn['min_sep'], bn['max_sep'], bn['bin_size'],
                          bn['b'], bn['a'], bn['max_n'])
    t = time.perf_counter() - t0
    M = bn['max_n']
    keys = ['weight', 'weight_im', 'meand2', 'meanlogd2', 'meand3', 'meanlogd3'] + \
        (['zeta%d' % i for i in range(8)] if d == 4 else [])
    shape = r['ntri'].shape
    ntri = float(r['ntri'][:, :, M].sum())
    return {"value": ntri / t, "unit": "triangles/s", "cores": int(m.GetOMPThreads()),
            "kind": "reference", "host": host_cores(), "seconds": t,
            "sample": f"the whole job (no sampling), one run: TreeCorr's C++ engine compiled from "
                      f"its sources (oracle/_ref), tree build + {'GGG' if d == 4 else 'NNN'} "
                      f"multipole of the same catalog, {t:.1f} s",
            "parity_vs_gpu": {
                "ntri_equal": bool(np.array_equal(gpu_out['ntri'].reshape(shape), r['ntri'])),
                "sums_normwise_max": max(normwise(gpu_out[k].reshape(shape), r[k]) for k in keys)}}


def run_mp_leg(args, L, hip):
    x, y, g1, g2, w = mp_flat_catalog(args.mp_n)
    bn = mp_binning()
    M = bn['max_n']
    dev = {k: hip.upload(v) for k, v in dict(x=x, y=y, g1=g1, g2=g2, w=w).items()}
    ntot = bn['nbins'] ** 2 * (2 * M + 1)
    out = {k: np.zeros(ntot) for k in KEYS3}
    res = {}
    for d, name in ((4, 'GGG'), (0, 'NNN')):
        print(f"# multipole {name} {args.mp_n}: tuned", file=sys.stderr, flush=True)
        step_s, cnt, (st_ms, st_l), (zt_ms, zt_l) = timed_mp(L, hip, dev, args.mp_n, bn, d, out,
                                                             args.steps, args.warmup)
        ntri = float(out['ntri'].reshape(bn['nbins'], bn['nbins'], 2 * M + 1)[:, :, M].sum())
        gpu_out = {k: out[k].copy() for k in out}
        L.tc_set_multipole_plain(1)
        print(f"# multipole {name}: plain form", file=sys.stderr, flush=True)
        plain_s, _, (pst_ms, _), _ = timed_mp(L, hip, dev, args.mp_n, bn, d, out, 1, 0)
        L.tc_set_multipole_plain(0)
        flops = cnt['gn_calls'] * gn_ops(d, M)
        st_s = st_ms * 1e-3 / args.steps
        achieved = flops / st_s / 1e12 if st_s > 0 else 0.
        zfl = zeta_flops(d, M, cnt['zeta_cells'])
        zt_s = zt_ms * 1e-3 / args.steps
        traffic = read_traffic(f'k_mp_stage_{name.lower()}_500k.txt')
        leg = {
            "workload": f"{name} LogMultipole max_n={M}, {args.mp_n} points (70% uniform + 30% "
                        f"in 500 clumps) in a 5x5 deg flat field, 20 log bins 5-100 arcmin, "
                        f"tree build included",
            "value": ntri / step_s, "unit": "triangles/s", "ms_per_step": 1e3 * step_s,
            "gn_pairs_per_s": cnt['gn_calls'] / step_s,
            "plain_gpu": {"ms_per_step": 1e3 * plain_s, "value": ntri / plain_s,
                          "stage_kernel_ms": pst_ms,
                          "form": "Gn accumulation with every lane adding the multipoles in "
                                  "order (lanes of one bin contend for each LDS address)"},
            "tuned_over_plain": plain_s / step_s,
            "counters": cnt,
            "roofline": {"bound": "fp64_valu", "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
                         "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS,
                         "traffic": traffic, "kernel": f"k_mp_stage<{d},Flat>",
                         "flops_per_step": flops, "kernel_ms_per_step": st_ms / args.steps,
                         "launches_per_step": st_l / args.steps},
            "zeta": {"kernel": f"k_mp_zeta<{d}> + k_mp_zeta_reduce",
                     "kernel_ms_per_step": zt_ms / args.steps,
                     "achieved_tflops": zfl / zt_s / 1e12 if zt_s > 0 else 0.,
                     "flops_per_step": zfl},
        }
        # the zeta contraction on the matrix cores, against the FP64 contraction above
        nn = 2 * M + 1
        shape = (bn['nbins'], bn['nbins'], nn)
        pairs = [('weight', 'weight_im')] + ([('zeta%d' % (2 * i), 'zeta%d' % (2 * i + 1))
                                               for i in range(4)] if d == 4 else [])
        leg["zeta_matrix_cores"] = {}
        for mode, mname in ((1, 'fp32_mfma'), (2, 'bf16x3_mfma')):
            L.tc_set_zeta_mode(mode)
            print(f"# multipole {name}: zeta on the matrix cores, {mname}", file=sys.stderr,
                  flush=True)
            try:
                m_s, _, _, (mzt_ms, _) = timed_mp(L, hip, dev, args.mp_n, bn, d, out, args.steps, 1)
            finally:
                L.tc_set_zeta_mode(0)
            errs = {}
            for kr, ki in pairs:
                zr = (gpu_out[kr] + 1j * gpu_out[ki]).reshape(shape)
                zg = (out[kr] + 1j * out[ki]).reshape(shape)
                per_n = []
                for i in range(nn):
                    den = np.max(np.abs(zr[:, :, i]))
                    per_n.append(float(np.max(np.abs(zg[:, :, i] - zr[:, :, i])) / den)
                                 if den > 0 else 0.)
                errs[kr] = {"normwise_max_over_n": max(per_n), "normwise_per_n": per_n}
            leg["zeta_matrix_cores"][mname] = {
                "ms_per_step": 1e3 * m_s, "zeta_kernels_ms_per_step": mzt_ms / args.steps,
                "ntri_equal": bool(np.array_equal(out['ntri'], gpu_out['ntri'])),
                "error_vs_fp64": errs,
                "error_definition": "per multipole n: max over (k2,k3) of |Z - Z_fp64| / max "
                                    "over (k2,k3) of |Z_fp64|, Z the complex weight or gam_i"}
        if not args.no_cpu_baseline and d == 4:
            print(f"# CPU baseline, {name} (the reference engine, all host cores)",
                  file=sys.stderr, flush=True)
            leg["cpu_baseline"] = cpu_baseline_mp(x, y, g1, g2, w, bn, d, gpu_out)
        res[name] = leg
    for v in dev.values():
        hip.free(v)
    return res


# ---------------------------------------------------------------------------
# Patch workload: NN data x randoms and KK auto over 100 patches, jackknife
# ---------------------------------------------------------------------------
def lognormal_box(n, seed, box=1000., ngrid=256):
    """n points in a box from a log-normal field: white noise on ngrid^3 smoothed by a
    2-cell Gaussian, normalised to sigma 1, rho = exp(G - 1/2); cell occupation
    proportional to rho, uniform within the cell."""
    from scipy.ndimage import gaussian_filter
    rng = np.random.default_rng(seed)
    g = gaussian_filter(rng.standard_normal((ngrid,) * 3), 2.0, mode='wrap')
    g /= g.std()
    rho = np.exp(g - 0.5).ravel()
    del g
    cells = rng.choice(rho.size, size=n, p=rho / rho.sum())
    ix, iy, iz = np.unravel_index(cells, (ngrid,) * 3)
    h = box / ngrid
    return ((ix + rng.random(n)) * h, (iy + rng.random(n)) * h, (iz + rng.random(n)) * h)


def kmeans_centers(pts_fit, npatch, seed, iters=10):
    """Patch centres from Lloyd's k-means on a subsample (workload preparation; the
    catalogs then get their patches from these centres, as with TreeCorr's
    patch_centers)."""
    from scipy.spatial import cKDTree
    rng = np.random.default_rng(seed)
    cen = pts_fit[rng.choice(len(pts_fit), npatch, replace=False)]
    for it in range(iters):
        lab = cKDTree(cen).query(pts_fit, workers=-1)[1]
        print(f"# k-means iteration {it}", file=sys.stderr, flush=True)
        for k in range(npatch):
            m = lab == k
            if np.any(m):
                cen[k] = pts_fit[m].mean(axis=0)
    return cen


def patch_workload(nd, nr, npatch, seed=2026_12_03 + 3):
    x, y, z = lognormal_box(nd, seed)
    print("# log-normal data drawn", file=sys.stderr, flush=True)
    rng = np.random.default_rng(seed + 1)
    k = rng.normal(0, 0.03, nd)
    rx, ry, rz = (rng.uniform(0, 1000., nr) for _ in range(3))
    pick = rng.choice(nd, min(nd, 200_000), replace=False)
    sub = np.column_stack([x[pick], y[pick], z[pick]])
    cen = kmeans_centers(sub, npatch, seed + 2)
    return dict(x=x, y=y, z=z, k=k, rx=rx, ry=ry, rz=rz, centers=cen)


PATCH_BIN = dict(nbins=30, min_sep=1., max_sep=150., bin_slop=0.05, min_top=4)


def patch_step(tc, D, R):
    """process(D, R) for NN and process(D) for KK over all patch pairs, then the KK
    jackknife covariance."""
    nn = tc.NNCorrelation(angle_slop=0.1, var_method='jackknife', **PATCH_BIN)
    kk = tc.KKCorrelation(angle_slop=1.0, var_method='jackknife', **PATCH_BIN)
    nn.process(D, R)
    kk.process(D)
    cov = kk.estimate_cov('jackknife', cross_patch_weight='simple')
    return nn, kk, cov


def patch_cpu_baseline(tc, w, nn, kk, target_s=20.):
    """The reference's own C++ engine (TreeCorr's _treecorr module compiled from its
    sources into oracle/_ref; the oracle where it is not built) on a random sample of the
    patch-pair jobs, each with its two tree builds, on all host cores; the full-job time
    is extrapolated by the share of pairs sampled.  Every sampled pair's npairs is
    compared with the GPU's."""
    from oracle import ref as R
    m = R.module()
    if m is not None:
        nthreads = int(m.GetOMPThreads())
        kind = "reference"

        def field(d, x, y, z, k, kw):
            return R.make_field(d, 2, x, y, z, k=k, split_method=2, **kw)

        def corr(d, f1, f2, bn, bin_size, b, a):
            return R.corr2(d, f1, f2, bn['nbins'], bn['min_sep'], bn['max_sep'], bin_size, b, a)
    else:
        from oracle import oracle as O
        nthreads = int(os.environ.get('OMP_NUM_THREADS', '0')) or (os.cpu_count() or 1)
        kind = "port"

        def field(d, x, y, z, k, kw):
            return O.Field(d, 2, x, y, z, k=k, **kw)

        def corr(d, f1, f2, bn, bin_size, b, a):
            return O.corr2(d, 2, f1, f2, bn['nbins'], bn['min_sep'], bn['max_sep'], bin_size,
                           b, a, exact=False, nthreads=nthreads)
    bn = PATCH_BIN
    bin_size = math.log(bn['max_sep'] / bn['min_sep']) / bn['nbins']
    jobs = [('nn', ij) for ij, r in nn.results.items() if r.nonzero] + \
           [('kk', ij) for ij, r in kk.results.items() if r.nonzero]
    rng = np.random.default_rng(7)
    rng.shuffle(jobs)
    tot = float(nn.npairs.sum() + kk.npairs.sum())
    done, spent, njob, equal = 0., 0., 0, True
    for kd, (i, j) in jobs:
        a = 0.1 if kd == 'nn' else 1.0
        b = bin_size * bn['bin_slop']
        b1 = min(a, b)
        kw = dict(min_size=bn['min_sep'] * b1 / (2 + 3 * b1), max_size=bn['max_sep'] * b1,
                  min_top=bn['min_top'])
        t0 = time.perf_counter()
        di = w['ld'] == i
        if kd == 'nn':
            rj = w['lr'] == j
            f1 = field(0, w['x'][di], w['y'][di], w['z'][di], None, kw)
            f2 = field(0, w['rx'][rj], w['ry'][rj], w['rz'][rj], None, kw)
            r = corr(0, f1, f2, bn, bin_size, b, a)
            mine = nn.results[(i, j)]
        else:
            dj = w['ld'] == j
            f1 = field(1, w['x'][di], w['y'][di], w['z'][di], w['k'][di], kw)
            f2 = None if i == j else field(1, w['x'][dj], w['y'][dj], w['z'][dj], w['k'][dj], kw)
            r = corr(1, f1, f2, bn, bin_size, b, a)
            mine = kk.results[(i, j)]
        spent += time.perf_counter() - t0
        done += mine.npairs.sum()
        njob += 1
        equal = equal and bool(np.array_equal(mine.npairs, r['npairs']))
        if spent > target_s:
            break
    t_full = spent * tot / max(done, 1.)
    what = ("TreeCorr's C++ engine compiled from its sources (oracle/_ref)" if kind == "reference"
            else "the oracle (an OpenMP restatement of TreeCorr's engine)")
    return {"value": tot / t_full, "unit": "pairs/s", "cores": nthreads, "kind": kind,
            "sample": (f"{what} on {njob} random patch-pair jobs, {100 * done / tot:.2f}% of the "
                       f"pairs, each with its two tree builds ({spent:.1f} s); full-job time "
                       f"extrapolated by pair count = {t_full:.0f} s; jackknife not included"),
            "parity_vs_gpu": {"sampled_pairs_npairs_equal": equal}}


def run_patch_workload(args):
    import treecorr_amd as tc
    from treecorr_amd import _lib
    L = _lib.lib()
    w = patch_workload(args.patch_nd, args.patch_nr, args.npatch)
    print(f"# patch workload ready: {args.patch_nd} data, {args.patch_nr} randoms, "
          f"{args.npatch} patches", file=sys.stderr, flush=True)

    def cats():
        # patches from the centres (QuickAssign on the GPU), for both catalogs
        return (tc.Catalog(x=w['x'], y=w['y'], z=w['z'], k=w['k'], patch_centers=w['centers']),
                tc.Catalog(x=w['rx'], y=w['ry'], z=w['rz'], patch_centers=w['centers']))
    L.tc_set_graph_mode(1)
    tuned = None
    if args.autotune:
        from treecorr_amd import tune
        D, R = cats()
        best, times = tune.autotune(lambda: patch_step(tc, D, R), min_jobs=(1 << 24, 1 << 26),
                                    grab_div=(4, 16))
        tuned = {"best": {"min_jobs": best[0], "grab_div": best[1]},
                 "step_s": {f"{k[0]}/{k[1]}": v for k, v in times.items()}}
        print(f"# autotuned: {tuned}", file=sys.stderr, flush=True)
    for i in range(args.warmup):
        patch_step(tc, *cats())
        print(f"# warmup {i} done", file=sys.stderr, flush=True)
    sets = [cats() for _ in range(args.steps)]
    w['ld'], w['lr'] = sets[0][0].patch, sets[0][1].patch
    L.tc_reset_kernel_stats()
    L.tc_set_timing(1)
    t0 = time.perf_counter()
    for i in range(args.steps):
        nn, kk, cov = patch_step(tc, *sets[i])
        print(f"# step {i} done", file=sys.stderr, flush=True)
    t1 = time.perf_counter()
    L.tc_set_timing(0)
    ms = ctypes.c_double()
    launches = ctypes.c_longlong()
    _lib.check(L.tc_kernel_stats(0, ctypes.byref(ms), ctypes.byref(launches)))
    npairs = float(nn.npairs.sum() + kk.npairs.sum())
    step_s = (t1 - t0) / args.steps
    res = {
        "metric": f"pairs/s (NN data x randoms + KK auto over {args.npatch} patches, jackknife)",
        "value": npairs / step_s, "unit": "pairs/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * step_s, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded; log-normal data in a 1000^3 box, uniform randoms, "
                "patches from k-means centres, assigned on the GPU)",
        "config": {"workload": f"NN {args.patch_nd} data x {args.patch_nr} randoms and KK auto "
                               f"of the data, {args.npatch} patches (every patch pair in one "
                               f"traversal), 30 log bins 1-150, bin_slop=0.05, KK jackknife "
                               f"covariance; host catalogs, patch split and tree builds included",
                   "nd": args.patch_nd, "nr": args.patch_nr, "npatch": args.npatch},
        "traversal_ms_per_step": ms.value / args.steps,
        "traversal_launches_per_step": launches.value / args.steps,
        "jobs": {"nn_pairs_stored": len(nn.results), "kk_pairs_stored": len(kk.results)},
        "counters": {"nn": nn.counters, "kk": kk.counters},
        "cov_shape": list(cov.shape),
        "graph_mode": True,
        "autotune": tuned,
    }
    if not args.no_cpu_baseline:
        res["cpu_baseline"] = patch_cpu_baseline(tc, w, nn, kk)
    print(json.dumps(res))


# ---------------------------------------------------------------------------
# Learned scheduler workload: KK of 20M points at (RA, Dec, r); the cost model of the
# persistent traversal trained on job logs of seeded runs shaped like the other workloads
# ---------------------------------------------------------------------------
def corr_step(L, D, C, dev, n, bn, out, min_top=None):
    """Tree build from device columns + one 2-point auto correlation through the C ABI;
    returns the work counters (out receives the raw sums)."""
    from treecorr_amd import _lib
    fh = ctypes.c_void_p()
    _lib.check(L.tc_field_create_device(D, C, dev['x'], dev['y'], dev.get('z'), dev.get('k'),
                                        dev.get('g2'), dev.get('w'), None, n, bn['min_size'],
                                        bn['max_size'], 2, 0, 0,
                                        bn['min_top'] if min_top is None else min_top,
                                        bn['max_top'], ctypes.byref(fh)))
    ch = ctypes.c_void_p()
    arr = [out[k] for k in ('xi0', 'xi1', 'xi2', 'xi3', 'meanr', 'meanlogr', 'weight', 'npairs')]
    for a in arr:
        a[:] = 0
    _lib.check(L.tc_corr2_create(D, D, 0, bn['min_sep'], bn['max_sep'], bn['nbins'],
                                 bn['bin_size'], bn['b'], bn['a'], -1.7976931348623157e308,
                                 1.7976931348623157e308, 0., 0., 0.,
                                 *[_lib.dptr(a) for a in arr], ctypes.byref(ch)))
    _lib.check(L.tc_corr2_process_auto(ch, fh, 0, 0, 1))
    c = (ctypes.c_longlong * 4)()
    _lib.check(L.tc_corr2_counters(ch, c))
    L.tc_corr2_destroy(ch)
    L.tc_field_destroy(fh)
    return dict(visits=c[0], accepted=c[1], leafpairs=c[2], refined=c[3])


def log_binning(nbins, min_sep, max_sep, bin_slop, angle_slop=0.1, min_top=4, max_top=10):
    bin_size = math.log(max_sep / min_sep) / nbins
    b = bin_size * bin_slop
    b1 = min(angle_slop, b)
    return dict(nbins=nbins, min_sep=min_sep, max_sep=max_sep, bin_size=bin_size, b=b,
                a=angle_slop, min_size=min_sep * b1 / (2 + 3 * b1), max_size=max_sep * b1,
                min_top=min_top, max_top=max_top)


def lightcone_catalog(n, seed, ra_deg=30., dec_deg=30., rmin=500., rmax=2000., cell=8.):
    """n points at (RA, Dec, r): RA uniform over ra_deg, Dec uniform on the sphere over
    dec_deg (a cap about the x axis), r in [rmin, rmax]; the density a log-normal field
    (white noise on a grid of `cell` Mpc smoothed by a 2-cell Gaussian, sigma_ln 1).
    kappa = 0.02 x the smoothed Gaussian field at the point + N(0, 0.05^2).  Returns
    x, y, z (Mpc) and kappa."""
    from scipy.ndimage import gaussian_filter
    rng = np.random.default_rng(seed)
    ha, hd = math.radians(ra_deg) / 2, math.radians(dec_deg) / 2
    lo = np.array([rmin * math.cos(ha) * math.cos(hd), -rmax * math.sin(ha), -rmax * math.sin(hd)])
    hi = np.array([rmax, rmax * math.sin(ha), rmax * math.sin(hd)])
    ng = np.maximum(1, np.ceil((hi - lo) / cell).astype(int))
    g = gaussian_filter(rng.standard_normal(tuple(ng)), 2.0, mode='wrap')
    g /= g.std()
    rho = np.exp(g - 0.5).ravel()
    # the cap fills about 43% of its box: draw 2.6 n candidates, keep those inside
    xs, ys, zs, gs = [], [], [], []
    kept = 0
    while kept < n:
        m = int(2.6 * (n - kept)) + 1000
        cells = rng.choice(rho.size, size=m, p=rho / rho.sum())
        ix, iy, iz = np.unravel_index(cells, tuple(ng))
        x = lo[0] + (ix + rng.random(m)) * cell
        y = lo[1] + (iy + rng.random(m)) * cell
        z = lo[2] + (iz + rng.random(m)) * cell
        r = np.sqrt(x * x + y * y + z * z)
        ra = np.arctan2(y, x)
        sd = z / r
        ok = (r >= rmin) & (r < rmax) & (np.abs(ra) < ha) & (np.abs(sd) < math.sin(hd))
        xs.append(x[ok]); ys.append(y[ok]); zs.append(z[ok]); gs.append(g.ravel()[cells[ok]])
        kept += int(ok.sum())
    x, y, z, gv = (np.concatenate(a)[:n] for a in (xs, ys, zs, gs))
    kappa = 0.02 * gv + rng.normal(0, 0.05, n)
    return x, y, z, kappa


def training_runs():
    """Seeded runs shaped like the other workloads (seeds of their own): NN flat, GG
    sphere, KK 3-d log-normal box, KK flat clumpy field.  Yields (D, C, columns, binning)."""
    rng = np.random.default_rng(2026_12_03 + 50)
    n1 = 100_000
    yield (0, 1, dict(x=rng.uniform(0, 1000., n1), y=rng.uniform(0, 1000., n1), w=np.ones(n1)),
           log_binning(20, 1., 100., 0.0))
    ra, dec, g1, g2, w = gg_sphere_catalog(500_000, seed=2026_12_03 + 51)
    x, y, z = radec_to_xyz(ra, dec)
    yield (4, 3, dict(x=x, y=y, z=z, k=g1, g2=g2, w=w), gg_binning())
    x, y, z = lognormal_box(1_000_000, 2026_12_03 + 52, ngrid=128)
    k = np.random.default_rng(53).normal(0, 0.03, len(x))
    yield (1, 2, dict(x=x, y=y, z=z, k=k, w=np.ones(len(x))), log_binning(30, 1., 150., 0.05, 1.0))
    x, y, g1, g2, w = mp_flat_catalog(500_000, seed=2026_12_03 + 54)
    yield (1, 1, dict(x=x, y=y, k=g1, w=w), log_binning(20, 5., 100., 0.1))


def run_learned_workload(args):
    from treecorr_amd import _lib, cost
    L = _lib.lib()
    hip = Hip()
    res = {"workload": "learned scheduler", "metric": "pairs/s (KK 3-d, 20M points at RA/Dec/r)"}
    # 1. job logs of the training runs (two job granularities each)
    Xs, ys = [], []
    t0 = time.perf_counter()
    for D, C, cols, bn in training_runs():
        n = len(cols['x'])
        dev = {k: hip.upload(np.ascontiguousarray(v, np.float64)) for k, v in cols.items()}
        out = {k: np.zeros(bn['nbins']) for k in ('xi0', 'xi1', 'xi2', 'xi3', 'meanr',
                                                  'meanlogr', 'weight', 'npairs')}
        for mj in (1 << 18, 1 << 22):
            _lib.check(L.tc_set_persist_params(0, mj))
            cost.set_mode(cost.DYNAMIC)
            cost.start_log(150_000)
            corr_step(L, D, C, dev, n, bn, out)
            X, v = cost.stop_log()
            Xs.append(X)
            ys.append(cost.target(v))
        for p in dev.values():
            hip.free(p)
        print(f"# logged D={D} C={C} n={n}: {sum(len(a) for a in ys)} jobs so far",
              file=sys.stderr, flush=True)
    _lib.check(L.tc_set_persist_params(0, 0))
    X = np.concatenate(Xs)
    y = np.concatenate(ys)
    t1 = time.perf_counter()
    model = cost.CostModel(seed=7)
    hip.sync()
    t2 = time.perf_counter()
    loss = model.train(X, y, steps=args.train_steps, batch=4096, lr=2e-3)
    hip.sync()
    t3 = time.perf_counter()
    res["training"] = {"samples": int(len(y)), "runs": "NN flat 100k, GG sphere 500k, "
                       "KK 3-d log-normal 1M, KK flat clumpy 500k (seeds of their own); jobs "
                       "logged at 2^18 and 2^22 jobs per run",
                       "logging_s": t1 - t0, "train_s": t3 - t2, "steps": args.train_steps,
                       "batch": 4096, "mse_log2": loss, "target_var": float(np.var(y))}
    print(f"# trained: mse {loss:.3f} (target variance {np.var(y):.3f}) in {t3 - t2:.2f} s",
          file=sys.stderr, flush=True)
    # 2. the 20M-point catalog
    x, yy, z, kap = lightcone_catalog(args.lc_n, 2026_12_03 + 5)
    bn = log_binning(40, 0.5, 200., 0.1, 0.1)
    dev = {k: hip.upload(v) for k, v in dict(x=x, y=yy, z=z, k=kap, w=np.ones(len(x))).items()}
    n = len(x)
    out = {k: np.zeros(bn['nbins']) for k in ('xi0', 'xi1', 'xi2', 'xi3', 'meanr', 'meanlogr',
                                              'weight', 'npairs')}
    print(f"# light-cone catalog: {n} points", file=sys.stderr, flush=True)
    # held-out prediction error: a logged run on this catalog
    print("# logged run on the light cone ...", file=sys.stderr, flush=True)
    cost.set_mode(cost.DYNAMIC)
    cost.start_log(2_000_000)
    tl = time.perf_counter()
    corr_step(L, 1, 2, dev, n, bn, out)
    Xh, vh = cost.stop_log()
    print(f"# logged run on the light cone: {time.perf_counter() - tl:.2f} s, "
          f"npairs {out['npairs'].sum():.4g}", file=sys.stderr, flush=True)
    pred = model.predict(Xh)
    th = cost.target(vh)
    resid = pred - th
    res["prediction"] = {"jobs_sampled": int(len(th)), "r2": float(1 - np.var(resid) / np.var(th)),
                         "corr": float(np.corrcoef(pred, th)[0, 1]),
                         "mean_abs_err_log2": float(np.mean(np.abs(resid))),
                         "median_abs_err_log2": float(np.median(np.abs(resid)))}
    # 3. the modes, timed (tree build included, as each is one process call)
    model.install()
    modes = [(cost.STATIC, "static equal-count shares (fallback and judge)"),
             (cost.LEARNED_STATIC, "static shares of equal predicted work"),
             (cost.DYNAMIC, "jobs by a shared counter, natural order"),
             (cost.LEARNED_DYNAMIC, "jobs by a shared counter, longest predicted first"),
             (cost.CHUNKED, "chunked launches driven from the host")]
    ref = None
    res["modes"] = {}
    for mode, what in modes:
        cost.set_mode(mode)
        for i in range(args.warmup):
            corr_step(L, 1, 2, dev, n, bn, out)
            print(f"# mode {mode}: warm-up {i} done", file=sys.stderr, flush=True)
        L.tc_reset_kernel_stats()
        L.tc_set_timing(1)
        hip.sync()
        ta = time.perf_counter()
        for i in range(args.steps):
            cnt = corr_step(L, 1, 2, dev, n, bn, out)
            print(f"# mode {mode}: step {i} done", file=sys.stderr, flush=True)
        hip.sync()
        tb = time.perf_counter()
        L.tc_set_timing(0)
        ms = ctypes.c_double()
        la = ctypes.c_longlong()
        _lib.check(L.tc_kernel_stats(0, ctypes.byref(ms), ctypes.byref(la)))
        got = {k: v.copy() for k, v in out.items()}
        if ref is None:
            ref = got
        ent = {"what": what, "ms_per_step": 1e3 * (tb - ta) / args.steps,
               "traversal_ms_per_step": ms.value / args.steps,
               "pairs_per_s": float(got['npairs'].sum()) / ((tb - ta) / args.steps),
               "npairs_equal_to_static": bool(np.array_equal(got['npairs'], ref['npairs'])),
               "sums_normwise_vs_static": max(normwise(got[k], ref[k]) for k in
                                              ('xi0', 'meanr', 'meanlogr', 'weight'))}
        res["modes"][str(mode)] = ent
        print(f"# mode {mode}: {ent['ms_per_step']:.1f} ms/step", file=sys.stderr, flush=True)
    cost.set_mode(cost.DYNAMIC)
    cost.clear_model()
    res["counters"] = cnt
    st = res["modes"][str(cost.STATIC)]["traversal_ms_per_step"]
    res["learned_static_speedup_over_static"] = st / res["modes"][str(cost.LEARNED_STATIC)]["traversal_ms_per_step"]
    res["learned_dynamic_speedup_over_dynamic"] = (
        res["modes"][str(cost.DYNAMIC)]["traversal_ms_per_step"] /
        res["modes"][str(cost.LEARNED_DYNAMIC)]["traversal_ms_per_step"])
    res["config"] = {"workload": f"KK 3-d auto, {n} points at (RA, Dec, r): RA, Dec over a 30 x 30 "
                                 f"deg cap, r 500-2000 Mpc from a log-normal density field, "
                                 f"kappa = smoothed field + N(0, 0.05^2); 40 log bins 0.5-200 Mpc, "
                                 f"bin_slop 0.1; tree build included", "n": n}
    res["dtype"] = "f64 (model FP32)"
    res["data"] = "synthetic (seeded)"
    print(json.dumps(res))


def read_traffic(name):
    """HBM bytes per launch measured with rocprofv3 --pmc (committed summary), or None."""
    path = os.path.join(ROOT, 'profiles', name)
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        for line in fh:
            if line.startswith('traffic_bytes_per_launch:'):
                return float(line.split(':', 1)[1])
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=3)
    ap.add_argument('--warmup', type=int, default=1)
    ap.add_argument('--n', type=int, default=2_000_000)
    ap.add_argument('--no-cpu-baseline', action='store_true')
    ap.add_argument('--calibrate-copy', type=int, default=0, metavar='MIB',
                    help='only run the plain copy kernel over MIB MiB (HBM counter calibration)')
    ap.add_argument('--workload', choices=('gg_sphere', 'patch_cov_3d', 'learned_kk3d'),
                    default='gg_sphere',
                    help='gg_sphere: the headline GG workload; patch_cov_3d: NN data x randoms '
                         'and KK auto over k-means patches with a jackknife covariance; '
                         'learned_kk3d: the traversal cost model trained on job logs, then '
                         'KK of 20M points at RA/Dec/r in every traversal mode')
    ap.add_argument('--lc-n', type=int, default=20_000_000)
    ap.add_argument('--train-steps', type=int, default=4000)
    ap.add_argument('--patch-nd', type=int, default=4_000_000)
    ap.add_argument('--patch-nr', type=int, default=40_000_000)
    ap.add_argument('--npatch', type=int, default=100)
    ap.add_argument('--autotune', action='store_true',
                    help='patch workload: autotune job granularity and queue depth first')
    ap.add_argument('--mp-n', type=int, default=500_000,
                    help='points of the multipole (GGG, NNN) leg of the headline line')
    ap.add_argument('--no-mp', action='store_true', help='leave out the multipole leg')
    ap.add_argument('--only-mp', action='store_true',
                    help='run only the multipole leg (its own JSON line, for profiling)')
    args = ap.parse_args()
    if args.calibrate_copy:
        from treecorr_amd import _lib
        hip = Hip()
        n = args.calibrate_copy * 1024 * 1024 // 8
        src = hip.upload(np.ones(n))
        dst = hip.upload(np.zeros(n))
        _lib.check(_lib.lib().tc_copy_calibration(src, dst, n))
        print(json.dumps({"calibration_copy_bytes_read": n * 8, "bytes_written": n * 8}))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "treecorr_amd runs on one GPU; --gpus must be 1 "
                                   "(the multi-GPU sharding is described, not built)"}))
        sys.exit(2)
    if args.workload == 'patch_cov_3d':
        run_patch_workload(args)
        return
    if args.workload == 'learned_kk3d':
        run_learned_workload(args)
        return
    from treecorr_amd import _lib
    L = _lib.lib()
    hip = Hip()
    if args.only_mp:
        print(json.dumps({"multipole": run_mp_leg(args, L, hip)}))
        return
    ra, dec, g1, g2, w = gg_sphere_catalog(args.n)
    x, y, z = radec_to_xyz(ra, dec)
    bn = gg_binning()
    dev = {k: hip.upload(v) for k, v in dict(x=x, y=y, z=z, g1=g1, g2=g2, w=w).items()}
    nb = bn['nbins']
    out = {k: np.zeros(nb) for k in ('xi0', 'xi1', 'xi2', 'xi3', 'meanr', 'meanlogr', 'weight',
                                     'npairs')}
    for i in range(args.warmup):
        cnt = gg_step(L, dev, args.n, bn, out)
        print(f"# warmup {i} done: npairs={out['npairs'].sum():.4g}", file=sys.stderr, flush=True)
    L.tc_reset_kernel_stats()
    L.tc_set_timing(1)
    hip.sync()
    t0 = time.perf_counter()
    for i in range(args.steps):
        cnt = gg_step(L, dev, args.n, bn, out)
        print(f"# step {i} done", file=sys.stderr, flush=True)
    hip.sync()
    t1 = time.perf_counter()
    L.tc_set_timing(0)
    ms = ctypes.c_double()
    launches = ctypes.c_longlong()
    _lib.check(L.tc_kernel_stats(0, ctypes.byref(ms), ctypes.byref(launches)))
    pms, plaunch = kernel_ms(L, 4)                    # the persistent launches alone
    pc = (ctypes.c_longlong * 4)()
    _lib.check(L.tc_persist_counters(pc))             # the last step's persistent launch
    npairs = float(out['npairs'].sum())
    ms_per_step = 1e3 * (t1 - t0) / args.steps
    value = npairs / ((t1 - t0) / args.steps)
    flops = (OPS_VISIT * cnt['visits'] + OPS_ACCEPT_GG_SPHERE * cnt['accepted'] +
             OPS_REFINED * cnt['refined'])
    # roofline of the dominant kernel, k_traverse_persist: its own counters (the chunked
    # split before it is counted apart) over its average launch time
    flops_p = OPS_VISIT * pc[0] + OPS_ACCEPT_GG_SPHERE * pc[1] + OPS_REFINED * pc[3]
    avg_launch_s = pms * 1e-3 / max(1, plaunch)
    achieved = flops_p / avg_launch_s / 1e12 if avg_launch_s > 0 else 0.
    launches_per_step = launches.value / args.steps
    traffic = read_traffic('k_traverse_persist_gg_sphere_2M.txt')
    gpu_out = {k: v.copy() for k, v in out.items()}
    # the plain form: the tree built level by level in reference order (no speculative
    # top levels) and the traversal in chunked launches driven from the host (one
    # read-back per launch)
    L.tc_set_tree_speculation(0)
    L.tc_set_traversal_mode(0)
    hip.sync()
    tp0 = time.perf_counter()
    for i in range(args.steps):
        gg_step(L, dev, args.n, bn, out)
    hip.sync()
    tp1 = time.perf_counter()
    L.tc_set_tree_speculation(1)
    L.tc_set_traversal_mode(1)
    plain_s = (tp1 - tp0) / args.steps
    res = {
        "metric": "pairs/s (GG xi+/xi- 2M galaxies, bin_slop=0.1) + GGG multipole 500k",
        "value": value,
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded; 70% uniform + 30% clumped, spherical cap)",
        "config": {"workload": "GG shear-shear, 2,000,000 galaxies, 10x10 deg cap, 25 log bins "
                               "1-200 arcmin, bin_slop=0.1, tree build included",
                   "n": args.n, "nbins": nb, "bin_slop": 0.1, "min_top": bn['min_top'],
                   "max_top": bn['max_top']},
        "roofline": {"bound": "fp64_valu", "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS,
                     "traffic": traffic,
                     "kernel": "k_traverse_persist<GData,Sphere,Item>",
                     "flops_per_launch": flops_p,
                     "counters_per_launch": dict(zip(('visits', 'accepted', 'leafpairs',
                                                      'refined'), (int(v) for v in pc))),
                     "avg_launch_ms": 1e3 * avg_launch_s,
                     "launches_per_step": plaunch / args.steps,
                     "hbm_GBs_from_counters": traffic / avg_launch_s / 1e9 if traffic else None,
                     "note": "compute bound on the FP64 vector ALUs; priced at the 78.6 TF/s "
                             "FP64 peak of MI355X"},
        "plain_gpu": {"ms_per_step": 1e3 * plain_s, "value": npairs / plain_s,
                      "form": "tree built level by level in reference order (no speculative "
                              "top levels); traversal in chunked launches with one host "
                              "read-back each (no persistent launch)"},
        "tuned_over_plain": plain_s / ((t1 - t0) / args.steps),
        "counters": cnt,
        "traversal_ms_per_step": ms.value / args.steps,
        "traversal_launches_per_step": launches_per_step,
        "whole_traversal_tflops": flops / (ms.value * 1e-3 / args.steps) / 1e12,
    }
    if not args.no_cpu_baseline:
        print("# CPU baseline (the reference engine, all host cores)", file=sys.stderr,
              flush=True)
        base = cpu_baseline_reference(x, y, z, g1, g2, w, bn, gpu_out)
        res["cpu_baseline"] = base if base is not None else \
            cpu_baseline(x, y, z, g1, g2, w, bn, cnt, npairs)
        if res["cpu_baseline"] is not None:
            res["cpu_baseline"]["gpu_over_cpu"] = value / res["cpu_baseline"]["value"]
    for v in dev.values():
        hip.free(v)
    if not args.no_mp:
        res["multipole"] = run_mp_leg(args, L, hip)
    print(json.dumps(res))


if __name__ == '__main__':
    main()
