#!/usr/bin/env python3
"""Benchmark of the MI355X FFT-VIE solve path (one GPU).

A step is one mixed-precision solve (FP32 matvec + 2-level circulant
right preconditioner + GMRES(restart) inside FP64 iterative refinement) of the
workload to a true FP64 relative residual <= tol, with the operator,
preconditioner and right-hand side resident in HBM.  The JSON line also
carries the plain FP64 GPU solve rate, FP64/FP32 matvec times with GB/s
against the canonical byte model, p50/p99 latency, the roofline of the
dominant kernel (HIP events on the library's stream) and a CPU baseline
(the FP64 NumPy/SciPy oracle matvec timed on the host cores).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = 'solves/s at 1e-6 residual (mixed vs FP64); ms/matvec + GB/s; p50/p99 latency'
HBM_PEAK = 8.0e12        # MI355X HBM3E, bytes/s


def _prec_dtype():
    half = os.environ.get('VP_PC_HALF', '1') != '0'
    blocks = 'binary16-stored blocks (scaled, FP32 arithmetic)' if half else 'FP32 blocks'
    return (f'c64 inner (FP32 matvec, FP32 GMRES, FP32 preconditioner FFTs with {blocks}), '
            'c128 outer residual and refinement')


PREC_DTYPE = _prec_dtype()

WORKLOADS = {
    'bragg_10nm': dict(
        desc='Bragg grating, Si 500 nm +/-50 nm corrugation, 220 nm, 320 nm x 100 periods, quadratic '
             'absorbers (Im eps 0->1 over 2 um) at both ends, SiO2, lambda 1550 nm, 10 nm voxels, '
             'grid 3600x100x44 (FFT 7200x200x88), y-polarised Gaussian (w0 300 nm)',
        tol=1e-6, restart=100, inner_tol=1e-3, inner_maxit=6, max_outer=12, prec='face', axis=1, damping=0.0),
    'straight_wg_20nm': dict(
        desc='straight Si waveguide 500x220 nm in SiO2, lambda 1550 nm, 20 nm voxels, grid 200x48x24 '
             '(FFT 400x96x48), y-polarised Gaussian (w0 300 nm)',
        tol=1e-6, restart=50, inner_tol=1e-2, inner_maxit=2, max_outer=30, prec='scalar', damping=0.0),
    'sweep_bragg_20nm': dict(
        desc='wavelength sweep of the 20 nm Bragg grating (1800x50x22, FFT 3600x100x44): 101 wavelengths '
             '1500-1600 nm, each with its own Green tensor and preconditioner, 2 right-hand sides per wavelength '
             '(y- and z-polarised Gaussians), solved as batches of 4 wavelengths x 2 RHS; GMRES(150), recycled '
             'space of 30 harmonic Ritz directions',
        tol=1e-6, restart=150, inner_tol=1e-2, inner_maxit=4, max_outer=12, prec='face', axis=1, damping=0.0,
        base='bragg_20nm', nlam=101, lam0=1500e-9, lam1=1600e-9, chunk=4, pols=('y', 'z'), recycle_width=30),
    'mor_bragg_20nm': dict(
        desc='transmission spectrum of the 20 nm Bragg grating (1800x50x22, FFT 3600x100x44) by model order '
             'reduction: 5 Gauss-Legendre sample wavelengths in 1500-1600 nm, order 5 (25 solves per structure), '
             'reduced models of the grating and of the straight waveguide, 1001 wavelengths at 0.1 nm; '
             'complex-point-source beam (y-polarised, 440 nm waist parameter) as the reference example; basis '
             'solves to 1e-4 as Bragg_MOR_script.m, accuracy checks against full solves to 1e-6',
        tol=1e-4, restart=100, inner_tol=1e-2, inner_maxit=4, max_outer=12, prec='face', axis=1, damping=0.0,
        base='bragg_20nm', nG=5, order=5, lam0=1500e-9, lam1=1600e-9, dlam=0.1e-9, check_tol=1e-6,
        check=(1512.3e-9, 1526.9e-9, 1587.1e-9)),
    'dr_bragg_20nm': dict(
        desc='GMRES with deflated restarts (pgmresDR, GCRO-DR) against restarted GMRES (pgmres) with the same '
             'short restart, FP64, on the 20 nm Bragg grating at 1550 nm (1800x50x22) with the face-averaged '
             '2-level right preconditioner',
        tol=1e-6, restart=40, ritz=15, maxit=75, prec='face', axis=1, damping=0.0, base='bragg_20nm'),
    'serve_mix': dict(
        desc='serving stream: seeded Poisson arrivals (mean 50 req/s) of mixed solves at 20 nm voxels, '
             'lambda U[1500,1600] nm: 40% straight waveguides of length U[2,8] um, 40% Bragg gratings of '
             'U{20..150} periods, 20% microdisks of radius U[2,5] um (220 nm thick); requests admitted by the '
             'server HBM budget, refinement rounds batched per geometry; GMRES(150), recycled space of 30 '
             'harmonic Ritz directions per request',
        tol=1e-6, restart=150, inner_tol=1e-2, inner_maxit=4, max_outer=20, nreq=40, rate=50.0, seed=4,
        budget_gb=96, max_active=32, setup_workers=8, fft_sizes='min', recycle_width=30),
    'microdisk_15nm': dict(
        desc='Si microdisk r = 10 um, 220 nm thick, with a 500 nm bus waveguide at a 200 nm gap (quadratic '
             'absorbers, Im eps 0->1 over 2 um, at both bus ends), SiO2, lambda 1550 nm, 15 nm voxels, grid '
             '1400x1500x30 (63 M voxels), y-polarised Gaussian (w0 300 nm) centred on the bus; block-Jacobi '
             'composite of 2-level circulants: face-averaged and exact across the bus on the bus box, core '
             'contrast on the disk box',
        tol=1e-6, restart=100, inner_tol=1e-2, inner_maxit=1, max_outer=40, outer_restart=30, prec='composite',
        damping=0.0, fields=False, cpu_baseline=False),
    'bragg_20nm': dict(
        desc='Bragg grating (100 periods) at 20 nm voxels, grid 1800x50x22 (FFT 3600x100x44), lambda 1550 nm',
        tol=1e-6, restart=100, inner_tol=1e-3, inner_maxit=6, max_outer=12, prec='face', axis=1, damping=0.0),
}


def mv_model_bytes(prob, s):
    """Canonical algorithmic bytes of one matvec: s*(6P + 12Q + 3n + nS) + 8*nS."""
    L, M, N = prob.shape
    P = 8 * L * M * N
    Q = 2 * L * M * N
    nS = int(prob.mask.sum())
    n = 3 * nS
    return s * (6 * P + 12 * Q + 3 * n + nS) + 8 * nS


def unknown_box(prob):
    """Extents (Lb, Mb, Nb) of the bounding box of the unknown voxels."""
    ext = []
    for ax in range(3):
        other = tuple(a for a in range(3) if a != ax)
        nz = np.flatnonzero(prob.mask.any(axis=other))
        ext.append(int(nz[-1] - nz[0] + 1))
    return tuple(ext)


def kernel_model_bytes(prob, kernel, s, fsym=True, pc_box=None, fft=None):
    """Algorithmic bytes per launch of one matvec or preconditioner kernel.

    The passes run on the unknowns' bounding box Lb x Mb x Nb inside the
    Lp x Mp x Np FFT grid (``fft``: the sizes the operator's plan runs at,
    default the reference's 2L x 2M x 2N): Qb = Lp*Mb*Nb x-spectra and
    Yb = Lp*Mp*Nb y-spectra per component.  The split z pass reads the parity octant of
    fN, Pc = (Lp/2+1)(Mp/2+1)(Np/2+1) entries per component (fsym), the
    fused plane kernel the full P = Lp*Mp*Np.

    Preconditioner kernels (pc_box = (Lc, Mc, Nc): two circulant axes and
    the dense one): the quadrant form stores (Lc/2+1)(Mc/2+1) blocks of
    (3 Nc)^2 entries, read once per apply ('pc_block', which also reads and
    writes the 3 Lc Mc Nc spectra; FP32 applies read binary16 blocks unless
    VP_PC_HALF=0); 'pc_yrows' reads and writes the spectra.
    """
    if kernel.startswith('pc_'):
        if pc_box is None:
            return None
        Lc, Mc, Nc = pc_box
        V3 = 3 * Lc * Mc * Nc
        n = prob.n
        if kernel == 'pc_block':   # FP32 applies stream the blocks as binary16 pairs (4 bytes)
            bs = 4 if s == 8 and os.environ.get('VP_PC_HALF', '1') != '0' else s
            return bs * (Lc // 2 + 1) * (Mc // 2 + 1) * (3 * Nc) ** 2 + s * 2 * V3
        if kernel == 'pc_yrows':
            return s * 2 * V3
        if kernel == 'pc_xfwd':
            return s * (n + V3)
        if kernel == 'pc_xinv':
            return s * (V3 + 2 * n)
        return None
    L, M, N = prob.shape
    Lp, Mp, Np = fft if fft is not None else (2 * L, 2 * M, 2 * N)
    Lb, Mb, Nb = unknown_box(prob)
    P = Lp * Mp * Np
    Pc = (Lp // 2 + 1) * (Mp // 2 + 1) * (Np // 2 + 1) if fsym else P
    Qb = Lp * Mb * Nb
    Yb = Lp * Mp * Nb
    nS = int(prob.mask.sum())
    n = 3 * nS
    if kernel == 'plane':     # fused: fN once, x-spectra read and written
        return s * (6 * P + 6 * Qb)
    if kernel == 'yfwd':      # x-spectra read, y-spectra written
        return s * (3 * Qb + 3 * Yb)
    if kernel == 'zpass':     # y-spectra read and written, fN octant once
        return s * (6 * Pc + 6 * Yb)
    if kernel == 'yinv':      # y-spectra read, x-spectra written
        return s * (3 * Yb + 3 * Qb)
    if kernel == 'xfwd':      # J read, x-spectra written, voxel map read
        return s * (n + 3 * Qb) + 4 * nS
    if kernel == 'xinv':      # x-spectra read, J and Mc read, output written, voxel map read
        return s * (3 * Qb + 2 * n + nS) + 4 * nS
    return None


def kernel_table(prob, stats, pc_box, fft=None):
    """Per-kernel launches, mean duration (HIP events) and, where the byte
    model covers the kernel, its FP32 algorithmic GB/s."""
    out = {}
    for k, (cnt, ms) in stats.items():
        avg = ms / 1e3 / max(cnt, 1)
        row = dict(launches=cnt, avg_us=round(avg * 1e6, 2))
        kb = kernel_model_bytes(prob, k, 8, pc_box=pc_box, fft=fft)
        if kb is not None and cnt:
            row['GBps'] = round(kb / avg / 1e9, 1)
        out[k] = row
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=3)
    ap.add_argument('--warmup', type=int, default=1)
    ap.add_argument('--workload', default='bragg_10nm', choices=sorted(WORKLOADS))
    ap.add_argument('--no-prec', action='store_true')
    ap.add_argument('--no-cpu-baseline', action='store_true')
    ap.add_argument('--no-fp64-solve', action='store_true')
    ap.add_argument('--no-fields', action='store_true', help='skip the excitation / total-field timing')
    ap.add_argument('--inner-tol', type=float, default=None, help='override the workload inner tolerance')
    ap.add_argument('--inner-maxit', type=int, default=None, help='override the workload inner restart cycles')
    ap.add_argument('--batch', type=int, default=None, help='sweep: wavelengths per batched solve (override)')
    ap.add_argument('--max-outer', type=int, default=None, help='override the workload refinement-step limit')
    ap.add_argument('--nlam', type=int, default=None, help='sweep: number of wavelengths (override)')
    ap.add_argument('--restart', type=int, default=None, help='override the workload GMRES restart')
    ap.add_argument('--recycle-width', type=int, default=None,
                    help='recycled Krylov space: harmonic Ritz directions kept (0: the whole cycle)')
    ap.add_argument('--reference-path', action='store_true',
                    help="also time one solve on the reference's FP64 GPU path (pgmres on host arrays, "
                         'transfers per operator application, operator at 2L x 2M x 2N)')
    ap.add_argument('--traffic-probe', action='store_true',
                    help='for rocprofv3 --pmc runs: a 1 GiB device copy (calibration), then 3 FP32 matvecs '
                         'and 3 FP32 preconditioner applies')
    args = ap.parse_args()
    if args.gpus != 1:
        print(f'bench.py measures one GPU only (asked for --gpus {args.gpus}); every solve of this path fits '
              'one MI355X, and multi-GPU sharding is described, not run.', file=sys.stderr)
        sys.exit(2)

    from voxphoton_amd import vie, configs
    from voxphoton_amd import _lib as L

    wl = dict(WORKLOADS[args.workload])
    if args.inner_tol is not None:
        wl['inner_tol'] = args.inner_tol
    if args.inner_maxit is not None:
        wl['inner_maxit'] = args.inner_maxit
    if args.max_outer is not None:
        wl['max_outer'] = args.max_outer
    if args.batch is not None and 'chunk' in wl:
        wl['chunk'] = args.batch
    if args.nlam is not None and 'nlam' in wl:
        wl['nlam'] = args.nlam
    if args.restart is not None:
        wl['restart'] = args.restart
    if args.recycle_width is not None:
        wl['recycle_width'] = args.recycle_width
    log = lambda *a: print('[bench]', *a, file=sys.stderr, flush=True)   # noqa: E731
    if 'nlam' in wl:
        run_sweep(args, wl, log)
        return
    if 'nG' in wl:
        run_mor(args, wl, log)
        return
    if 'ritz' in wl:
        run_dr(args, wl, log)
        return
    if 'nreq' in wl:
        run_serve(args, wl, log)
        return
    prob = configs.config_problem(args.workload)
    ctx = vie.default_context()
    n = prob.n
    log(f'workload {args.workload}: grid {prob.shape}, unknowns {n}')

    t0 = time.perf_counter()
    op = vie.NOperator.assembled(prob.shape, prob.dx, prob.ko, None, prob.Mc, prob.idx)
    t_op = time.perf_counter() - t0
    oinfo = op.info()
    fft32, fft64 = oinfo['fft_fp32'], oinfo['fft_fp64']
    log(f'operator assembled + embedded in {t_op:.2f} s; unknowns box {oinfo["box"]}, autotuned FFT sizes '
        f'FP32 {fft32} FP64 {fft64} (reference {tuple(2 * v for v in prob.shape)})')
    if args.traffic_probe:
        P, _ = make_preconditioner(prob, wl, vie, configs)
        traffic_probe(op, P, ctx, n, vie, L, log)
        return
    prec = None
    pdesc = None
    t_pc = 0.0
    if not args.no_prec:
        t0 = time.perf_counter()
        prec, pdesc = make_preconditioner(prob, wl, vie, configs)
        t_pc = time.perf_counter() - t0
        log(f'{pdesc} built in {t_pc:.2f} s')

    b = prob.gaussian_rhs('y')
    db = vie.DeviceArray(ctx, b.nbytes)
    dx = vie.DeviceArray(ctx, b.nbytes)
    db.upload(b)
    hp = prec.handle if prec is not None else None

    def solve_mixed():
        info = L.GmresInfo()
        args = [op.handle, hp, L.VP_PREC_RIGHT, db.ptr, dx.ptr, L.VP_MEM_DEVICE, wl['tol'], wl['restart'],
                wl['inner_tol'], wl['inner_maxit'], wl['max_outer']]
        if wl.get('outer_restart'):
            # flexible refinement: FGMRES in FP64 over the inner solves
            L.check(L.lib.vp_solve_fgmres_ir(*args, wl['outer_restart'], ctypes.byref(info)))
        else:
            L.check(L.lib.vp_solve_ir(*args, ctypes.byref(info)))
        return info

    for _ in range(args.warmup):
        info = solve_mixed()
        log(f'warmup solve: relres {info.relres:.3e}, IR steps {info.ir_steps}, inner its {info.total_iters}')
    def read_stats(names):
        out = {}
        for k in names:
            l_, m_ = ctypes.c_int64(), ctypes.c_double()
            L.check(L.lib.vp_ctx_kernel_stats(ctx.handle, k.encode(), ctypes.byref(l_), ctypes.byref(m_), 1))
            out[k] = (l_.value, m_.value)
        return out

    L.check(L.lib.vp_ctx_set_profiling(ctx.handle, 1))
    kernels = ['xfwd', 'plane', 'yfwd', 'zpass', 'yinv', 'xinv', 'pc_xfwd', 'pc_plane', 'pc_yrows', 'pc_block',
               'pc_xinv']
    for k in kernels:
        l_, m_ = ctypes.c_int64(), ctypes.c_double()
        L.check(L.lib.vp_ctx_kernel_stats(ctx.handle, k.encode(), ctypes.byref(l_), ctypes.byref(m_), 1))
    lat = []
    infos = []
    ctx.synchronize()
    t_start = time.perf_counter()
    for _ in range(args.steps):
        ts = time.perf_counter()
        infos.append(solve_mixed())
        ctx.synchronize()
        lat.append(time.perf_counter() - ts)
    ctx.synchronize()
    t_total = time.perf_counter() - t_start
    stats = {}
    for k in kernels:
        l_, m_ = ctypes.c_int64(), ctypes.c_double()
        L.check(L.lib.vp_ctx_kernel_stats(ctx.handle, k.encode(), ctypes.byref(l_), ctypes.byref(m_), 1))
        stats[k] = (l_.value, m_.value)
    L.check(L.lib.vp_ctx_set_profiling(ctx.handle, 0))
    solves_per_s = args.steps / t_total
    worst = max(i.relres for i in infos)
    log(f'mixed: {solves_per_s:.3f} solves/s, worst relres {worst:.3e}')

    # ---- matvec microbenchmark (device vectors, both precisions)
    x64 = vie.DeviceArray(ctx, n * 16)
    y64 = vie.DeviceArray(ctx, n * 16)
    rng = np.random.default_rng(1550)
    xv = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    x64.upload(xv)
    x32 = vie.DeviceArray(ctx, n * 8)
    y32 = vie.DeviceArray(ctx, n * 8)
    x32.upload(xv.astype(np.complex64))

    def time_mv(xd, yd, prec_id, reps=10):
        for _ in range(2):
            L.check(L.lib.vp_op_apply(op.handle, 0, xd.ptr, yd.ptr, L.VP_MEM_DEVICE, prec_id))
        ctx.synchronize()
        t = time.perf_counter()
        for _ in range(reps):
            L.check(L.lib.vp_op_apply(op.handle, 0, xd.ptr, yd.ptr, L.VP_MEM_DEVICE, prec_id))
        ctx.synchronize()
        return (time.perf_counter() - t) / reps

    def built_bytes(xd, yd, prec_id, s):
        # one profiled pass: which matvec kernels ran, summed as-built bytes
        L.check(L.lib.vp_ctx_set_profiling(ctx.handle, 1))
        mv_kernels = ['xfwd', 'plane', 'yfwd', 'zpass', 'yinv', 'xinv']
        read_stats(mv_kernels)
        L.check(L.lib.vp_op_apply(op.handle, 0, xd.ptr, yd.ptr, L.VP_MEM_DEVICE, prec_id))
        ctx.synchronize()
        ran = read_stats(mv_kernels)
        L.check(L.lib.vp_ctx_set_profiling(ctx.handle, 0))
        fft = fft64 if s == 16 else fft32
        return sum(kernel_model_bytes(prob, k, s, fft=fft) for k, (cnt, _) in ran.items() if cnt)

    t_mv64 = time_mv(x64, y64, L.VP_FP64)
    t_mv32 = time_mv(x32, y32, L.VP_FP32)
    B64 = mv_model_bytes(prob, 16)
    B32 = mv_model_bytes(prob, 8)
    Bb64 = built_bytes(x64, y64, L.VP_FP64, 16)
    Bb32 = built_bytes(x32, y32, L.VP_FP32, 8)
    log(f'matvec fp64 {t_mv64 * 1e3:.3f} ms ({B64 / t_mv64 / 1e9:.0f} GB/s canonical model, '
        f'{Bb64 / t_mv64 / 1e9:.0f} GB/s as built), fp32 {t_mv32 * 1e3:.3f} ms ({B32 / t_mv32 / 1e9:.0f} GB/s '
        f'canonical, {Bb32 / t_mv32 / 1e9:.0f} GB/s as built)')

    # ---- plain FP64 GPU solve (serial), into its own buffer: the mixed
    # solution stays in dx and is compared with it
    fp64 = None
    x_err = None
    if not args.no_fp64_solve:
        dx64 = vie.DeviceArray(ctx, b.nbytes)
        info = L.GmresInfo()
        ctx.synchronize()
        t = time.perf_counter()
        L.check(L.lib.vp_gmres(op.handle, hp, L.VP_PREC_RIGHT, db.ptr, None, dx64.ptr, L.VP_MEM_DEVICE, L.VP_FP64,
                               wl['restart'], wl['tol'], 400, ctypes.byref(info)))
        ctx.synchronize()
        t64 = time.perf_counter() - t
        fp64 = dict(solves_per_s=1.0 / t64, seconds=t64, iterations=info.total_iters, relres=info.relres,
                    flag=info.flag)
        log(f'plain fp64: {t64:.3f} s/solve, {info.total_iters} its, relres {info.relres:.3e}')
        xm = dx.download(np.empty(n, dtype=complex))
        x64s = dx64.download(np.empty(n, dtype=complex))
        x_err = float(np.linalg.norm(xm - x64s) / np.linalg.norm(x64s))
        # both solutions have true residuals <= tol, so they differ by at most
        # about cond(A) * 2 tol; the measured difference is reported beside it
        log(f'mixed vs plain FP64 solution: relative difference {x_err:.3e}')
        del xm, x64s, dx64

    # ---- the reference's own FP64 GPU path (opt-in: host Arnoldi, transfers per application)
    refp = None
    if args.reference_path and prec is not None:
        from voxphoton_amd import refpath
        _, ri = refpath.solve_reference_path(vie, prob, prec, b, wl['restart'], wl['tol'], 40, ctx=ctx)
        refp = dict(ri, fft=list(ri['fft']), solves_per_s=1.0 / ri['seconds'])
        log(f'reference FP64 GPU path: {ri["seconds"]:.2f} s, {ri["iterations"]} its, relres {ri["relres"]:.2e}')

    # ---- roofline of the dominant kernel
    dom = max(stats, key=lambda k: stats[k][1])
    launches, ms = stats[dom]
    avg_s = ms / 1e3 / max(launches, 1)
    pc_box = getattr(prec, 'box_shape', None)
    kb = kernel_model_bytes(prob, dom, 8, pc_box=pc_box, fft=fft32)
    roof = None
    if kb is not None and launches:
        achieved = kb / avg_s / 1e9
        traffic = None
        tf = os.path.join(ROOT, 'profiles', f'traffic_{dom}_{args.workload}.json')
        if os.path.exists(tf):
            with open(tf) as f:
                traffic = json.load(f).get('bytes_per_launch')
        roof = dict(kernel=dom, bound='hbm', achieved=round(achieved, 1), peak=HBM_PEAK / 1e9, unit='GB/s',
                    frac=round(achieved / (HBM_PEAK / 1e9), 4), traffic=traffic, avg_us=round(avg_s * 1e6, 2),
                    launches=launches, model_bytes_per_launch=kb, precision='fp32')

    # ---- excitation and post-processing on the same grid
    fields = None
    if not args.no_fields and wl.get('fields', True):
        fields = field_block(prob, vie, L, ctx, dx, log)

    # ---- CPU baseline (FP64 oracle on the host cores); the oracle's matvec
    # of the seeded vector also checks the device matvecs of the same vector
    cpu = None
    mv_check = None
    if not args.no_cpu_baseline and wl.get('cpu_baseline', True):
        cpu, y_ref = cpu_baseline(prob, vie, fp64, log, pc_box=getattr(prec, 'box_shape', None),
                                  restart=wl['restart'])
        yd64 = y64.download(np.empty(n, dtype=complex))
        L.check(L.lib.vp_op_apply(op.handle, 0, x32.ptr, y32.ptr, L.VP_MEM_DEVICE, L.VP_FP32))
        ctx.synchronize()
        yd32 = y32.download(np.empty(n, dtype=np.complex64)).astype(complex) * op.fp32_scale
        nref = np.linalg.norm(y_ref)
        mv_check = dict(fp64_rel_err=float(np.linalg.norm(yd64 - y_ref) / nref),
                        fp32_rel_err=float(np.linalg.norm(yd32 - y_ref) / nref),
                        bound_fp64=1e-12, bound_fp32=1e-5,
                        vector='default_rng(1550) N(0,1) + i N(0,1), complex128')
        log(f'device matvec vs FP64 oracle: fp64 {mv_check["fp64_rel_err"]:.2e}, fp32 {mv_check["fp32_rel_err"]:.2e}')
        del yd64, yd32, y_ref
        if mv_check['fp64_rel_err'] > 1e-12 or mv_check['fp32_rel_err'] > 1e-5:
            raise SystemExit(f'device matvec disagrees with the FP64 oracle: {mv_check}')

    lat_ms = np.array(lat) * 1e3
    line = {
        'metric': METRIC,
        'value': round(solves_per_s, 4),
        'unit': 'solves/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': round(t_total / args.steps * 1e3, 3),
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': PREC_DTYPE,
        'data': 'synthetic (device-assembled Green tensor, generated geometry)',
        'config': {'workload': args.workload, 'description': wl['desc'], 'unknowns': n, 'tol': wl['tol'],
                   'restart': wl['restart'], 'inner_tol': wl['inner_tol'], 'inner_maxit': wl['inner_maxit'],
                   'preconditioner': None if prec is None else pdesc, 'preconditioning': 'right',
                   'refinement': (f'flexible (FGMRES in FP64 over the inner solves, cycles of {wl["outer_restart"]})'
                                  if wl.get('outer_restart') else 'one correction per step'),
                   'parallelism': 'single GPU'},
        'relres_worst': worst,
        'ir_steps': [i.ir_steps for i in infos],
        'inner_iterations': [i.total_iters for i in infos],
        'latency_ms': {'p50': round(float(np.percentile(lat_ms, 50)), 3),
                       'p99': round(float(np.percentile(lat_ms, 99)), 3)},
        'fp64_plain': fp64,
        'mixed_vs_fp64': None if fp64 is None else round(solves_per_s / fp64['solves_per_s'], 3),
        'reference_fp64_gpu_path': refp,
        'mixed_vs_reference_fp64_gpu_path': None if refp is None else round(solves_per_s * refp['seconds'], 2),
        'x_err_vs_fp64': x_err,
        'matvec_vs_oracle': mv_check,
        'matvec': {'fp64_ms': round(t_mv64 * 1e3, 4), 'fp32_ms': round(t_mv32 * 1e3, 4),
                   'model_bytes_fp64': B64, 'model_bytes_fp32': B32,
                   'built_bytes_fp64': Bb64, 'built_bytes_fp32': Bb32,
                   'fp64_GBps_built': round(Bb64 / t_mv64 / 1e9, 1), 'fp32_GBps_built': round(Bb32 / t_mv32 / 1e9, 1),
                   'fp64_GBps': round(B64 / t_mv64 / 1e9, 1), 'fp32_GBps': round(B32 / t_mv32 / 1e9, 1),
                   'fp64_frac_hbm': round(B64 / t_mv64 / HBM_PEAK, 4),
                   'fp32_frac_hbm': round(B32 / t_mv32 / HBM_PEAK, 4)},
        'kernel_ms': kernel_table(prob, stats, pc_box, fft=fft32),
        'fft_sizes': {'fp32': fft32, 'fp64': fft64, 'reference': [2 * v for v in prob.shape],
                      'unknowns_box': oinfo['box'], 'chosen_by': 'autotuner (timed per box and precision)'},
        'setup_s': {'operator': round(t_op, 3), 'preconditioner': round(t_pc, 3)},
        'roofline': roof,
        'fields': fields,
        'cpu_baseline': cpu,
    }
    print(json.dumps(line), flush=True)


def run_sweep(args, wl, log):
    """Batched wavelength sweep: every wavelength has its own operator (Green
    tensor assembled on the device) and preconditioner; the right-hand sides
    of ``chunk`` wavelengths are solved together by vp_solve_ir_batched (one
    launch per kernel per Krylov step for all of them).  The baseline is the
    plain FP64 GPU path run serially, timed on the first and last wavelength
    and scaled to all solves."""
    from voxphoton_amd import vie, configs
    from voxphoton_amd import _lib as L
    import dataclasses
    base = configs.config_problem(wl['base'])
    lams = np.linspace(wl['lam0'], wl['lam1'], wl['nlam'])
    pols = wl['pols']
    ctx = vie.default_context()
    ctx.set_recycle_width(wl.get('recycle_width', -1))
    n = base.n
    nsolve = len(lams) * len(pols)
    log(f'sweep: {len(lams)} wavelengths x {len(pols)} RHS on grid {base.shape}, unknowns {n}')

    def build(lam):
        prob = dataclasses.replace(base, lam=float(lam))
        op = vie.NOperator.assembled(prob.shape, prob.dx, prob.ko, None, prob.Mc, prob.idx)
        P, _ = make_preconditioner(prob, wl, vie, configs)
        rhs = [prob.gaussian_rhs(pol) for pol in pols]
        return prob, op, P, rhs

    t_setup = 0.0
    t_solve = 0.0
    infos_all = []
    chunk = wl['chunk']
    lat = []

    def solve_batch(ops, precs, B):
        ns = B.shape[0]
        db = vie.DeviceArray(ctx, B.nbytes)
        dxv = vie.DeviceArray(ctx, B.nbytes)
        db.upload(B)
        hops = (ctypes.c_void_p * ns)(*[o.handle.value for o in ops])
        hpre = (ctypes.c_void_p * ns)(*[p.handle.value for p in precs])
        infos = (L.GmresInfo * ns)()
        args = [ctypes.cast(hops, ctypes.c_void_p), ctypes.cast(hpre, ctypes.c_void_p), ns, L.VP_PREC_RIGHT, db.ptr,
                dxv.ptr, wl['tol'], wl['restart'], wl['inner_tol'], wl['inner_maxit'], wl['max_outer']]
        ctx.synchronize()
        t0 = time.perf_counter()
        L.check(L.lib.vp_solve_ir_batched(*args, ctypes.cast(infos, ctypes.POINTER(L.GmresInfo))))
        ctx.synchronize()
        return time.perf_counter() - t0, [infos[i] for i in range(ns)]

    for c0 in range(0, len(lams), chunk):
        t0 = time.perf_counter()
        items = [build(lam) for lam in lams[c0:c0 + chunk]]
        ctx.synchronize()
        t_setup += time.perf_counter() - t0
        dt, infos = solve_batch([it[1] for it in items for _ in pols], [it[2] for it in items for _ in pols],
                                np.array([r for it in items for r in it[3]]))
        t_solve += dt
        lat.append(dt)
        infos_all += [dict(flag=i.flag, relres=i.relres, ir=i.ir_steps, its=i.total_iters) for i in infos]
        log(f'  wavelengths {c0}..{c0 + len(items) - 1}: {len(infos)} solves in {dt:.2f} s '
            f'(worst relres {max(i.relres for i in infos):.2e}, inner its '
            f'{[i.total_iters for i in infos]}), setup {t_setup:.1f} s so far')
        del items
    value = nsolve / t_solve
    worst = max(i['relres'] for i in infos_all)
    line = {
        'metric': METRIC, 'value': round(value, 4), 'unit': 'solves/s', 'n_gpus': 1,
        'steps': nsolve, 'warmup': 0, 'ms_per_step': round(t_solve / nsolve * 1e3, 3),
        'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None,
        'dtype': PREC_DTYPE,
        'data': 'synthetic (device-assembled Green tensor per wavelength, generated geometry)',
        'config': {'workload': args.workload, 'description': wl['desc'], 'unknowns': n, 'solves': nsolve,
                   'tol': wl['tol'], 'restart': wl['restart'], 'inner_tol': wl['inner_tol'],
                   'recycle_width': wl.get('recycle_width', 0),
                   'batch': chunk * len(pols), 'preconditioning': 'right', 'parallelism': 'single GPU'},
        'relres_worst': worst, 'all_converged': all(i['flag'] == 0 for i in infos_all),
        'inner_iterations_mean': float(np.mean([i['its'] for i in infos_all])),
        'batch_seconds': [round(v, 3) for v in lat],
        'setup_s_per_wavelength': round(t_setup / len(lams), 3),
    }
    if args.no_fp64_solve:
        print(json.dumps(line), flush=True)
        return
    # plain FP64 serial baseline on the first and last wavelength
    t64 = []
    its64 = []
    for lam in (lams[0], lams[-1]):
        prob, op, P, rhs = build(lam)
        dbb = vie.DeviceArray(ctx, rhs[0].nbytes)
        dxx = vie.DeviceArray(ctx, rhs[0].nbytes)
        dbb.upload(rhs[0])
        info = L.GmresInfo()
        ctx.synchronize()
        t0 = time.perf_counter()
        L.check(L.lib.vp_gmres(op.handle, P.handle, L.VP_PREC_RIGHT, dbb.ptr, None, dxx.ptr, L.VP_MEM_DEVICE,
                               L.VP_FP64, wl['restart'], wl['tol'], 400, ctypes.byref(info)))
        ctx.synchronize()
        t64.append(time.perf_counter() - t0)
        its64.append(info.total_iters)
        log(f'  plain FP64 at {lam * 1e9:.0f} nm: {t64[-1]:.3f} s, {info.total_iters} its, relres {info.relres:.2e}')
    fp64_serial = nsolve * float(np.mean(t64))
    # the reference's own FP64 GPU path (JVIE_Solver_const GPU_flag = 1):
    # pgmres on host arrays, every operator and preconditioner application
    # uploads, runs on the device (operator at 2L x 2M x 2N) and downloads;
    # timed on the same first and last wavelength (first right-hand side)
    from voxphoton_amd import refpath
    rinfos = []
    for lam in (lams[0], lams[-1]):
        prob, op, P, rhs = build(lam)
        _, ri = refpath.solve_reference_path(vie, prob, P, rhs[0], wl['restart'], wl['tol'], 40, ctx=ctx)
        rinfos.append(ri)
        log(f'  reference FP64 GPU path at {lam * 1e9:.0f} nm: {ri["seconds"]:.2f} s, {ri["iterations"]} its, '
            f'relres {ri["relres"]:.2e}, FFT {ri["fft"]}')
        del op, P
    rsec = float(np.mean([ri['seconds'] for ri in rinfos]))
    line.update({
        'fp64_plain_serial': {'seconds_per_solve': round(float(np.mean(t64)), 4), 'iterations': its64,
                              'solves_per_s': round(nsolve / fp64_serial, 4)},
        'mixed_batched_vs_fp64_serial': round(value / (nsolve / fp64_serial), 3),
        'reference_fp64_gpu_path': {
            'description': "the reference's FP64 GPU solve path (JVIE_Solver_const GPU_flag=1): pgmres "
                           'on host arrays (numpy, the host cores), every operator and preconditioner '
                           'application uploaded, run in FP64 on the device (operator at 2L x 2M x 2N on the '
                           'full grid) and downloaded; same preconditioner, restart and tolerance; the '
                           'first and last wavelength (first right-hand side), serial',
            'seconds_per_solve': round(rsec, 3), 'seconds': [round(ri['seconds'], 3) for ri in rinfos],
            'iterations': [ri['iterations'] for ri in rinfos], 'relres': [ri['relres'] for ri in rinfos],
            'flags': [ri['flag'] for ri in rinfos], 'fft': list(rinfos[0]['fft']),
            'solves_per_s': round(1.0 / rsec, 4)},
        'mixed_batched_vs_reference_fp64_gpu_path': round(value * rsec, 2),
    })
    print(json.dumps(line), flush=True)


def run_mor(args, wl, log):
    """Transmission spectrum by model order reduction, as the reference's
    example_Bragg_1500nm_to_1600nm.m with Bragg_MOR_script.m and
    straight_WG_sweep_heat.m: per structure, operators, preconditioners and
    field operators at nG Gauss-Legendre wavelengths, nG*order device solves
    (mixed precision) for the moment basis, projection, then every swept
    wavelength from the reduced model (excitation, reduced solve, total field
    through the interpolated field operator, power through a cross-section),
    all on the device.  The accuracy check solves the full grating at a few
    swept wavelengths and compares power and residual."""
    from voxphoton_amd import vie, configs, mor
    import dataclasses
    ctx = vie.default_context()
    ctx.set_recycle_width(wl.get('recycle_width', -1))
    base = configs.config_problem(wl['base'])
    straight = configs.bragg_grating(dx=base.dx, corr=0.0, M=base.M, N=base.N)
    xg, _ = np.polynomial.legendre.leggauss(wl['nG'])
    s = 0.5 * (wl['lam0'] + wl['lam1']) + 0.5 * (wl['lam1'] - wl['lam0']) * xg
    lams = wl['lam0'] + wl['dlam'] * np.arange(int(round((wl['lam1'] - wl['lam0']) / wl['dlam'])) + 1)
    lam_t = 440e-9
    len_abs = base.extra['len_abs']
    src = np.array([0.25 * lam_t + len_abs - 1j * lam_t, 0.0, 0.0])
    origin = (0.5 * base.dx, -0.5 * (base.M - 1) * base.dx, -0.5 * (base.N - 1) * base.dx)
    x_thru = base.L - int(round((len_abs + 450e-9) / base.dx))
    # voxFromEdge = round(delW/2)/res of the example rounds delW/2 in metres: 0
    m_edge = 0
    log(f'MOR: samples {np.round(s * 1e9, 2)} nm, order {wl["order"]}, {lams.size} swept wavelengths, grid '
        f'{base.shape}, unknowns {base.n} (grating) {straight.n} (straight)')
    out = {}
    t_all = time.perf_counter()
    models = {}
    for name, prob0 in (('grating', base), ('straight', straight)):
        t0 = time.perf_counter()
        ops, precs, fops = [], [], []
        for lam in s:
            prob = dataclasses.replace(prob0, lam=float(lam))
            ops.append(vie.NOperator.assembled(prob.shape, prob.dx, prob.ko, None, prob.Mc, prob.idx))
            P, _ = make_preconditioner(prob, wl, vie, configs)
            precs.append(P)
            fops.append(vie.NOperator.for_fields(prob.shape, dx=prob.dx, ko=prob.ko))
        ctx.synchronize()
        t_setup = time.perf_counter() - t0
        sw = mor.SpectrumSweep(s, fops, prob0.Mc, prob0.idx, prob0.gram, origin, prob0.dx, [0, 1, 0], src,
                               0.5 * lam_t, x_thru, m_edge, n_ext=prob0.n_ext)
        t0 = time.perf_counter()
        rhs = [sw.rhs_copy(lam) for lam in s]
        rm = mor.ReducedModel(ops, s, rhs, order=wl['order'], precs=precs, tol=wl['tol'], restart=wl['restart'],
                              inner_tol=wl['inner_tol'], inner_maxit=wl['inner_maxit'], max_outer=wl['max_outer'],
                              log=log)
        ctx.synchronize()
        t_model = time.perf_counter() - t0
        del rhs
        t0 = time.perf_counter()
        power = np.array([sw.point(lam, rm) for lam in lams])
        ctx.synchronize()
        t_sweep = time.perf_counter() - t0
        log(f'  {name}: setup {t_setup:.2f} s, reduced model (K={rm.K}) {t_model:.2f} s '
            f'{ {k: round(v, 2) for k, v in rm.times.items()} }, sweep {t_sweep:.2f} s '
            f'({t_sweep / lams.size * 1e3:.2f} ms/wavelength)')
        models[name] = (prob0, ops, precs, sw, rm)
        out[name] = dict(setup_s=round(t_setup, 3), model_s=round(t_model, 3), sweep_s=round(t_sweep, 3),
                         ms_per_wavelength=round(t_sweep / lams.size * 1e3, 3), K=rm.K,
                         solves=len(rm.solve_stats), all_converged=all(f == 0 for f, _, _ in rm.solve_stats),
                         inner_iterations_mean=float(np.mean([it for _, it, _ in rm.solve_stats])),
                         stage_s={k: round(v, 3) for k, v in rm.times.items()}, power=power)
    t_total = time.perf_counter() - t_all
    tn = out['grating']['power'] / out['straight']['power']
    tn = tn / tn.max()
    # accuracy: full solves of the grating at swept wavelengths off the samples
    prob0, ops, precs, sw, rm = models['grating']
    checks = []
    for lam in wl['check']:
        i = int(np.argmin(np.abs(lams - lam)))
        lam = float(lams[i])
        prob = dataclasses.replace(prob0, lam=lam)
        op = vie.NOperator.assembled(prob.shape, prob.dx, prob.ko, None, prob.Mc, prob.idx)
        P, _ = make_preconditioner(prob, wl, vie, configs)
        b = sw.rhs(lam)
        xr = mor.vec_lincomb(ctx, [rm.solve(lam, b)[0]], [[1.0]], sw.n)[0]
        # residual of the reduced solution under the full operator at lam
        y = vie.DeviceArray(ctx, sw.n * 16)
        op.apply_device(xr.ptr.value, y.ptr.value)
        r = mor.vec_lincomb(ctx, [b, y], [[1.0], [-1.0]], sw.n)[0]
        rr = float(np.sqrt(mor.vec_dots(ctx, [r], [r], sw.n)[0, 0].real / mor.vec_dots(ctx, [b], [b], sw.n)[0, 0].real))
        t0 = time.perf_counter()
        info = L_solve(op, P, b, sw.d_x, dict(wl, tol=wl['check_tol']))
        t_full = time.perf_counter() - t0
        sw.field(lam)
        p_full = sw.power()
        d = mor.vec_lincomb(ctx, [sw.d_x, xr], [[1.0], [-1.0]], sw.n)[0]
        xrel = float(np.sqrt(mor.vec_dots(ctx, [d], [d], sw.n)[0, 0].real /
                             mor.vec_dots(ctx, [sw.d_x], [sw.d_x], sw.n)[0, 0].real))
        p_red = float(out['grating']['power'][i])
        checks.append(dict(lam_nm=round(lam * 1e9, 2), reduced_residual=rr, solution_rel_diff=xrel,
                           power_rel_diff=abs(p_red - p_full) / p_full, full_solve_s=round(t_full, 3),
                           full_relres=info['relres'], full_flag=info['flag']))
        log(f'  check {lam * 1e9:.1f} nm: reduced residual {rr:.2e}, |x_red - x|/|x| {xrel:.2e}, power rel diff '
            f'{checks[-1]["power_rel_diff"]:.2e}, full solve {t_full:.2f} s')
        del op, P
    t_full_mean = float(np.mean([c['full_solve_s'] for c in checks]))
    direct = 2 * lams.size * t_full_mean
    for k in ('grating', 'straight'):
        out[k].pop('power')
    line = {
        'metric': 'transmission spectrum wavelengths/s by model order reduction (grating + straight reference)',
        'value': round(lams.size / t_total, 3), 'unit': 'wavelengths/s', 'n_gpus': 1, 'steps': int(lams.size),
        'warmup': 0, 'ms_per_step': round(t_total / lams.size * 1e3, 3), 'higher_is_better': True,
        'scaling': 'weak', 'vs_baseline': None,
        'dtype': 'c64 inner / c128 outer solves, c128 reduced model and fields',
        'data': 'synthetic (device-assembled Green tensors, generated geometry)',
        'config': {'workload': args.workload, 'description': wl['desc'], 'unknowns': base.n,
                   'samples_nm': [round(v * 1e9, 3) for v in s], 'order': wl['order'], 'tol': wl['tol'],
                   'check_tol': wl['check_tol'], 'restart': wl['restart'],
                   'recycle_width': wl.get('recycle_width', 0),
                   'wavelengths': int(lams.size), 'parallelism': 'single GPU'},
        'total_s': round(t_total, 3), 'stages': out, 'checks': checks,
        'direct_full_solves_estimate_s': round(direct, 1),
        'speedup_vs_direct_full_solves': round(direct / t_total, 2),
        'normalised_transmission': {'min': float(tn.min()), 'argmin_nm': round(float(lams[np.argmin(tn)]) * 1e9, 2),
                                    'sampled_every_10nm': [round(float(v), 4) for v in tn[::100]]},
    }
    print(json.dumps(line), flush=True)


def run_dr(args, wl, log):
    """pgmresDR against pgmres at the same restart length (FP64, right
    preconditioner): iterations, time and true residual of both."""
    from voxphoton_amd import vie, configs
    from voxphoton_amd import _lib as L
    from voxphoton_amd.krylov_dr import pgmresDR
    ctx = vie.default_context()
    prob = configs.config_problem(wl['base'])
    op = vie.NOperator.assembled(prob.shape, prob.dx, prob.ko, None, prob.Mc, prob.idx, precisions=('fp64',))
    P, pdesc = make_preconditioner(prob, wl, vie, configs)
    b = prob.gaussian_rhs('y')
    log(f'deflated restarts: grid {prob.shape}, unknowns {prob.n}, restart {wl["restart"]}, ritz {wl["ritz"]}')
    res = {}
    for name in ('pgmres', 'pgmresDR'):
        ctx.synchronize()
        t0 = time.perf_counter()
        if name == 'pgmres':
            x, flag, relres, it, rv = vie.pgmres(op, b, wl['restart'], wl['tol'], wl['maxit'], R1=P)
        else:
            x, flag, relres, it, rv = pgmresDR(op, b, wl['restart'], wl['tol'], wl['maxit'], wl['ritz'], R1=P)
        ctx.synchronize()
        dt = time.perf_counter() - t0
        res[name] = dict(seconds=round(dt, 3), iterations=int(len(rv) - 1), flag=int(flag), relres=float(relres),
                         ms_per_iteration=round(dt / max(len(rv) - 1, 1) * 1e3, 3))
        log(f'  {name}: {len(rv) - 1} iterations, {dt:.2f} s, flag {flag}, relres {relres:.2e}')
    line = {
        'metric': 'FP64 solve time with deflated vs plain restarts', 'value': res['pgmresDR']['seconds'],
        'unit': 's', 'n_gpus': 1, 'steps': 1, 'warmup': 0, 'ms_per_step': res['pgmresDR']['seconds'] * 1e3,
        'higher_is_better': False, 'scaling': 'weak', 'vs_baseline': None, 'dtype': 'c128',
        'data': 'synthetic (device-assembled Green tensor, generated geometry)',
        'config': {'workload': args.workload, 'description': wl['desc'], 'unknowns': prob.n,
                   'restart': wl['restart'], 'ritz': wl['ritz'], 'tol': wl['tol'], 'preconditioner': pdesc,
                   'parallelism': 'single GPU'},
        'results': res,
        'iteration_ratio': round(res['pgmres']['iterations'] / max(res['pgmresDR']['iterations'], 1), 3),
    }
    print(json.dumps(line), flush=True)


def L_solve(op, P, b, x, wl):
    """Mixed-precision device solve of op x = b (right preconditioner P)."""
    from voxphoton_amd import _lib as L
    from voxphoton_amd.vie import _info
    info, _ = _info(wl['max_outer'] + 2)
    L.check(L.lib.vp_solve_ir(op.handle, P.handle, L.VP_PREC_RIGHT, b.ptr, x.ptr, L.VP_MEM_DEVICE, wl['tol'],
                              wl['restart'], wl['inner_tol'], wl['inner_maxit'], wl['max_outer'], ctypes.byref(info)))
    return dict(flag=info.flag, relres=info.relres)


def serve_requests(wl):
    """The seeded request stream: (arrival s, kind, size, wavelength)."""
    rng = np.random.default_rng(wl['seed'])
    t = np.cumsum(rng.exponential(1.0 / wl['rate'], wl['nreq']))
    out = []
    for i in range(wl['nreq']):
        u = rng.random()
        lam = float(rng.uniform(1500e-9, 1600e-9))
        if u < 0.4:
            out.append((float(t[i]), 'waveguide', float(rng.uniform(2e-6, 8e-6)), lam))
        elif u < 0.8:
            out.append((float(t[i]), 'bragg', int(rng.integers(20, 151)), lam))
        else:
            out.append((float(t[i]), 'disk', float(rng.uniform(2e-6, 5e-6)), lam))
    return out


def serve_problem(kind, size, lam, configs):
    dx = 20e-9
    # Preconditioners measured per kind: waveguides and gratings converge
    # fastest with the face-averaged form exact along y; disks with the
    # core-contrast 2-level form on the disk's box, undamped (152-189 inner
    # iterations for radii 2.2-4.8 um, against 468-1071 with a damped kernel
    # and GMRES(200)).
    if kind == 'waveguide':
        prob = configs.straight_waveguide(L=int(round(size / dx)), M=48, N=24, dx=dx, lam=lam)
        pc = dict(prec='face', axis=1, damping=0.0)
    elif kind == 'bragg':
        prob = configs.bragg_grating(periods=size, dx=dx, lam=lam, M=50, N=22)
        pc = dict(prec='face', axis=1, damping=0.0)
    else:
        prob = configs.disk_resonator(radius=size, dx=dx, lam=lam)
        pc = dict(prec='scalar', damping=0.0)
    return configs.smoothen(prob), pc


def run_serve(args, wl, log):
    """Serving stream through the request server (continuous batching).

    Client setup (geometry, operator, preconditioner, right-hand side) runs
    on a pool of setup threads, each with its own context and stream, from
    each request's arrival; the main thread submits requests as their setup
    finishes and collects finished requests, so setup overlaps both other
    setups and the server's solves on the device.
    Latency: submit -> finish (queue + solve) and arrival -> finish (with
    setup and any client backlog)."""
    import concurrent.futures
    import threading
    from voxphoton_amd import vie, configs
    reqs = serve_requests(wl)
    if wl.get('fft_sizes'):
        # every request has its own box: the shortest FFT sizes instead of
        # timing candidates on a device that is busy with other solves
        os.environ['VP_FFT_SIZES'] = wl['fft_sizes']
    ctx = vie.default_context()
    ctx.set_recycle_width(wl.get('recycle_width', -1))
    srv = vie.Server(int(wl['budget_gb'] * (1 << 30)), wl['max_active'], ctx=ctx)
    live = {}
    done = []
    t_setup = []
    tls = threading.local()
    contexts = []
    ctx_mu = threading.Lock()

    def setup(i, t_arr, kind, size, lam):
        c = getattr(tls, 'ctx', None)
        if c is None:
            c = tls.ctx = vie.Context(ctx.device)
            with ctx_mu:
                contexts.append(c)
        t0 = time.perf_counter()
        prob, pc = serve_problem(kind, size, lam, configs)
        op = vie.NOperator.assembled(prob.shape, prob.dx, prob.ko, None, prob.Mc, prob.idx, ctx=c)
        P, _ = make_preconditioner(prob, pc, vie, configs, ctx=c)
        b = prob.gaussian_rhs('y')
        db = vie.DeviceArray(c, b.nbytes)
        dxv = vie.DeviceArray(c, b.nbytes)
        db.upload(b)
        c.synchronize()
        return dict(i=i, kind=kind, size=size, n=prob.n, t_arr=t_arr, pc=pc, op=op, P=P, db=db, dxv=dxv,
                    setup=time.perf_counter() - t0)

    def collect():
        for tk in list(live):
            fin, info = srv.poll(tk)
            if fin:
                info = srv.wait(tk)
                rec = live.pop(tk)
                rec['info'] = info
                rec['t_done'] = time.perf_counter() - t_start
                rec.pop('keep', None)
                done.append(rec)

    pool = concurrent.futures.ThreadPoolExecutor(max_workers=wl['setup_workers'])
    pending = []
    nxt = 0
    t_start = time.perf_counter()
    try:
        while nxt < len(reqs) or pending or live:
            now = time.perf_counter() - t_start
            while nxt < len(reqs) and reqs[nxt][0] <= now:
                pending.append(pool.submit(setup, nxt, *reqs[nxt]))
                nxt += 1
            for f in [f for f in pending if f.done()]:
                pending.remove(f)
                r = f.result()
                pc = r['pc']
                t_setup.append(r['setup'])
                t_sub = time.perf_counter() - t_start
                tk = srv.submit(r['op'], r['P'], r['db'], r['dxv'], tol=wl['tol'], restart=pc.get('restart', wl['restart']),
                                inner_tol=wl['inner_tol'], inner_maxit=pc.get('inner_maxit', wl['inner_maxit']),
                                max_outer=pc.get('max_outer', wl['max_outer']), side='right')
                live[tk] = dict(i=r['i'], kind=r['kind'], size=r['size'], n=r['n'], t_arr=r['t_arr'], t_sub=t_sub,
                                keep=(r['op'], r['P'], r['db'], r['dxv']))
                log(f'  request {r["i"]} {r["kind"]} {r["size"]:.3g} n={r["n"]} arrival {r["t_arr"]:.3f} s submitted '
                    f'{t_sub:.3f} s (setup {r["setup"]:.2f} s), in flight {len(live)}, server {srv.stats()}')
            collect()
            time.sleep(0.001)
        pool.shutdown()
    except BaseException:
        # a client that leaves on an exception first stops the server, whose
        # executors may still be solving with the requests' operators and vectors
        pool.shutdown(wait=True)
        srv.close()
        raise
    t_total = max(r['t_done'] for r in done) - reqs[0][0]
    srv.close()
    lat = np.array([r['info']['seconds'] for r in done]) * 1e3
    e2e = np.array([r['t_done'] - r['t_arr'] for r in done]) * 1e3
    worst = max(r['info']['relres'] for r in done)
    kinds = {k: sum(1 for r in done if r['kind'] == k) for k in ('waveguide', 'bragg', 'disk')}
    line = {
        'metric': METRIC, 'value': round(len(done) / t_total, 4), 'unit': 'solves/s', 'n_gpus': 1,
        'steps': len(done), 'warmup': 0, 'ms_per_step': round(t_total / len(done) * 1e3, 3),
        'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None,
        'dtype': PREC_DTYPE,
        'data': 'synthetic (seeded request stream; device-assembled Green tensors)',
        'config': {'workload': args.workload, 'description': wl['desc'], 'requests': len(done), 'mix': kinds,
                   'unknowns_min': int(min(r['n'] for r in done)), 'unknowns_max': int(max(r['n'] for r in done)),
                   'tol': wl['tol'], 'restart': wl['restart'], 'recycle_width': wl.get('recycle_width', 0),
                   'hbm_budget_gb': wl['budget_gb'],
                   'max_active': wl['max_active'], 'parallelism': 'single GPU'},
        'relres_worst': worst, 'all_converged': all(r['info']['flag'] == 0 for r in done),
        'latency_ms': {'p50': round(float(np.percentile(lat, 50)), 3), 'p99': round(float(np.percentile(lat, 99)), 3),
                       'definition': 'submit to finish (queue + solve)'},
        'latency_e2e_ms': {'p50': round(float(np.percentile(e2e, 50)), 3),
                           'p99': round(float(np.percentile(e2e, 99)), 3),
                           'definition': 'scheduled arrival to finish (client setup, backlog, queue, solve)'},
        'setup_s_mean': round(float(np.mean(t_setup)), 3),
        'inner_iterations_mean': float(np.mean([r['info']['inner_iters'] for r in done])),
        'per_kind': {k: {'requests': sum(1 for r in done if r['kind'] == k),
                         'converged': sum(1 for r in done if r['kind'] == k and r['info']['flag'] == 0),
                         'latency_ms_p50': round(float(np.percentile(
                             [r['info']['seconds'] * 1e3 for r in done if r['kind'] == k] or [0.0], 50)), 1),
                         'inner_iterations_mean': round(float(np.mean(
                             [r['info']['inner_iters'] for r in done if r['kind'] == k] or [0.0])), 1)}
                     for k in ('waveguide', 'bragg', 'disk')},
    }
    print(json.dumps(line), flush=True)


def traffic_probe(op, P, ctx, n, vie, L, log):
    """Kernels for the HBM-traffic counters: torch's device copy of 1 GiB
    (read 1 GiB, write 1 GiB: the calibration of FETCH_SIZE / WRITE_SIZE),
    then three FP32 matvecs and three FP32 preconditioner applies."""
    import torch
    src = torch.ones(1 << 28, dtype=torch.float32, device='cuda')
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    x32 = vie.DeviceArray(ctx, n * 8)
    y32 = vie.DeviceArray(ctx, n * 8)
    x32.upload((np.random.default_rng(1550).standard_normal(n) + 0j).astype(np.complex64))
    for _ in range(3):
        L.check(L.lib.vp_op_apply(op.handle, 0, x32.ptr, y32.ptr, L.VP_MEM_DEVICE, L.VP_FP32))
    L.check(L.lib.vp_prec_set_fp32_scale(P.handle, 1.0))
    for _ in range(3):
        L.check(L.lib.vp_prec_apply(P.handle, x32.ptr, y32.ptr, L.VP_MEM_DEVICE, L.VP_FP32))
    ctx.synchronize()
    log('traffic probe done: 3 x 1 GiB copies, 3 FP32 matvecs, 3 FP32 preconditioner applies')


def field_block(prob, vie, L, ctx, xsol, log):
    """Steps either side of the solve on the workload grid (FP64, device
    vectors, HIP-synchronised wall time):

    * the complex-point-source beam of example_Bragg_1500nm_to_1600nm.m:62-65
      (440 nm, source lambda/4 past the absorber, imaginary offset -lambda,
      cutoff lambda/2, y-polarised) at every voxel centre;
    * the total field Einc + (N(J) - Gram J)/(Gram j w eps0) of the solution
      (E_field_Nop_const) through a field operator over the whole grid, with
      the workload's Gaussian incident field;
    * the transmitted power through the cross-section one wavelength inside
      the far absorber (example_Bragg_1500nm_to_1600nm.m:108-113).
    """
    Lg, Mg, Ng = prob.shape
    V = Lg * Mg * Ng
    dx = prob.dx
    lam_t = 440e-9
    len_abs = prob.extra.get('len_abs', 2e-6)
    src = np.array([lam_t / 4 + len_abs - 1j * lam_t, Mg * dx / 2, Ng * dx / 2], dtype=complex)
    Jd = np.array([0, 1, 0], dtype=complex)
    origin = np.array([0.5 * dx, 0.5 * dx, 0.5 * dx])
    Ed = vie.DeviceArray(ctx, 3 * V * 16)

    def dgf():
        L.check(L.lib.vp_field_dgf_grid(ctx.handle, L.ptr(Jd), L.ptr(src), 1, L.ptr(origin), dx, Lg, Mg, Ng,
                                        float(prob.ko), 0.5 * lam_t, Ed.ptr, L.VP_MEM_DEVICE))
    dgf()
    ctx.synchronize()
    t = time.perf_counter()
    for _ in range(5):
        dgf()
    ctx.synchronize()
    t_dgf = (time.perf_counter() - t) / 5
    t = time.perf_counter()
    fop = vie.NOperator.for_fields(prob.shape, dx=dx, ko=prob.ko, ctx=ctx)
    t_setup = time.perf_counter() - t
    xs = xsol.download(np.empty(prob.n, dtype=complex))
    J = np.zeros(3 * V, dtype=complex)
    J[prob.idx - 1] = xs
    Jdev = vie.DeviceArray(ctx, J.nbytes)
    Jdev.upload(J)
    Einc = vie.DeviceArray(ctx, J.nbytes)
    Einc.upload(np.ascontiguousarray(prob.gaussian_field('y').reshape(-1, order='F')))
    Et = vie.DeviceArray(ctx, J.nbytes)

    def total():
        L.check(L.lib.vp_field_total(fop.handle, Jdev.ptr, Einc.ptr, float(prob.freq), Et.ptr, L.VP_MEM_DEVICE))
    total()
    ctx.synchronize()
    t = time.perf_counter()
    for _ in range(3):
        total()
    ctx.synchronize()
    t_tot = (time.perf_counter() - t) / 3
    E = Et.download(np.empty(3 * V, dtype=complex)).reshape(Lg, Mg, Ng, 3, order='F')
    x_thru = Lg - int(round((len_abs + 450e-9) / dx))
    p_thru = vie.power_through(E, x_thru, 0)
    p_in = vie.power_through(E, int(round((len_abs + 450e-9) / dx)), 0)
    fop.close()
    log(f'fields: source beam {t_dgf * 1e3:.2f} ms over {V} voxels, field operator {t_setup:.2f} s, '
        f'total field {t_tot * 1e3:.2f} ms, power through x={x_thru}: {p_thru:.4e} (x={Lg - x_thru}: {p_in:.4e})')
    return dict(points=V, source_beam_ms=round(t_dgf * 1e3, 3), field_operator_setup_s=round(t_setup, 3),
                total_field_ms=round(t_tot * 1e3, 3), power_thru=p_thru, power_near_source_side=p_in,
                x_thru=x_thru)


def make_preconditioner(prob, wl, vie, configs, ctx=None):
    """2-level circulant preconditioner of the workload on its core box.

    'face': contrast averaged over the two circulant axes, exact along
    wl['axis'] (1 = y, 2 = z); 'scalar': the core contrast on the box.
    """
    n = prob.n
    if wl['prec'] == 'composite':
        # block Jacobi over the sub-boxes (prec_fun_disk_idx.m): a 2-level
        # circulant on each box, embedded on its unknowns.  Part forms:
        # 'face_y' face-averaged contrast, exact across the box's y extent (the
        # waveguide form of Bragg_MOR_script.m:61-72); 'core' the core contrast,
        # exact along z; 'field' the box's own contrast array folded into the
        # Toeplitz tensor, as resonator_MOR_scripts.m:165-170 builds the disk square
        forms = wl.get('part_forms') or prob.extra.get('part_forms') or ('core',) * len(prob.extra['parts'])
        parts = []
        for box, form in zip(prob.extra['parts'], forms):
            if form == 'face_y':
                shape, bidx, sub, face = configs.face_box(prob, 1, box=box)
                q = vie.Preconditioner.level2_face(shape, prob.dx, prob.ko, face, bidx, damping=wl['damping'],
                                                   ctx=ctx)
            else:
                shape, bidx, sub = prob.box_problem(box)
                mc = prob.Mc[box] if form == 'field' else prob.core_contrast()
                q = vie.Preconditioner.level2_assembled(shape, prob.dx, prob.ko, mc, 1.0, bidx,
                                                        damping=wl['damping'], ctx=ctx)
            q.set_embedding(n, sub)
            q.box_shape = tuple(int(v) for v in shape)
            parts.append(q)
        P = vie.Preconditioner.composite(parts, ctx=ctx)
        P.box_shape = parts[-1].box_shape
        desc = (f'block-Jacobi composite of 2-level circulants on the boxes '
                f'{[f"{q.box_shape} {f}" for q, f in zip(parts, forms)]}, damping {wl["damping"]}')
        return P, desc
    if wl['prec'] == 'face':
        shape, pidx, sub, face = configs.face_box(prob, wl['axis'])
        P = vie.Preconditioner.level2_face(shape, prob.dx, prob.ko, face, pidx, damping=wl['damping'], ctx=ctx)
        desc = (f'2-level circulant, face-averaged contrast, exact along {"yz"[wl["axis"] - 1]}, box {shape}, '
                f'damping {wl["damping"]}')
    else:
        shape, bidx, sub = prob.box_problem()
        P = vie.Preconditioner.level2_assembled(shape, prob.dx, prob.ko, prob.core_contrast(), 1.0, bidx,
                                                damping=wl['damping'], ctx=ctx)
        desc = f'2-level circulant, core contrast, box {shape}, damping {wl["damping"]}'
    P.set_embedding(n, sub)
    P.box_shape = tuple(int(v) for v in shape)
    return P, desc


def cpu_baseline(prob, vie, fp64, log, pc_box=None, restart=100):
    """CPU baseline: the FP64 oracle (NumPy/SciPy restatement of the
    reference's mv_AN_const, chan_2_mvp_for_parallel_idx and pgmres_iter) on
    the host cores, on a bounded sample of the workload:

    * one full matvec of the seeded vector (its output is returned, so the
      caller checks the device matvecs against it);
    * the 2-level preconditioner apply on a slab of kx frequencies of the
      preconditioner box with blocks of the box's size, scaled to all kx;
    * one classical Gram-Schmidt step against 8 basis vectors, scaled to the
      mean width (restart + 1) / 2 of a GMRES(restart) cycle.

    Extrapolated to one plain FP64 GMRES solve with the GPU-measured
    iteration count (plus the final residual matvec)."""
    from oracle import vie as ovie, precond as opre
    cores = min(len(os.sched_getaffinity(0)), 16)
    G = vie.assemble_toeplitz(prob.shape, prob.dx, prob.ko)
    fN = ovie.fN_from_toeplitz(G, workers=cores)
    del G
    rng = np.random.default_rng(1550)
    x = rng.standard_normal(prob.n) + 1j * rng.standard_normal(prob.n)
    Mr = np.ones(prob.shape, complex)
    t = time.perf_counter()
    y_ref = ovie.mv_AN_const(x, fN, Mr, prob.Mc, prob.gram, 'notransp', prob.idx, workers=cores)
    t_mv = time.perf_counter() - t
    del fN
    log(f'cpu oracle matvec {t_mv:.2f} s on {cores} threads')
    t_pc = 0.0
    pc_note = 'no preconditioner'
    if pc_box is not None:
        Lc, Mc_, Nc = pc_box
        Ls = min(Lc, 32)
        rs = np.random.default_rng(7)
        nb = 3 * Nc
        B = (rs.standard_normal((Ls, Mc_, nb, nb)) + 1j * rs.standard_normal((Ls, Mc_, nb, nb))) / nb
        v = rs.standard_normal(3 * Ls * Mc_ * Nc) + 1j * rs.standard_normal(3 * Ls * Mc_ * Nc)
        t = time.perf_counter()
        opre.chan_2_mvp(B, v, Ls, Mc_, Nc)
        t_pc = (time.perf_counter() - t) * Lc / Ls
        pc_note = f'2-level apply on {Ls} of {Lc} kx of the {Lc}x{Mc_}x{Nc} box ({nb}x{nb} blocks), scaled'
        del B, v
    j_s = 8
    jm = (restart + 1) / 2
    Vb = np.ones((j_s, prob.n), dtype=complex)
    w = x.copy()
    t = time.perf_counter()
    h = Vb.conj() @ w
    w -= Vb.T @ h
    np.linalg.norm(w)
    t_cgs = (time.perf_counter() - t) * jm / j_s
    del Vb, w
    log(f'cpu oracle preconditioner apply {t_pc:.2f} s (sampled), Gram-Schmidt step {t_cgs:.2f} s (sampled)')
    n_it = None if fp64 is None else fp64['iterations']
    per_solve = None if n_it is None else n_it * (t_mv + t_pc + t_cgs) + t_mv
    value = None if per_solve is None else 1.0 / per_solve
    cpu = dict(value=value, unit='solves/s', cores=cores, kind='port',
               sample=f'FP64 oracle on {cores} host threads: one full matvec ({t_mv:.2f} s); {pc_note} '
                      f'({t_pc:.2f} s); one Gram-Schmidt step against 8 vectors scaled to {jm:.1f} '
                      f'({t_cgs:.2f} s); extrapolated to the {n_it} iterations of one plain FP64 GMRES solve '
                      f'plus its final residual matvec',
               matvec_s=round(t_mv, 3), precond_apply_s=round(t_pc, 3), gram_schmidt_s=round(t_cgs, 3),
               seconds_per_solve=None if per_solve is None else round(per_solve, 2))
    return cpu, y_ref


if __name__ == '__main__':
    main()
