"""The benchmark configurations at their real sizes, through the C ABI,
against the FP64 oracle.

* straight waveguide 200x48x24 at 20 nm: FP64 / FP32 matvec of the seeded
  N(0,1) + iN(0,1) vector against the oracle's mv_AN_const (1e-12 / 1e-5),
  and plain FP64 GMRES(50) solves with the 2-level and with the 1-level
  preconditioner to 1e-6 whose true residual the oracle recomputes;
* Bragg grating 3600x100x44 at 10 nm: the same matvec checks, the
  mixed-precision solve's true residual by the oracle, and its difference
  from the plain FP64 solve;
* the 20 nm grating 1800x50x22: two wavelengths x two right-hand sides as
  one batched solve, oracle residuals, batched = solo;
* the serving stream: seeded requests (waveguides, gratings, a disk) through
  the request server, oracle residuals, each result equal to the same
  request served alone, and a request with a NaN right-hand side that fails
  without changing the others.
* the microdisk with a bus at 15 nm, 1400x1500x30 (64.9 M unknowns): the
  matvec against direct Toeplitz sums at sampled rows, FP32 against FP64,
  and against the reference's 2L x 2M x 2N embedding; the mixed solve to
  1e-6 (flexible refinement), whose reported residual is the true FP64
  residual.

The oracle's tensor is the device-assembled Toeplitz tensor (its cubature is
checked against the oracle's at small grids); everything after it (the
circulant embedding, FFTs, contrast, indexing, residuals) is the oracle's.
"""
import ctypes
import dataclasses
import os
import time

import numpy as np
import pytest

from oracle import vie as ovie
from tests.helpers import rel

pytestmark = pytest.mark.gpu

WORKERS = min(len(os.sched_getaffinity(0)), 16)
SEED_VECTOR = 1550


class OracleOperator:
    """The oracle's JVIE-II operator A = Gram - Mc N on a problem's unknowns."""

    def __init__(self, prob, ctx):
        from voxphoton_amd import vie
        G = vie.assemble_toeplitz(prob.shape, prob.dx, prob.ko, ctx=ctx)
        self.fN = ovie.fN_from_toeplitz(G, workers=WORKERS)
        del G
        self.prob = prob
        self.Mr = np.ones(prob.shape, complex)
        self.Mc = prob.Mc
        self.idx = prob.idx

    def __call__(self, x):
        return ovie.mv_AN_const(x, self.fN, self.Mr, self.Mc, self.prob.gram, 'notransp', self.idx,
                                workers=WORKERS)

    def true_relres(self, x, b):
        return float(np.linalg.norm(b - self(x)) / np.linalg.norm(b))


def _seeded(n):
    rng = np.random.default_rng(SEED_VECTOR)
    return rng.standard_normal(n) + 1j * rng.standard_normal(n)


def _say(*a):
    print(*a, flush=True)


def _check_matvec(op, oracle, n, name):
    x = _seeded(n)
    t = time.perf_counter()
    ref = oracle(x)
    t_ref = time.perf_counter() - t
    e64 = rel(op.apply(x), ref)
    e32 = rel(op.apply(x, prec='fp32'), ref)
    _say(f'{name}: matvec vs oracle fp64 {e64:.2e} fp32 {e32:.2e} (oracle {t_ref:.1f} s, n={n})')
    assert e64 < 1e-12
    assert e32 < 1e-5


# ---------------------------------------------------------------------------
# straight waveguide, 200 x 48 x 24 at 20 nm, plain FP64
# ---------------------------------------------------------------------------

@pytest.fixture(scope='module')
def wg20(vp_ctx):
    from voxphoton_amd import configs, vie
    prob = configs.straight_waveguide()
    assert prob.shape == (200, 48, 24) and prob.n == 3 * 200 * 25 * 11
    op = vie.NOperator.assembled(prob.shape, prob.dx, prob.ko, None, prob.Mc, prob.idx, ctx=vp_ctx)
    return prob, op, OracleOperator(prob, vp_ctx)


def test_straight_wg_matvec_matches_oracle(wg20):
    prob, op, oracle = wg20
    _check_matvec(op, oracle, prob.n, 'straight WG 20 nm')


def test_straight_wg_fp64_gmres50_true_residual(wg20, vp_ctx):
    from voxphoton_amd import configs, vie
    import bench
    prob, op, oracle = wg20
    wl = dict(bench.WORKLOADS['straight_wg_20nm'])
    P, desc = bench.make_preconditioner(prob, wl, vie, configs)
    b = prob.gaussian_rhs('y')
    t = time.perf_counter()
    x, flag, relres, it, rv = vie.pgmres(op, b, 50, 1e-6, 40, R1=P)
    dt = time.perf_counter() - t
    true = oracle.true_relres(x, b)
    _say(f'straight WG FP64 GMRES(50), {desc}: flag {flag}, {len(rv) - 1} iterations, {dt:.2f} s, '
         f'relres {relres:.3e}, oracle relres {true:.3e}')
    assert flag == 0
    assert true <= 1e-6 * (1 + 1e-6)
    assert abs(true - relres) <= 1e-9 * max(true, 1e-300) + 1e-15


@pytest.mark.timeout(600)
def test_straight_wg_fp64_gmres50_level1_true_residual(wg20, vp_ctx):
    # the reference's 1-level form (level_1_parallel_func, blocks of
    # 3MN = 825 on the core box, inverted by rocSOLVER) as the right
    # preconditioner of FP64 GMRES(50)
    from voxphoton_amd import vie
    prob, op, oracle = wg20
    shape, bidx, sub = prob.box_problem()
    t = time.perf_counter()
    P = vie.Preconditioner.level1_assembled(shape, prob.dx, prob.ko, prob.core_contrast(), 1.0, bidx,
                                            ctx=vp_ctx).set_embedding(prob.n, sub)
    t_setup = time.perf_counter() - t
    b = prob.gaussian_rhs('y')
    t = time.perf_counter()
    x, flag, relres, it, rv = vie.pgmres(op, b, 50, 1e-6, 40, R1=P)
    dt = time.perf_counter() - t
    true = oracle.true_relres(x, b)
    _say(f'straight WG FP64 GMRES(50), 1-level on box {shape} (setup {t_setup:.2f} s): flag {flag}, '
         f'{len(rv) - 1} iterations, {dt:.2f} s, relres {relres:.3e}, oracle relres {true:.3e}')
    assert flag == 0
    assert true <= 1e-6 * (1 + 1e-6)


# ---------------------------------------------------------------------------
# Bragg grating, 3600 x 100 x 44 at 10 nm, mixed precision
# ---------------------------------------------------------------------------

@pytest.fixture(scope='module')
def bragg10(vp_ctx):
    from voxphoton_amd import configs, vie
    prob = configs.bragg_grating()
    assert prob.shape == (3600, 100, 44)
    op = vie.NOperator.assembled(prob.shape, prob.dx, prob.ko, None, prob.Mc, prob.idx, ctx=vp_ctx)
    t = time.perf_counter()
    oracle = OracleOperator(prob, vp_ctx)
    _say(f'Bragg 10 nm oracle tensor in {time.perf_counter() - t:.1f} s')
    return prob, op, oracle


@pytest.mark.timeout(900)
def test_bragg10_matvec_matches_oracle(bragg10):
    prob, op, oracle = bragg10
    _check_matvec(op, oracle, prob.n, 'Bragg 10 nm')


@pytest.mark.timeout(900)
def test_bragg10_mixed_solve_true_residual_and_fp64_difference(bragg10, vp_ctx):
    from voxphoton_amd import configs, vie
    import bench
    prob, op, oracle = bragg10
    wl = dict(bench.WORKLOADS['bragg_10nm'])
    P, desc = bench.make_preconditioner(prob, wl, vie, configs)
    b = prob.gaussian_rhs('y')
    xm, info = vie.solve_ir(op, b, tol=wl['tol'], restart=wl['restart'], inner_tol=wl['inner_tol'],
                            inner_maxit=wl['inner_maxit'], max_outer=wl['max_outer'], M=P, side='right')
    x64, flag, relres64, it, rv = vie.pgmres(op, b, wl['restart'], wl['tol'], 400, R1=P)
    tm = oracle.true_relres(xm, b)
    t64 = oracle.true_relres(x64, b)
    # both residuals are <= tol, so A (xm - x64) is at most 2 tol |b|; the
    # solutions' difference is that times the conditioning of A
    dres = float(np.linalg.norm(oracle(xm - x64)) / np.linalg.norm(b))
    dx_rel = rel(xm, x64)
    _say(f'Bragg 10 nm, {desc}: mixed {info["ir_steps"]} refinement steps / {info["inner_iters"]} inner its, '
         f'relres {info["relres"]:.3e} (oracle {tm:.3e}); plain FP64 {len(rv) - 1} its, relres {relres64:.3e} '
         f'(oracle {t64:.3e}); |A(xm - x64)|/|b| {dres:.2e}; |xm - x64|/|x64| {dx_rel:.2e}')
    assert info['flag'] == 0 and flag == 0
    assert tm <= 1e-6 * (1 + 1e-6) and t64 <= 1e-6 * (1 + 1e-6)
    assert dres <= tm + t64 + 1e-12
    assert dx_rel < 1e-3


# ---------------------------------------------------------------------------
# 20 nm grating, 1800 x 50 x 22: wavelengths x right-hand sides, batched
# ---------------------------------------------------------------------------

@pytest.mark.timeout(900)
def test_bragg20_batched_sweep_slice(vp_ctx):
    from voxphoton_amd import configs, vie
    import bench
    wl = dict(bench.WORKLOADS['sweep_bragg_20nm'])
    base = configs.config_problem('bragg_20nm')
    assert base.shape == (1800, 50, 22)
    ops, precs, rhs, probs = [], [], [], []
    for lam in (1500e-9, 1600e-9):
        prob = dataclasses.replace(base, lam=lam)
        op = vie.NOperator.assembled(prob.shape, prob.dx, prob.ko, None, prob.Mc, prob.idx, ctx=vp_ctx)
        P, _ = bench.make_preconditioner(prob, wl, vie, configs)
        for pol in wl['pols']:
            ops.append(op)
            precs.append(P)
            rhs.append(prob.gaussian_rhs(pol))
            probs.append(prob)
    B = np.array(rhs)
    # the workload's recycled-space width (harmonic Ritz directions) on the operators' context
    vp_ctx.set_recycle_width(wl.get('recycle_width', -1))
    try:
        X, infos = vie.solve_ir_batched(ops, precs, B, tol=wl['tol'], restart=wl['restart'],
                                        inner_tol=wl['inner_tol'], inner_maxit=wl['inner_maxit'],
                                        max_outer=wl['max_outer'], side='right')
        # solo solve of the last problem: identical to its batched result
        Xs, infos_s = vie.solve_ir_batched(ops[-1:], precs[-1:], B[-1:], tol=wl['tol'], restart=wl['restart'],
                                           inner_tol=wl['inner_tol'], inner_maxit=wl['inner_maxit'],
                                           max_outer=wl['max_outer'], side='right')
    finally:
        vp_ctx.set_recycle_width(-1)
    oracles = {}
    for i, prob in enumerate(probs):
        if prob.lam not in oracles:
            oracles[prob.lam] = OracleOperator(prob, vp_ctx)
        true = oracles[prob.lam].true_relres(X[i], B[i])
        _say(f'batched sweep slice: lambda {prob.lam * 1e9:.0f} nm pol {wl["pols"][i % 2]}: flag {infos[i]["flag"]}, '
             f'{infos[i]["ir_steps"]} steps / {infos[i]["inner_iters"]} its, relres {infos[i]["relres"]:.3e}, '
             f'oracle {true:.3e}')
        assert infos[i]['flag'] == 0
        assert true <= wl['tol'] * (1 + 1e-6)
    assert rel(Xs[0], X[-1]) <= 1e-12
    assert infos_s[0]['relres'] == infos[-1]['relres']


# ---------------------------------------------------------------------------
# serving stream: seeded requests through the request server
# ---------------------------------------------------------------------------

SERVE_PICK = (0, 9, 19, 23)      # waveguide 4.2 um, disk r = 2.24 um, gratings of 20 and 69 periods


def _serve_setup(vp_ctx):
    from voxphoton_amd import configs, vie
    import bench
    wl = dict(bench.WORKLOADS['serve_mix'])
    stream = bench.serve_requests(wl)
    reqs = []
    for i in SERVE_PICK:
        _, kind, size, lam = stream[i]
        prob, pc = bench.serve_problem(kind, size, lam, configs)
        op = vie.NOperator.assembled(prob.shape, prob.dx, prob.ko, None, prob.Mc, prob.idx, ctx=vp_ctx)
        P, _ = bench.make_preconditioner(prob, pc, vie, configs)
        prm = dict(tol=wl['tol'], restart=pc.get('restart', wl['restart']), inner_tol=wl['inner_tol'],
                   inner_maxit=pc.get('inner_maxit', wl['inner_maxit']),
                   max_outer=pc.get('max_outer', wl['max_outer']), side='right')
        pols = ('y', 'z') if i == SERVE_PICK[0] else ('y',)
        for pol in pols:
            b = prob.gaussian_rhs(pol)
            reqs.append(dict(name=f'{kind} {size:.3g} {pol}', prob=prob, op=op, P=P, b=b, prm=prm))
    # a request sharing the first geometry whose right-hand side holds a NaN
    bad = dict(reqs[0])
    bb = reqs[0]['b'].copy()
    bb[len(bb) // 3] = np.nan
    bad.update(name='waveguide NaN', b=bb)
    reqs.insert(2, bad)
    for r in reqs:
        r['db'] = vie.DeviceArray(vp_ctx, r['b'].nbytes)
        r['db'].upload(r['b'])
    return reqs


@pytest.mark.timeout(900)
def test_serving_stream_against_oracle_and_alone(vp_ctx):
    from voxphoton_amd import vie
    reqs = _serve_setup(vp_ctx)

    def serve(batch):
        srv = vie.Server(hbm_budget_bytes=64 << 30, max_active=16, ctx=vp_ctx)
        try:
            xs = [vie.DeviceArray(vp_ctx, r['b'].nbytes) for r in batch]
            tk = [srv.submit(r['op'], r['P'], r['db'], x, **r['prm']) for r, x in zip(batch, xs)]
            infos = [srv.wait(t) for t in tk]
            st = srv.stats()
            assert st['pending'] == 0 and st['active'] == 0 and st['bytes_in_use'] == 0
        finally:
            srv.close()
        out = []
        for r, x, info in zip(batch, xs, infos):
            h = np.empty(r['b'].size, complex)
            x.download(h)
            out.append((h, info))
        return out

    t = time.perf_counter()
    mixed = serve(reqs)
    t_mix = time.perf_counter() - t
    alone = [serve([r])[0] for r in reqs]
    _say(f'serving: {len(reqs)} requests in the mix in {t_mix:.2f} s')
    oracles = {}
    for r, (x, info), (xa, ia) in zip(reqs, mixed, alone):
        if r['name'] == 'waveguide NaN':
            _say(f'  {r["name"]}: flag {info["flag"]}, relres {info["relres"]} (alone: flag {ia["flag"]})')
            assert info['flag'] == 1 and ia['flag'] == 1
            assert not np.isfinite(info['relres'])
            continue
        key = id(r['op'])
        if key not in oracles:
            oracles[key] = OracleOperator(r['prob'], vp_ctx)
        true = oracles[key].true_relres(x, r['b'])
        _say(f'  {r["name"]}: n={r["b"].size}, flag {info["flag"]}, {info["ir_steps"]} steps / '
             f'{info["inner_iters"]} its, relres {info["relres"]:.3e}, oracle {true:.3e}, '
             f'|x_mix - x_alone|/|x| {rel(x, xa):.1e}')
        assert info['flag'] == 0
        assert true <= r['prm']['tol'] * (1 + 1e-6)
        assert rel(x, xa) <= 1e-12
        assert info['relres'] == ia['relres']


# ---------------------------------------------------------------------------
# microdisk with a bus waveguide, 1400 x 1500 x 30 at 15 nm (64.9 M unknowns)
# ---------------------------------------------------------------------------

@pytest.mark.timeout(1100)
def test_microdisk15_full_size(vp_ctx, monkeypatch):
    # At this size the oracle's FFT product is out of reach on the host, so
    # the matvec is checked (a) entry by entry against direct Toeplitz sums
    # over all unknowns at sampled rows, (b) FP32 against FP64 and (c) against
    # the same operator embedded at the reference's circulant sizes
    # 2L x 2M x 2N (the product on idx does not depend on the embedding).
    # The solve runs a few refinement steps of the mixed GMRES(50) with the
    # block-Jacobi composite; its reported residual must be the true FP64
    # residual of the checked matvec.
    from voxphoton_amd import configs, vie
    import bench
    free0, total = vp_ctx.mem_info()
    t = time.perf_counter()
    prob = configs.config_problem('microdisk_15nm')
    assert prob.shape == (1400, 1500, 30)
    n = prob.n
    _say(f'microdisk 15 nm: n = {n} unknowns, geometry in {time.perf_counter() - t:.1f} s')
    t = time.perf_counter()
    op = vie.NOperator.assembled(prob.shape, prob.dx, prob.ko, None, prob.Mc, prob.idx, ctx=vp_ctx)
    info = op.info()
    _say(f'operator in {time.perf_counter() - t:.1f} s: box {info["box"]} at {info["box_origin"]}, FFT sizes '
         f'fp64 {info["fft_fp64"]} fp32 {info["fft_fp32"]}')
    x = _seeded(n)
    y64 = op.apply(x)
    y32 = op.apply(x, prec='fp32')
    e32 = rel(y32, y64)
    rms = np.linalg.norm(y64) / np.sqrt(n)
    _say(f'fp32 vs fp64 matvec {e32:.2e}')
    assert e32 < 1e-5

    # (a) direct sums at rows in the bus, the disk and on the disk's rim
    G = vie.assemble_toeplitz(info['box'], prob.dx, prob.ko, ctx=vp_ctx)
    rng = np.random.default_rng(15)
    rows = list(rng.choice(n, 5, replace=False)) + [0, n // 3 + 7, n - 1]
    t = time.perf_counter()
    ref = ovie.an_rows(G, 1.0, prob.Mc, prob.gram, prob.idx, x, rows, prob.shape)
    del G
    err64 = np.abs(y64[rows] - ref).max() / rms
    err32 = np.abs(y32[rows] - ref).max() / rms
    _say(f'direct sums at {len(rows)} rows ({time.perf_counter() - t:.1f} s): max error / rms(y) '
         f'fp64 {err64:.2e} fp32 {err32:.2e}')
    assert err64 < 1e-11
    assert err32 < 1e-4

    # (c) the reference's embedding
    monkeypatch.setenv('VP_FFT_SIZES', 'reference')
    t = time.perf_counter()
    op_ref = vie.NOperator.assembled(prob.shape, prob.dx, prob.ko, None, prob.Mc, prob.idx, precisions=('fp64',),
                                     ctx=vp_ctx)
    monkeypatch.delenv('VP_FFT_SIZES')
    assert op_ref.info()['fft_fp64'] == (2800, 3000, 60)
    e_emb = rel(op_ref.apply(x), y64)
    _say(f'embedding at 2800x3000x60 ({time.perf_counter() - t:.1f} s): relative difference {e_emb:.2e}')
    assert e_emb < 1e-12
    op_ref.close()
    del op_ref, y32

    # mixed solve to 1e-6 as the workload runs it: GMRES(100) inner solves with
    # the block-Jacobi composite inside flexible FP64 refinement; the
    # reported residual must be the true FP64 residual of the checked matvec
    wl = dict(bench.WORKLOADS['microdisk_15nm'])
    t = time.perf_counter()
    P, desc = bench.make_preconditioner(prob, wl, vie, configs)
    _say(f'{desc}: setup {time.perf_counter() - t:.1f} s')
    b = prob.gaussian_rhs('y')
    t = time.perf_counter()
    xm, sinfo = vie.solve_ir(op, b, tol=wl['tol'], restart=wl['restart'], inner_tol=wl['inner_tol'],
                             inner_maxit=wl['inner_maxit'], max_outer=wl['max_outer'], M=P, side='right',
                             outer_restart=wl['outer_restart'])
    dt = time.perf_counter() - t
    free1, _ = vp_ctx.mem_info()
    true = float(np.linalg.norm(b - op.apply(xm)) / np.linalg.norm(b))
    _say(f'mixed GMRES({wl["restart"]}), flexible refinement: {sinfo["ir_steps"]} inner solves '
         f'({sinfo["inner_iters"]} inner its, {dt:.1f} s): relres {sinfo["relres"]:.4e}, true FP64 relres '
         f'{true:.4e}, history {[float("%.3g" % v) for v in sinfo["history"]]}; device memory in use '
         f'{(total - free1) / 2**30:.1f} GiB of {total / 2**30:.0f} GiB ({(total - free0) / 2**30:.1f} GiB before)')
    assert abs(true - sinfo['relres']) <= 1e-9 * true
    assert sinfo['flag'] == 0 and true <= wl['tol']
