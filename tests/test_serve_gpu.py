"""Request server (continuous batching): every result equals the solo solve,
whatever was running beside it; the HBM budget limits admission."""
import time

import numpy as np
import pytest

from tests.helpers import rel

pytestmark = pytest.mark.gpu

PARAMS = dict(tol=1e-8, restart=20, inner_tol=1e-3, inner_maxit=3, max_outer=15, side='right')


def _requests(vp_ctx):
    """Three geometries (two waveguide lengths, one short grating), two
    wavelengths each, y- and z-polarised right-hand sides."""
    import dataclasses
    from voxphoton_amd import configs, vie
    bases = [configs.straight_waveguide(L=24, M=14, N=10, dx=40e-9),
             configs.straight_waveguide(L=30, M=14, N=10, dx=40e-9),
             configs.bragg_grating(periods=3, dx=40e-9, len_abs=0.4e-6, M=16, N=10)]
    reqs = []
    for base in bases:
        for lam in (1520e-9, 1580e-9):
            prob = dataclasses.replace(base, lam=lam)
            op = vie.NOperator.assembled(prob.shape, prob.dx, prob.ko, None, prob.Mc, prob.idx, ctx=vp_ctx)
            shape, bidx, sub = prob.box_problem()
            P = vie.Preconditioner.level2_assembled(shape, prob.dx, prob.ko, prob.core_contrast(), 1.0, bidx,
                                                    ctx=vp_ctx).set_embedding(prob.n, sub)
            for pol in ('y', 'z'):
                b = prob.gaussian_rhs(pol)
                db = vie.DeviceArray(vp_ctx, b.nbytes)
                db.upload(b)
                reqs.append((op, P, db, b.size))
    return reqs


def _solo(vp_ctx, reqs):
    # a served request keeps the recycled Krylov space of its first inner
    # solve between rounds, as vp_solve_ir keeps it between refinement steps;
    # requests that share a round carry zero columns for each other's
    # narrower spaces, which changes no result, so the solo solve is vp_solve_ir
    import ctypes
    from voxphoton_amd import _lib as L, vie
    out = []
    return _solo_runs(vp_ctx, reqs, L, vie, ctypes, out)


def _solo_runs(vp_ctx, reqs, L, vie, ctypes, out):
    for op, P, db, n in reqs:
        dx = vie.DeviceArray(vp_ctx, n * 16)
        info = L.GmresInfo()
        L.check(L.lib.vp_solve_ir(op.handle, P.handle, L.VP_PREC_RIGHT, db.ptr, dx.ptr, L.VP_MEM_DEVICE,
                                  PARAMS['tol'], PARAMS['restart'], PARAMS['inner_tol'], PARAMS['inner_maxit'],
                                  PARAMS['max_outer'], ctypes.byref(info)))
        x = np.empty(n, complex)
        dx.download(x)
        out.append((x, info.flag, info.relres))
    return out


@pytest.mark.parametrize('width', [-1, 4])
def test_server_results_equal_solo_solves(vp_ctx, monkeypatch, width):
    # width: the context's recycled-space width, which the server's rounds and
    # the solo solves both take (4: harmonic directions, narrower than a cycle)
    from voxphoton_amd import vie
    monkeypatch.delenv('VP_IR_RECYCLE_K', raising=False)
    vp_ctx.set_recycle_width(width)
    try:
        _server_equals_solo(vp_ctx, vie)
    finally:
        vp_ctx.set_recycle_width(-1)


def _server_equals_solo(vp_ctx, vie):
    reqs = _requests(vp_ctx)
    solo = _solo(vp_ctx, reqs)
    srv = vie.Server(hbm_budget_bytes=8 << 30, max_active=16, ctx=vp_ctx)
    try:
        xs = [vie.DeviceArray(vp_ctx, n * 16) for (_, _, _, n) in reqs]
        tickets = []
        # the first half starts alone; the second half arrives while it runs
        half = len(reqs) // 2
        for i in range(half):
            op, P, db, n = reqs[i]
            tickets.append(srv.submit(op, P, db, xs[i], **PARAMS))
        t0 = time.time()
        while srv.stats()['rounds'] < 2 and time.time() - t0 < 30:
            time.sleep(0.002)
        for i in range(half, len(reqs)):
            op, P, db, n = reqs[i]
            tickets.append(srv.submit(op, P, db, xs[i], **PARAMS))
        infos = [srv.wait(t) for t in tickets]
        st = srv.stats()
        print('server stats', st, flush=True)
        assert st['pending'] == 0 and st['active'] == 0 and st['bytes_in_use'] == 0
        for i, (info, (x_solo, flag, relres)) in enumerate(zip(infos, solo)):
            x = np.empty_like(x_solo)
            xs[i].download(x)
            print(i, info['relres'], relres, info['ir_steps'], rel(x, x_solo), flush=True)
            assert info['flag'] == 0 and flag == 0
            assert rel(x, x_solo) <= 1e-12          # same result alone and in the mix
            assert info['relres'] == relres
    finally:
        srv.close()


def test_server_budget_admits_one_at_a_time(vp_ctx):
    from voxphoton_amd import vie
    reqs = _requests(vp_ctx)[:4]
    solo = _solo(vp_ctx, reqs)
    srv = vie.Server(hbm_budget_bytes=1, max_active=16, ctx=vp_ctx)   # every request is oversize: one at a time
    try:
        xs = [vie.DeviceArray(vp_ctx, n * 16) for (_, _, _, n) in reqs]
        tickets = [srv.submit(op, P, db, xs[i], **PARAMS) for i, (op, P, db, n) in enumerate(reqs)]
        seen_active = 0
        while True:
            st = srv.stats()
            seen_active = max(seen_active, st['active'])
            if st['pending'] == 0 and st['active'] == 0:
                break
            time.sleep(0.001)
        assert seen_active <= 1
        for i, t in enumerate(tickets):
            info = srv.wait(t)
            x = np.empty(reqs[i][3], complex)
            xs[i].download(x)
            assert info['flag'] == 0 and rel(x, solo[i][0]) <= 1e-12
    finally:
        srv.close()


def test_server_rejects_mismatched_preconditioner(vp_ctx):
    from voxphoton_amd import vie
    from voxphoton_amd._lib import VoxPhotonError
    reqs = _requests(vp_ctx)
    op_a = reqs[0][0]
    P_b = reqs[4][1]          # preconditioner of another geometry
    srv = vie.Server(hbm_budget_bytes=1 << 30, ctx=vp_ctx)
    try:
        x = vie.DeviceArray(vp_ctx, reqs[0][3] * 16)
        with pytest.raises(VoxPhotonError):
            srv.submit(op_a, P_b, reqs[0][2], x, **PARAMS)
    finally:
        srv.close()


def _notched(vp_ctx, x_notch):
    """A waveguide with one cross-section column of the core removed at
    x_notch: equal grids, boxes and unknown counts, different voxel maps."""
    import dataclasses
    from voxphoton_amd import configs, vie
    base = configs.straight_waveguide(L=24, M=14, N=10, dx=40e-9)
    mask = base.mask.copy()
    ys, zs = np.nonzero(mask[x_notch])
    mask[x_notch, ys[0], zs[0]] = False
    prob = dataclasses.replace(base, mask=mask)
    op = vie.NOperator.assembled(prob.shape, prob.dx, prob.ko, None, prob.Mc, prob.idx, ctx=vp_ctx)
    shape, bidx, sub = prob.box_problem()
    P = vie.Preconditioner.level2_assembled(shape, prob.dx, prob.ko, prob.core_contrast(), 1.0, bidx,
                                            ctx=vp_ctx).set_embedding(prob.n, sub)
    b = prob.gaussian_rhs('y')
    db = vie.DeviceArray(vp_ctx, b.nbytes)
    db.upload(b)
    return op, P, db, b.size, prob


def test_equal_size_different_masks_are_never_batched_together(vp_ctx):
    # two geometries with the same grid, box and number of unknowns but
    # different voxel maps: the batched solve refuses to share one map, and
    # the server keeps them in separate groups, so each equals its solo solve
    from voxphoton_amd import vie
    from voxphoton_amd._lib import VoxPhotonError
    a = _notched(vp_ctx, 5)
    b = _notched(vp_ctx, 17)
    assert a[3] == b[3] and a[4].box == b[4].box and not np.array_equal(a[4].mask, b[4].mask)
    B = np.array([a[4].gaussian_rhs('y'), b[4].gaussian_rhs('y')])
    with pytest.raises(VoxPhotonError):
        vie.solve_ir_batched([a[0], b[0]], [a[1], b[1]], B, tol=PARAMS['tol'], restart=PARAMS['restart'],
                             inner_tol=PARAMS['inner_tol'], inner_maxit=PARAMS['inner_maxit'],
                             max_outer=PARAMS['max_outer'], side='right')
    reqs = [a[:4], b[:4]]
    solo = _solo(vp_ctx, reqs)
    srv = vie.Server(hbm_budget_bytes=8 << 30, max_active=16, ctx=vp_ctx)
    try:
        xs = [vie.DeviceArray(vp_ctx, n * 16) for (_, _, _, n) in reqs]
        tickets = [srv.submit(op, P, db, xs[i], **PARAMS) for i, (op, P, db, n) in enumerate(reqs)]
        for i, t in enumerate(tickets):
            info = srv.wait(t)
            x = np.empty(reqs[i][3], complex)
            xs[i].download(x)
            print('notched', i, info['relres'], solo[i][2], rel(x, solo[i][0]), flush=True)
            assert info['flag'] == 0 and rel(x, solo[i][0]) <= 1e-12 and info['relres'] == solo[i][2]
    finally:
        srv.close()


def test_concurrent_groups_retire_only_their_own_requests(vp_ctx):
    # six geometries (six groups) at once, so the executors run rounds side by
    # side; the client collects each request as soon as it is done and drops
    # its vectors at once, as a serving client does.  A request is reported
    # done only after its own round has finished with it: every result equals
    # its solo solve, in both waves of submissions.
    from voxphoton_amd import vie
    geos = [_notched(vp_ctx, x) for x in (3, 7, 11, 15, 19, 21)]
    reqs = [g[:4] for g in geos]
    solo = _solo(vp_ctx, reqs)
    srv = vie.Server(hbm_budget_bytes=8 << 30, max_active=16, ctx=vp_ctx)
    try:
        for wave in range(2):
            live = {}
            for i, (op, P, db, n) in enumerate(reqs):
                x = vie.DeviceArray(vp_ctx, n * 16)
                live[srv.submit(op, P, db, x, **PARAMS)] = (i, x)
            t0 = time.time()
            while live and time.time() - t0 < 120:
                for t in list(live):
                    done, _ = srv.poll(t)
                    if not done:
                        continue
                    info = srv.wait(t)
                    i, x = live.pop(t)
                    h = np.empty(reqs[i][3], complex)
                    x.download(h)
                    del x                      # the client drops the vector at once
                    assert info['flag'] == 0 and info['relres'] == solo[i][2], (wave, i)
                    assert rel(h, solo[i][0]) <= 1e-12, (wave, i)
                time.sleep(0.001)
            assert not live
            st = srv.stats()
            assert st['active'] == 0 and st['bytes_in_use'] == 0
    finally:
        srv.close()
