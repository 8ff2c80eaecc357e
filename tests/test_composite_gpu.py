"""Block-Jacobi composite preconditioner (prec_fun_disk_idx.m:1-23) through
the C ABI: the apply against the oracle's composite, and resonator solves
(disk plus bus waveguide) preconditioned by it, against the oracle's residual."""
import numpy as np
import pytest

from oracle import vie as ovie, precond as opre
from tests.helpers import rel

pytestmark = pytest.mark.gpu


def _disk_bus_sets(rng):
    """Outer problem of a bus box (all voxels) and a disk box (some voxels)."""
    L, M, N = 6, 7, 2
    mask = np.zeros((L, M, N), bool)
    mask[:, 0:2, :] = True
    mask[1:5, 3:7, :] = rng.random((4, 4, N)) < 0.8
    idxS3 = ovie.idx_from_mask(mask)
    V = L * M * N
    full_vox = np.flatnonzero(mask.reshape(-1, order='F'))
    nS = full_vox.size

    def embed(box_vox_field):
        # field index (c*V + voxel) -> position in the outer unknown vector
        c, v = np.divmod(box_vox_field, V)
        pos = np.searchsorted(full_vox, v)
        assert np.array_equal(full_vox[pos], v)
        return c * nS + pos + 1

    Lw, Mw, Nw = L, 2, N
    bl, bm, bn = np.meshgrid(np.arange(Lw), np.arange(Mw), np.arange(Nw), indexing='ij')
    bvox = (bl + L * (bm + M * bn)).reshape(-1, order='F')
    idx_T = np.concatenate([bvox + c * V for c in range(3)]) + 1
    Ld, Md, Nd = 4, 4, N
    dmask = mask[1:5, 3:7, :]
    idx_Sq = ovie.idx_from_mask(dmask)
    dl, dm, dn = np.meshgrid(np.arange(Ld), np.arange(Md), np.arange(Nd), indexing='ij')
    dvox = ((dl + 1) + L * ((dm + 3) + M * dn)).reshape(-1, order='F')
    dfield = np.concatenate([dvox + c * V for c in range(3)])
    idx_R = dfield[idx_Sq - 1] + 1
    return dict(L=L, M=M, N=N, idxS3=idxS3, idx_T=idx_T, idx_R=idx_R, idx_Sq=idx_Sq, box_w=(Lw, Mw, Nw),
                box_d=(Ld, Md, Nd), sub_T=embed(idx_T - 1), sub_R=embed(idx_R - 1))


def test_composite_apply_matches_oracle(vp_ctx):
    from voxphoton_amd.vie import Preconditioner
    rng = np.random.default_rng(23)
    d = _disk_bus_sets(rng)
    Lw, Mw, Nw = d['box_w']
    Ld, Md, Nd = d['box_d']
    nw = 3 * Mw * Nw
    Bw = (rng.standard_normal((Lw, nw, nw)) + 1j * rng.standard_normal((Lw, nw, nw))) / nw
    Bd = (rng.standard_normal((Ld, Md, 3 * Nd, 3 * Nd)) + 1j * rng.standard_normal((Ld, Md, 3 * Nd, 3 * Nd))) / Nd
    n = d['idxS3'].size
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    ref = opre.prec_fun_disk_idx(x, Bw, d['idx_T'], d['idx_R'], Lw, Mw, Nw, Bd, Ld, Md, Nd, d['L'], d['M'], d['N'],
                                 d['idxS3'], d['idx_Sq'])
    # the bus box is whole: its unknowns are all its voxels
    pw = Preconditioner.chan1(Bw, Lw, Mw, Nw, ovie.idx_from_mask(np.ones((Lw, Mw, Nw), bool)), ctx=vp_ctx)
    pw.set_embedding(n, d['sub_T'])
    pd = Preconditioner.chan2(Bd, Ld, Md, Nd, d['idx_Sq'], ctx=vp_ctx)
    pd.set_embedding(n, d['sub_R'])
    P = Preconditioner.composite([pw, pd], ctx=vp_ctx)
    e64 = rel(P.apply(x), ref)
    e32 = rel(P.apply(x, 'fp32'), ref)
    print('composite apply vs oracle', e64, e32, flush=True)
    assert e64 < 1e-12 and e32 < 1e-5
    # order of the parts does not matter (disjoint unknowns)
    P2 = Preconditioner.composite([pd, pw], ctx=vp_ctx)
    assert rel(P2.apply(x), ref) < 1e-12


def test_composite_refuses_overlapping_parts(vp_ctx):
    from voxphoton_amd.vie import Preconditioner
    from voxphoton_amd._lib import VoxPhotonError
    rng = np.random.default_rng(5)
    d = _disk_bus_sets(rng)
    Ld, Md, Nd = d['box_d']
    Bd = (rng.standard_normal((Ld, Md, 3 * Nd, 3 * Nd)) + 1j * rng.standard_normal((Ld, Md, 3 * Nd, 3 * Nd)))
    n = d['idxS3'].size
    pa = Preconditioner.chan2(Bd, Ld, Md, Nd, d['idx_Sq'], ctx=vp_ctx).set_embedding(n, d['sub_R'])
    pb = Preconditioner.chan2(Bd, Ld, Md, Nd, d['idx_Sq'], ctx=vp_ctx).set_embedding(n, d['sub_R'])
    with pytest.raises(VoxPhotonError):
        Preconditioner.composite([pa, pb], ctx=vp_ctx)
    no_embed = Preconditioner.chan2(Bd, Ld, Md, Nd, d['idx_Sq'], ctx=vp_ctx)
    with pytest.raises(VoxPhotonError):
        Preconditioner.composite([pa, no_embed], ctx=vp_ctx)


def _resonator(vp_ctx, lam=1550e-9):
    import dataclasses
    from voxphoton_amd import configs, vie
    import bench
    prob = configs.microdisk_bus(radius=0.3e-6, dx=20e-9, L=40, M=80, N=16)
    prob = dataclasses.replace(prob, lam=lam)
    op = vie.NOperator.assembled(prob.shape, prob.dx, prob.ko, None, prob.Mc, prob.idx, ctx=vp_ctx)
    P, desc = bench.make_preconditioner(prob, dict(prec='composite', damping=0.0), vie, configs)
    return prob, op, P, desc


def test_resonator_solve_with_composite_matches_oracle(vp_ctx):
    from voxphoton_amd import vie
    prob, op, P, desc = _resonator(vp_ctx)
    b = prob.gaussian_rhs('y')
    x, info = vie.solve_ir(op, b, tol=1e-8, restart=40, inner_tol=1e-3, inner_maxit=6, max_outer=30, M=P,
                           side='right')
    G = vie.assemble_toeplitz(prob.shape, prob.dx, prob.ko, ctx=vp_ctx)
    r = b - ovie.mv_AN_const(x, ovie.fN_from_toeplitz(G), np.ones(prob.shape, complex), prob.Mc, prob.gram,
                             'notransp', prob.idx)
    true = np.linalg.norm(r) / np.linalg.norm(b)
    # the composite is a preconditioner: GMRES with it converges faster than without
    x0, info0 = vie.solve_ir(op, b, tol=1e-8, restart=40, inner_tol=1e-3, inner_maxit=6, max_outer=30, M=None,
                             side='right')
    print(f'resonator {prob.shape} n={prob.n}, {desc}: {info["ir_steps"]} steps / {info["inner_iters"]} its '
          f'(unpreconditioned {info0["inner_iters"]}), relres {info["relres"]:.2e}, oracle {true:.2e}', flush=True)
    assert info['flag'] == 0 and true <= 1e-8 * (1 + 1e-6)
    assert info['inner_iters'] < info0['inner_iters']


def test_resonator_batched_with_composite_equals_solo(vp_ctx):
    from voxphoton_amd import vie
    prob, op, P, _ = _resonator(vp_ctx)
    B = np.array([prob.gaussian_rhs('y'), prob.gaussian_rhs('z')])
    prm = dict(tol=1e-8, restart=40, inner_tol=1e-3, inner_maxit=6, max_outer=30, side='right')
    X, infos = vie.solve_ir_batched([op, op], [P, P], B, **prm)
    for i in range(2):
        Xs, infos_s = vie.solve_ir_batched([op], [P], B[i:i + 1], **prm)
        print('batched vs solo', i, infos[i]['relres'], infos_s[0]['relres'], rel(X[i], Xs[0]), flush=True)
        assert infos[i]['flag'] == 0
        assert rel(X[i], Xs[0]) <= 1e-12
