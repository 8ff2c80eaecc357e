"""BiCGSTAB and the one-call harness (JVIE_Solver_const) through the C ABI,
against the FP64 oracle."""
import numpy as np
import pytest

from oracle import vie as ovie, krylov as okry
from tests.helpers import rel

pytestmark = pytest.mark.gpu


def _waveguide(vp_ctx):
    from voxphoton_amd import configs, vie
    prob = configs.straight_waveguide(L=24, M=14, N=10, dx=40e-9)
    op = vie.NOperator.assembled(prob.shape, prob.dx, prob.ko, None, prob.Mc, prob.idx, ctx=vp_ctx)
    G = vie.assemble_toeplitz(prob.shape, prob.dx, prob.ko, ctx=vp_ctx)
    fN = ovie.fN_from_toeplitz(G)
    ones = np.ones(prob.shape, complex)
    A = lambda v: ovie.mv_AN_const(v, fN, ones, prob.Mc, prob.gram, 'notransp', prob.idx)   # noqa: E731
    return prob, op, A


def test_bicgstab_matches_oracle(vp_ctx):
    from voxphoton_amd import vie
    prob, op, A = _waveguide(vp_ctx)
    b = prob.gaussian_rhs('y')
    x, flag, relres, it, rv = vie.bicgstab(op, b, 1e-8, 300)
    xo, flago, relreso, ito, rvo = okry.bicgstab(A, b, 1e-8, 300)
    true = np.linalg.norm(b - A(x)) / np.linalg.norm(b)
    print(f'bicgstab: device {it} its relres {relres:.2e} (oracle residual {true:.2e}); oracle {ito} its '
          f'relres {relreso:.2e}; history rel diff (first 12) {np.abs(rv[:12] - rvo[:12]).max():.1e}', flush=True)
    assert flag == 0 and flago == 0
    assert true <= 1e-8 * (1 + 1e-6)
    assert abs(it - ito) <= 1.0
    k = min(len(rv), len(rvo), 12)
    assert np.abs(rv[:k] - rvo[:k]).max() <= 1e-6 * rv[0]
    assert rel(x, xo) < 1e-6


def test_bicgstab_preconditioned_and_fp32(vp_ctx):
    from voxphoton_amd import configs, vie
    prob, op, A = _waveguide(vp_ctx)
    shape, bidx, sub = prob.box_problem()
    P = vie.Preconditioner.level2_assembled(shape, prob.dx, prob.ko, prob.core_contrast(), 1.0, bidx,
                                            ctx=vp_ctx).set_embedding(prob.n, sub)
    b = prob.gaussian_rhs('z')
    x, flag, relres, it, rv = vie.bicgstab(op, b, 1e-8, 300, M=P)
    # the oracle's iteration with the same preconditioner callback
    xo, flago, relreso, ito, rvo = okry.bicgstab(A, b, 1e-8, 300, M=P.apply)
    x0, flag0, _, it0, _ = vie.bicgstab(op, b, 1e-8, 300)
    true = np.linalg.norm(b - A(x)) / np.linalg.norm(b)
    print(f'preconditioned bicgstab: {it} its (oracle {ito}, unpreconditioned {it0}), oracle residual {true:.2e}',
          flush=True)
    assert flag == 0 and flago == 0 and true <= 1e-8 * (1 + 1e-6)
    assert abs(it - ito) <= 1.0 and it < it0
    # FP32 arithmetic: converges to a single-precision tolerance
    x32, flag32, relres32, it32, _ = vie.bicgstab(op, b, 1e-4, 300, M=P, prec='fp32')
    t32 = np.linalg.norm(b - A(x32)) / np.linalg.norm(b)
    print(f'fp32 bicgstab: {it32} its, relres {relres32:.2e}, oracle residual {t32:.2e}', flush=True)
    assert flag32 == 0 and t32 <= 2e-4


def test_bicgstab_nan_rhs_fails_cleanly(vp_ctx):
    from voxphoton_amd import vie
    prob, op, A = _waveguide(vp_ctx)
    b = prob.gaussian_rhs('y')
    b[5] = np.nan
    x, flag, relres, it, rv = vie.bicgstab(op, b, 1e-8, 50)
    assert flag != 0 and not np.isfinite(relres)


@pytest.mark.parametrize('solver', ['B', 'G', 'D'])
def test_jvie_solver_const_harness(vp_ctx, solver):
    # the harness on a small grid: reference-style inputs (r grid, relative
    # permittivity, incident field, frequency and fN), JVIE-II formulation
    from voxphoton_amd import configs, vie
    prob = configs.straight_waveguide(L=16, M=12, N=8, dx=40e-9)
    L, M, N = prob.shape
    xs, ys, zs = (np.arange(v) * prob.dx for v in (L, M, N))
    r = np.stack(np.meshgrid(xs, ys, zs, indexing='ij'), axis=-1)
    e_r = np.where(prob.mask, prob.eps / configs.EPS_SIO2, 1.0 + 0j)
    Einc = prob.gaussian_field('y')
    G = vie.assemble_toeplitz(prob.shape, prob.dx, prob.ko, ctx=vp_ctx)
    fN = ovie.fN_from_toeplitz(G)
    Je, flag, relres, it, resvec, solvert, vsol = vie.JVIE_Solver_const(Einc, e_r, r, prob.freq, fN, tol=1e-8,
                                                                         solver=solver, inner_it=20, outer_it=40,
                                                                         max_it=400, ritz=5, ctx=vp_ctx)
    b = ovie.jvie2_rhs(Einc, prob.Mc, prob.gram, prob.freq, prob.idx)
    A = lambda v: ovie.mv_AN_const(v, fN, np.ones(prob.shape, complex), prob.Mc, prob.gram, 'notransp',   # noqa
                                   prob.idx)
    true = np.linalg.norm(b - A(vsol)) / np.linalg.norm(b)
    print(f'harness {solver}: flag {flag}, relres {relres:.2e}, oracle residual {true:.2e}, iter {it}', flush=True)
    assert flag == 0 and true <= 1e-8 * (1 + 1e-6)
    # Je carries the solution at idxS3 and zero elsewhere
    flat = Je.reshape(-1, order='F')
    assert np.array_equal(flat[prob.idx - 1], vsol)
    mask = np.ones(flat.size, bool)
    mask[prob.idx - 1] = False
    assert not np.any(flat[mask])
