"""GPU: model-order-reduction primitives and the reduced model through the C
ABI against numpy and the MOR oracle."""
import numpy as np
import pytest

from oracle import mor as omor, vie as ovie, greens as ogr
from tests.helpers import random_problem, random_toeplitz, rel

pytestmark = pytest.mark.gpu


def test_vector_products(vp_ctx):
    from voxphoton_amd.vie import DeviceArray
    from voxphoton_amd.mor import vec_dots, vec_lincomb
    rng = np.random.default_rng(1)
    n = 10007
    X = [rng.standard_normal(n) + 1j * rng.standard_normal(n) for _ in range(5)]
    Y = [rng.standard_normal(n) + 1j * rng.standard_normal(n) for _ in range(11)]
    dX, dY = [], []
    for a, d in ((X, dX), (Y, dY)):
        for v in a:
            arr = DeviceArray(vp_ctx, n * 16)
            arr.upload(v)
            d.append(arr)
    G = vec_dots(vp_ctx, dX, dY, n)
    assert rel(G, np.array(X).conj() @ np.array(Y).T) < 1e-13
    C = rng.standard_normal((5, 3)) + 1j * rng.standard_normal((5, 3))
    Z = vec_lincomb(vp_ctx, dX, C, n)
    want = np.array(X).T @ C
    for j in range(3):
        assert rel(Z[j].download(np.empty(n, complex)), want[:, j]) < 1e-14


def test_operator_lincomb_is_linear(vp_ctx):
    from voxphoton_amd.vie import NOperator
    from voxphoton_amd.mor import op_lincomb
    rng = np.random.default_rng(2)
    G0, Mr, Mc, Gram, idx, x = random_problem(8, 6, 4, seed=2)
    Gs = [G0] + [random_toeplitz(8, 6, 4, rng) for _ in range(2)]
    ops = [NOperator.from_toeplitz(G, None, Mc, Gram, idx, ctx=vp_ctx) for G in Gs]
    w = np.array([0.3 - 0.2j, -1.1, 0.7 + 0.5j])
    op = op_lincomb(ops, w)
    # A_w x = Gram x - Mc N_w x with N_w = sum w_i N_i
    N = sum(wi * (Gram * x - o.apply(x)) for wi, o in zip(w, ops))      # sum w_i Mc N_i x
    want = Gram * x - N
    assert rel(op.apply(x), want) < 1e-12
    assert rel(op.apply(x, prec='fp32'), want) < 1e-5


def _bragg_family(L, M, N, s_um, seed=0):
    # silicon-in-oxide voxels at wavelengths s_um (micrometres), 20 nm voxels
    dx = 20e-9
    rng = np.random.default_rng(seed)
    idx = ovie.idx_from_mask(np.ones((L, M, N), bool))
    er = 12.11 / 2.085
    Mc = np.full((L, M, N), (er - 1) / er + 0j)
    Gs = [ogr.assemble_G(L, M, N, dx, 2 * np.pi * np.sqrt(2.085) / (t * 1e-6)) for t in s_um]
    b = (rng.standard_normal(idx.size) + 1j * rng.standard_normal(idx.size)) * dx ** 3
    return Gs, Mc, dx ** 3, idx, b


def test_reduced_model_matches_oracle(vp_ctx):
    from voxphoton_amd.vie import NOperator, Preconditioner
    from voxphoton_amd.mor import ReducedModel
    L, M, N = 10, 5, 3
    s = np.array([1.50, 1.55, 1.60])
    Gs, Mc, Gram, idx, b = _bragg_family(L, M, N, s, seed=4)
    ops = [NOperator.from_toeplitz(G, None, Mc, Gram, idx, ctx=vp_ctx) for G in Gs]
    precs = [Preconditioner.level2(G, Mc.flat[0], 1.0, Gram, idx, ctx=vp_ctx) for G in Gs]
    rm = ReducedModel(ops, s, [b] * 3, order=3, precs=precs, tol=1e-12, restart=40, inner_tol=1e-4,
                      max_outer=30)
    print('MOR solves', rm.solve_stats, 'K', rm.K, flush=True)
    assert rm.K == 9 and all(f == 0 for f, _, _ in rm.solve_stats)
    ones = np.ones(Mc.shape, complex)
    A = [ovie.dense_AN(G, ones, Mc, Gram, idx) for G in Gs]
    E = [sum(w * Ai for w, Ai in zip(omor.lagrange_deriv_weights(t, s), A)) for t in s]
    U = omor.mor_basis(A, E, [b] * 3, order=3)
    Q = omor.mor_project(A, U)
    # the basis spans the same space: reduced solutions agree on and off the samples
    for S in (1.50, 1.527, 1.58):
        xd = rm.solve(S, b)[0].download(np.empty(idx.size, complex))
        xo = omor.mor_solve(Q, s, S, U, b)
        print('reduced solution at', S, 'rel', rel(xd, xo), flush=True)
        assert rel(xd, xo) < 1e-8
    # exact at a sample (the solution is a basis column)
    xd = rm.solve(1.55, b)[0].download(np.empty(idx.size, complex))
    assert rel(xd, np.linalg.solve(A[1], b)) < 1e-10
    # the projected matrices are U^H A_i U of the device basis
    Ud = np.column_stack([u.download(np.empty(idx.size, complex)) for u in rm.U])
    assert np.abs(Ud.conj().T @ Ud - np.eye(rm.K)).max() < 1e-12
    for Ai, Qi in zip(A, rm.Q):
        assert rel(Qi, Ud.conj().T @ Ai @ Ud) < 1e-12
