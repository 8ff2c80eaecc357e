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


def test_gather_scatter_absmax(vp_ctx):
    from voxphoton_amd._lib import VoxPhotonError
    from voxphoton_amd.vie import DeviceArray
    from voxphoton_amd.mor import vec_absmax, vec_gather, vec_scatter
    rng = np.random.default_rng(8)
    nx, n = 5003, 1777
    X = rng.standard_normal(nx) + 1j * rng.standard_normal(nx)
    idx = np.sort(rng.choice(nx, n, replace=False)).astype(np.int64)
    w = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    dev = lambda a: (lambda d: (d.upload(np.ascontiguousarray(a)), d)[1])(DeviceArray(vp_ctx, a.nbytes))  # noqa
    dX, didx, dw, dY = dev(X), dev(idx), dev(w), DeviceArray(vp_ctx, n * 16)
    vec_gather(vp_ctx, dX, nx, didx, n, dY, w=dw, alpha=0.3 - 2j)
    # complex products round differently under fused multiply-add
    assert rel(dY.download(np.empty(n, complex)), (0.3 - 2j) * w * X[idx]) < 1e-15
    vec_gather(vp_ctx, dX, nx, didx, n, dY)
    assert np.array_equal(dY.download(np.empty(n, complex)), X[idx])
    dZ = DeviceArray(vp_ctx, nx * 16)
    dZ.upload(np.full(nx, 7 + 7j))
    vec_scatter(vp_ctx, dY, didx, n, dZ, nx)
    want = np.zeros(nx, complex)
    want[idx] = X[idx]
    assert np.array_equal(dZ.download(np.empty(nx, complex)), want)
    # per-vector maxima, also through offset pointers into one buffer
    m = vec_absmax(vp_ctx, [dX, dX.ptr.value + 16 * 3000], 2000)
    assert m[0] == np.abs(X[:2000]).max() and m[1] == np.abs(X[3000:5000]).max()
    # an index outside the source is refused after the kernel skipped it
    bad = dev(np.array([0, nx], dtype=np.int64))
    with pytest.raises(VoxPhotonError):
        vec_gather(vp_ctx, dX, nx, bad, 2, dY)
    with pytest.raises(VoxPhotonError):
        vec_scatter(vp_ctx, dY, bad, 2, dZ, nx)


def test_lincomb_into_equals_new_operator(vp_ctx):
    from voxphoton_amd.vie import NOperator
    from voxphoton_amd.mor import op_lincomb, op_lincomb_into
    rng = np.random.default_rng(9)
    shape = (7, 5, 3)
    ops = [NOperator.for_fields(shape, Gram=0.7, G=random_toeplitz(*shape, rng), precisions=('fp64', 'fp32'),
                                ctx=vp_ctx) for _ in range(3)]
    w = [0.2 + 0.1j, -0.4, 1.3j]
    a = op_lincomb(ops, w)
    b = op_lincomb(ops, [1, 0, 0])
    op_lincomb_into(b, ops, w)
    x = rng.standard_normal(ops[0].n) + 1j * rng.standard_normal(ops[0].n)
    for prec in ('fp64', 'fp32'):
        assert np.array_equal(a.apply(x, prec=prec), b.apply(x, prec=prec))
    from voxphoton_amd._lib import VoxPhotonError
    with pytest.raises(VoxPhotonError):
        op_lincomb_into(ops[0], ops, w)          # the target may not be a source


def test_spectrum_point_matches_oracle(vp_ctx):
    from oracle import fields as ofl
    from oracle.vie import C0, em_constants, jvie2_rhs
    from voxphoton_amd.vie import NOperator, Preconditioner
    from voxphoton_amd.mor import ReducedModel, SpectrumSweep
    L, M, N = 12, 6, 4
    dx = 20e-9
    n_ext = np.sqrt(2.085)
    s = np.array([1.50e-6, 1.55e-6, 1.60e-6])
    mask = np.zeros((L, M, N), bool)
    mask[:, 1:5, 1:3] = True
    mask[4:8, 0, 1:3] = True                      # a corrugation
    idx = ovie.idx_from_mask(mask)
    er = 12.11 / 2.085
    Mc = np.where(mask, (er - 1) / er, 0).astype(complex)
    Gram = dx ** 3
    Gs = [ogr.assemble_G(L, M, N, dx, 2 * np.pi * n_ext / t) for t in s]
    ops = [NOperator.from_toeplitz(G, None, Mc, Gram, idx, ctx=vp_ctx) for G in Gs]
    precs = [Preconditioner.level2(G, (er - 1) / er, 1.0, Gram, idx, ctx=vp_ctx) for G in Gs]
    fops = [NOperator.for_fields((L, M, N), Gram=Gram, G=G, ctx=vp_ctx) for G in Gs]
    lam_t = 440e-9
    origin = (0.4e-6, -2.5 * dx, -1.5 * dx)
    src = np.array([0.25 * lam_t + 0.3e-6 - 1j * lam_t, 0, 0])
    # the sample right-hand sides as the sweep forms them (example:89-95)
    fN = [ovie.fN_from_toeplitz(G) for G in Gs]
    A = [ovie.dense_AN(G, np.ones(Mc.shape, complex), Mc, Gram, idx) for G in Gs]
    rhs = []
    for t in s:
        freq = C0 * n_ext / t
        Ei = ofl.E_field_DGF_mollified(np.array([0, 1, 0], complex), src[None], ofl.grid_coords(origin, dx, L, M, N),
                                       em_constants(freq)[1], 0.5 * lam_t).reshape(L, M, N, 3, order='F')
        assert np.count_nonzero(Ei[..., 1]) > 0.3 * Ei[..., 1].size
        rhs.append(jvie2_rhs(Ei / np.abs(Ei).max(axis=(0, 1, 2)), Mc, Gram, freq, idx))
    rm = ReducedModel(ops, s, rhs, order=3, precs=precs, tol=1e-12, restart=40, inner_tol=1e-4, max_outer=30)
    assert rm.K == 9 and all(f == 0 for f, _, _ in rm.solve_stats)
    E = [sum(w * Ai for w, Ai in zip(omor.lagrange_deriv_weights(t, s), A)) for t in s]
    U = omor.mor_basis(A, E, rhs, order=3)
    Q = omor.mor_project(A, U)
    sw = SpectrumSweep(s, fops, Mc, idx, Gram, origin, dx, [0, 1, 0], src, 0.5 * lam_t, x_thru=10, m_edge=1,
                       n_ext=n_ext)
    # the device right-hand side at a sample is the host one
    assert rel(sw.rhs_copy(s[1]).download(np.empty(idx.size, complex)), rhs[1]) < 1e-13
    for S in (1.55e-6, 1.521e-6, 1.597e-6):
        p = sw.point(S, rm)
        po, xo, Eo = omor.spectrum_point(Q, s, U, S, fN, Mc, idx, Gram, origin, dx, [0, 1, 0], src, 0.5 * lam_t,
                                         10, 1, n_ext)
        xd = sw.d_x.download(np.empty(idx.size, complex))
        Ed = sw.d_E.download(np.empty(3 * L * M * N, complex)).reshape(L, M, N, 3, order='F')
        print('spectrum point', S, 'power', p, po, 'x rel', rel(xd, xo), 'E rel', rel(Ed, Eo), flush=True)
        assert rel(xd, xo) < 1e-8 and rel(Ed, Eo) < 1e-8
        assert abs(p - po) <= 1e-8 * po
