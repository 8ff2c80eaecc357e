"""CPU tests pinning the FP64 oracle against closed forms and brute force."""
import numpy as np
import pytest

from oracle import vie as ovie, precond as opre, krylov as okry, greens as ogr
from tests.helpers import random_problem, random_toeplitz, rel


@pytest.mark.parametrize('shape', [(5, 4, 3), (6, 5, 4), (3, 1, 2), (1, 1, 1), (4, 3, 1)])
def test_mv_AN_const_matches_dense_toeplitz(shape):
    G, Mr, Mc, Gram, idx, x = random_problem(*shape, seed=sum(shape), with_mr=True)
    fN = ovie.fN_from_toeplitz(G)
    y = ovie.mv_AN_const(x, fN, Mr, Mc, Gram, 'notransp', idx)
    A = ovie.dense_AN(G, Mr, Mc, Gram, idx)
    assert rel(y, A @ x) < 1e-13


def test_mv_AN_const_transp_is_adjoint():
    G, Mr, Mc, Gram, idx, x = random_problem(5, 4, 3, seed=3, with_mr=True)
    fN = ovie.fN_from_toeplitz(G)
    A = ovie.dense_AN(G, Mr, Mc, Gram, idx)
    yt = ovie.mv_AN_const(x, fN, Mr, Mc, Gram, 'transp', idx)
    assert rel(yt, A.conj().T @ x) < 1e-13


def test_zero_contrast_is_gram_identity():
    G, Mr, Mc, Gram, idx, x = random_problem(4, 3, 2, seed=5, with_mr=True)
    y = ovie.mv_AN_const(x, ovie.fN_from_toeplitz(G), Mr, 0 * Mc, Gram, 'notransp', idx)
    mr = np.tile(Mr.reshape(-1, order='F'), 3)[idx - 1]
    assert rel(y, Gram * mr * x) < 1e-15


def test_N_is_complex_symmetric():
    rng = np.random.default_rng(7)
    G = random_toeplitz(4, 3, 3, rng)
    D = ovie.dense_N(G)
    assert np.abs(D - D.T).max() < 1e-14 * np.abs(D).max()


def test_circulant_signs_match_table():
    # the parity table reproduces gperiodic_coeff_nop for every reflected cube
    for cube, axes in [('L', [0]), ('M', [1]), ('N', [2]), ('LM', [0, 1]), ('LN', [0, 2]),
                       ('MN', [1, 2]), ('LMN', [0, 1, 2])]:
        expect = np.prod(ovie.PARITY[:, axes], axis=1)
        assert np.array_equal(expect, ovie.gperiodic_coeff_nop(cube))


def test_chan_apply_matches_dense_fourier_form():
    rng = np.random.default_rng(11)
    L, M, N = 4, 3, 2
    B = rng.standard_normal((L, M, 3 * N, 3 * N)) + 1j * rng.standard_normal((L, M, 3 * N, 3 * N))
    mask = rng.random((L, M, N)) < 0.7
    idx = ovie.idx_from_mask(mask)
    x = rng.standard_normal(idx.size) + 1j * rng.standard_normal(idx.size)
    y = opre.chan_2_mvp_for_parallel_idx(B, x, L, M, N, idx)
    # dense: F^-1 blkdiag F on the 3LMN field
    V = L * M * N
    n3 = 3 * V
    Fx = np.fft.fft(np.eye(L), axis=0)
    Fy = np.fft.fft(np.eye(M), axis=0)
    # field index l + L*(m + M*(n + N*c))
    F = np.kron(np.eye(3 * N), np.kron(Fy, Fx))
    D = np.zeros((n3, n3), dtype=complex)
    for i in range(L):
        for j in range(M):
            for a in range(3 * N):
                for b in range(3 * N):
                    D[i + L * (j + M * a), i + L * (j + M * b)] = B[i, j, a, b]
    A = np.linalg.inv(F) @ D @ F
    i0 = idx - 1
    full = np.zeros(n3, complex)
    full[i0] = x
    assert rel(y, (A @ full)[i0]) < 1e-13


def test_chan_exact_for_circulant_operator():
    # For blocks equal to the exact inverse of a block-circulant operator, the
    # preconditioned operator is the identity.
    rng = np.random.default_rng(2)
    L, M, N = 4, 3, 2
    Bf = rng.standard_normal((L, M, 3 * N, 3 * N)) + 1j * rng.standard_normal((L, M, 3 * N, 3 * N)) \
        + 6 * np.eye(3 * N)
    Binv = np.linalg.inv(Bf)
    v = rng.standard_normal(3 * L * M * N) + 1j * rng.standard_normal(3 * L * M * N)
    w = opre.chan_2_mvp(Bf, opre.chan_2_mvp(Binv, v, L, M, N), L, M, N)
    assert rel(w, v) < 1e-12


def test_level2_direct_equals_literal():
    rng = np.random.default_rng(4)
    L, M, N = 4, 3, 3
    G = random_toeplitz(L, M, N, rng)
    a = opre.level_2_blocks_literal(G, 0.83 + 0.1j, 1.0, L, M, N, 0.7)
    b = opre.level_2_blocks_direct(G, 0.83 + 0.1j, 1.0, L, M, N, 0.7)
    assert np.abs(a - b).max() <= 1e-14 * np.abs(a).max()


def test_chan_is_frobenius_optimal():
    # Chan's circulant minimises ||C - T||_F over circulants (checked by least squares)
    rng = np.random.default_rng(9)
    L = 6
    t = rng.standard_normal(2 * L - 1) + 1j * rng.standard_normal(2 * L - 1)   # T(d), d = -(L-1)..L-1
    T = np.array([[t[(i - j) + L - 1] for j in range(L)] for i in range(L)])
    basis = [np.roll(np.eye(L), k, axis=0) for k in range(L)]
    Amat = np.stack([b.reshape(-1) for b in basis], axis=1)
    c_ls = np.linalg.lstsq(Amat, T.reshape(-1), rcond=None)[0]
    # oracle formula: c_i = ((L-i) T_i + i T_{i-L}) / L
    c = np.array([((L - i) * t[i + L - 1] + i * t[i - L + L - 1]) / L if i else t[L - 1] for i in range(L)])
    assert rel(c, c_ls) < 1e-13


def test_pgmres_matches_direct_solve():
    rng = np.random.default_rng(1)
    n = 30
    A = np.eye(n) * 12 + rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    for orth in ('cgs', 'cgs2'):
        x, flag, relres, it, rv = okry.pgmres(lambda v: A @ v, b, 10, 1e-10, 50, orth=orth)
        assert flag == 0
        assert rel(x, np.linalg.solve(A, b)) < 1e-8


def test_pgmres_terminates_in_n_steps_and_rank_one():
    rng = np.random.default_rng(8)
    n = 12
    A = np.eye(n) + np.outer(rng.standard_normal(n), rng.standard_normal(n)) * 0.1
    b = rng.standard_normal(n) + 0j
    x, flag, relres, it, rv = okry.pgmres(lambda v: A @ v, b, n, 1e-12, 1)
    assert flag == 0 and len(rv) - 1 <= 3      # I + rank-1 converges in <= 2 steps
    B = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)) + 5 * np.eye(n)
    x, flag, relres, it, rv = okry.pgmres(lambda v: B @ v, b, n, 1e-13, 1)
    assert len(rv) - 1 <= n and relres < 1e-12


def test_green_static_closed_form():
    I = ogr._parallel((0, 0, 0), 0, 0, 0, 0.0, 10)
    exact = 4 * np.log(1 + np.sqrt(2)) - 4 / 3 * (np.sqrt(2) - 1)
    assert abs(I - exact) < 1e-13


def test_green_singular_rules_converge():
    base = ogr.face_pair_integrals((1, 0, 1), 0.117)
    old = (ogr.ORDER_SINGULAR, ogr.ORDER_NEAR)
    try:
        ogr.ORDER_SINGULAR, ogr.ORDER_NEAR = 20, 16
        hi = ogr.face_pair_integrals((1, 0, 1), 0.117)
    finally:
        ogr.ORDER_SINGULAR, ogr.ORDER_NEAR = old
    assert np.abs(base - hi).max() < 1e-12 * np.abs(hi).max()


def test_green_far_matches_brute_force():
    m, kappa = (6, 2, 1), 0.2
    I = ogr.face_pair_integrals(m, kappa)
    t, w = ogr.gauss01(8)
    B = np.zeros((6, 6), complex)
    for p, (ap, sp) in enumerate(ogr.FACES):
        for q, (aq, sq) in enumerate(ogr.FACES):
            op_ = [a for a in range(3) if a != ap]
            oq = [a for a in range(3) if a != aq]
            U1, U2, U3, U4 = np.meshgrid(t, t, t, t, indexing='ij')
            W = w[:, None, None, None] * w[None, :, None, None] * w[None, None, :, None] * w[None, None, None, :]
            rp = [None] * 3
            rq = [None] * 3
            rp[ap] = m[ap] + sp
            rp[op_[0]] = m[op_[0]] + U1
            rp[op_[1]] = m[op_[1]] + U2
            rq[aq] = sq
            rq[oq[0]] = U3
            rq[oq[1]] = U4
            rho = np.sqrt(sum((rq[a] - rp[a]) ** 2 for a in range(3)))
            B[p, q] = np.sum(W * np.exp(-1j * kappa * rho) / rho)
    assert np.abs(I - B).max() < 1e-7 * np.abs(B).max()


def test_green_projection_and_symmetry():
    G = ogr.assemble_G(3, 3, 2, 20e-9, 5.85e6)
    # diagonal components equal at the self cell (cubic symmetry)
    assert abs(G[0, 0, 0, 0] - G[0, 0, 0, 3]) < 1e-12 * abs(G[0, 0, 0, 0])
    assert abs(G[0, 0, 0, 0] - G[0, 0, 0, 5]) < 1e-12 * abs(G[0, 0, 0, 0])
    # xx at offset (0,1,0) equals zz at offset (0,1,0) by symmetry of x<->z
    assert abs(G[0, 1, 0, 0] - G[0, 1, 0, 5]) < 1e-12 * abs(G[0, 1, 0, 0])
    # xy at (1,1,0) equals yz at (0,1,1) (axis relabelling)
    assert abs(G[1, 1, 0, 1] - G[0, 1, 1, 4]) < 1e-12 * abs(G[1, 1, 0, 1])
