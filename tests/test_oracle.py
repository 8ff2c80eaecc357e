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


# ---------------------------------------------------------------------------
# Circulant preconditioner producers pinned by what the mathematics fixes:
# the Chan circulant of a matrix is its orthogonal (Frobenius) projection onto
# the circulants, so the block producers must equal that projection of the
# dense JVIE-II operator, computed here without any of their sign tables.
# ---------------------------------------------------------------------------

def _frobenius_fit(A, L, M, N, levels):
    """Orthogonal projection of the 3V x 3V field-order matrix A (field index
    l + L*(m + M*(n + N*c))) onto matrices that are block circulant in x
    (levels=1: blocks indexed by (m, n, c)) or in x and y (levels=2: blocks
    indexed by (n, c)), returned in the Fourier basis the applies use:
    levels=2: B[kx, ky] (3N x 3N, index n + N*c); levels=1: B[kx] (3MN x 3MN,
    index m + M*(n + N*c)).  The coefficient of the circulant shift (p, q) is
    the mean of A over the entries at that cyclic offset."""
    if levels == 2:
        A6 = A.reshape(3 * N, M, L, 3 * N, M, L)
        cf = np.zeros((L, M, 3 * N, 3 * N), complex)
        mi, li = np.arange(M)[:, None], np.arange(L)[None, :]
        for p in range(L):
            for q in range(M):
                S = A6[:, (mi + q) % M, (li + p) % L, :, mi, li]      # (M, L, 3N, 3N)
                cf[p, q] = S.sum(axis=(0, 1)) / (L * M)
        return np.fft.fft2(cf, axes=(0, 1))
    A4 = A.reshape(3 * M * N, L, 3 * M * N, L)
    cf = np.zeros((L, 3 * M * N, 3 * M * N), complex)
    li = np.arange(L)
    for p in range(L):
        cf[p] = A4[:, (li + p) % L, :, li].sum(axis=0) / L
    return np.fft.fft(cf, axis=0)


def _full_operator(G, mc, mr0, gram):
    L, M, N, _ = G.shape
    idx = ovie.idx_from_mask(np.ones((L, M, N), bool))
    return ovie.dense_AN(G, np.full((L, M, N), mr0 + 0j), np.full((L, M, N), mc + 0j), gram, idx), idx


@pytest.mark.parametrize('shape', [(4, 3, 2), (3, 4, 2), (5, 3, 3), (2, 2, 1)])
def test_level2_blocks_are_frobenius_optimal(shape):
    L, M, N = shape
    rng = np.random.default_rng(sum(shape) + 40)
    G = random_toeplitz(L, M, N, rng)
    mc, mr0, gram = 0.83 + 0.1j, 1.2, 0.7
    A, _ = _full_operator(G, mc, mr0, gram)
    fit = _frobenius_fit(A, L, M, N, 2)
    blocks = opre.level_2_blocks_direct(G, mc, mr0, L, M, N, gram)
    assert np.abs(blocks - fit).max() <= 1e-13 * np.abs(fit).max()


def test_level2_frobenius_pin_catches_a_sign_slip(monkeypatch):
    # a flipped sign in the y-Toeplitz conventions of one odd component must
    # fail the projection check above (planted in the oracle, on the CPU)
    L, M, N = 4, 3, 2
    G = random_toeplitz(L, M, N, np.random.default_rng(44))
    A, _ = _full_operator(G, 0.83, 1.0, 0.7)
    fit = _frobenius_fit(A, L, M, N, 2)
    good = opre._y_first_row_col
    for k_bad in (1, 2, 4):
        def slipped(ck_x, k, N_, M_, k_bad=k_bad):
            col, row = good(ck_x, k, N_, M_)
            if k == k_bad:
                row = row.copy()
                row[..., 1:] *= -1
            return col, row
        monkeypatch.setattr(opre, '_y_first_row_col', slipped)
        blocks = opre.level_2_blocks_direct(G, 0.83, 1.0, L, M, N, 0.7)
        assert np.abs(blocks - fit).max() > 1e-3 * np.abs(fit).max(), k_bad
    monkeypatch.setattr(opre, '_y_first_row_col', good)


@pytest.mark.parametrize('shape', [(4, 3, 2), (5, 2, 2)])
def test_level1_blocks_are_frobenius_optimal(shape):
    L, M, N = shape
    rng = np.random.default_rng(sum(shape) + 50)
    G = random_toeplitz(L, M, N, rng)
    mc, mr0, gram = 0.6 - 0.2j, 1.0, 0.9
    A, _ = _full_operator(G, mc, mr0, gram)
    fit = _frobenius_fit(A, L, M, N, 1)
    blocks = opre.level_1_blocks(opre.folded_toeplitz(G, mc, mr0, gram))
    assert np.abs(blocks - fit).max() <= 1e-13 * np.abs(fit).max()
    # and the inverse producer is the inverse of those blocks
    inv = opre.level_1_parallel_func(G, mc, mr0, L, M, N, gram)
    assert np.abs(inv - np.linalg.inv(fit)).max() <= 1e-10 * np.abs(inv).max()


def test_level1_face_blocks_from_bare_operator():
    # Bragg_MOR_script.m:61-72 form: inv(Gram I - diag(face3) circ_N) where
    # circ_N is the x-circulant projection of the bare N operator
    L, M, N = 4, 3, 2
    rng = np.random.default_rng(61)
    G = random_toeplitz(L, M, N, rng)
    Mr_mid = 1.0 + rng.random((L, M, N)) * 4
    gram = 0.8
    D = ovie.dense_N(G)
    circ = _frobenius_fit(D, L, M, N, 1)
    hey = Mr_mid.sum(axis=0) / L
    face = ((hey - 1) / hey).reshape(-1, order='F')
    face3 = np.tile(face, 3)
    expect = np.linalg.inv(gram * np.eye(3 * M * N)[None] - face3[None, :, None] * circ)
    got = opre.level_1_parallel_func_N_face(G, Mr_mid, gram)
    assert np.abs(got - expect).max() <= 1e-10 * np.abs(expect).max()


def _periodic_toeplitz(L, M, N, rng):
    """Tensor whose Toeplitz operator is exactly circulant along x and y:
    T(d - L) = T(d), i.e. G[L - i] = parity * G[i] along x (and along y)."""
    G = random_toeplitz(L, M, N, rng)
    for ax, n in ((0, L), (1, M)):
        for i in range(1, n):
            j = n - i
            if j <= i:
                continue
            src = [slice(None)] * 3
            dst = [slice(None)] * 3
            src[ax], dst[ax] = i, j
            G[tuple(dst)] = G[tuple(src)] * ovie.PARITY[:, ax]
        if n % 2 == 0:    # the middle plane is its own mirror: odd components vanish there
            mid = [slice(None)] * 3
            mid[ax] = n // 2
            G[tuple(mid) + (ovie.PARITY[:, ax] < 0,)] = 0.0
    return G


@pytest.mark.parametrize('shape', [(4, 3, 2), (5, 4, 3)])
def test_level2_inverts_a_circulant_operator(shape):
    L, M, N = shape
    rng = np.random.default_rng(sum(shape) + 70)
    G = _periodic_toeplitz(L, M, N, rng)
    assert all(np.abs(G[..., k]).max() > 0.1 for k in range(6))
    mc, mr0, gram = 0.83 + 0.05j, 1.0, 2.5
    A, idx = _full_operator(G, mc, mr0, gram)
    # the operator is block circulant in x and y: its projection is itself
    assert np.abs(_frobenius_fit(A, L, M, N, 2) - opre.level_2_blocks_direct(G, mc, mr0, L, M, N, gram)).max() \
        <= 1e-13 * np.abs(A).max() * L * M
    Binv = opre.level_2_parallel_func(G, mc, mr0, L, M, N, gram)
    x = rng.standard_normal(idx.size) + 1j * rng.standard_normal(idx.size)
    y = opre.chan_2_mvp_for_parallel_idx(Binv, A @ x, L, M, N, idx)
    assert rel(y, x) < 1e-12


def test_chan1_apply_matches_dense_fourier_form():
    rng = np.random.default_rng(13)
    L, M, N = 4, 3, 2
    n3 = 3 * M * N
    B = rng.standard_normal((L, n3, n3)) + 1j * rng.standard_normal((L, n3, n3))
    idx = ovie.idx_from_mask(rng.random((L, M, N)) < 0.7)
    x = rng.standard_normal(idx.size) + 1j * rng.standard_normal(idx.size)
    y = opre.chan_mvp_idx(B, x, L, M, N, idx)
    V = L * M * N
    F = np.kron(np.eye(3 * M * N), np.fft.fft(np.eye(L), axis=0))
    D = np.zeros((3 * V, 3 * V), complex)
    for i in range(L):
        for a in range(n3):
            for b in range(n3):
                D[i + L * a, i + L * b] = B[i, a, b]
    A = np.linalg.inv(F) @ D @ F
    full = np.zeros(3 * V, complex)
    full[idx - 1] = x
    assert rel(y, (A @ full)[idx - 1]) < 1e-13


@pytest.mark.parametrize('k', [0, 1])
def test_chan_x_is_frobenius_optimal(k):
    # component k even (xx) or odd (xy) in x: T(-a) = parity * T(a); chan_x's
    # output must be the DFT of the least-squares circulant of the Toeplitz matrix
    rng = np.random.default_rng(9 + k)
    L = 6
    half = rng.standard_normal(L) + 1j * rng.standard_normal(L)
    if ovie.PARITY[k, 0] < 0:
        half[0] = 0.0
    t = {a: half[a] for a in range(L)}
    t.update({-a: ovie.PARITY[k, 0] * half[a] for a in range(1, L)})
    T = np.array([[t[i - j] for j in range(L)] for i in range(L)])
    basis = [np.roll(np.eye(L), s, axis=0) for s in range(L)]
    Amat = np.stack([b.reshape(-1) for b in basis], axis=1)
    c_ls = np.linalg.lstsq(Amat, T.reshape(-1), rcond=None)[0]
    A = np.zeros((L, 1, 1), complex)
    A[:, 0, 0] = half
    got = opre.chan_x(A, 0 if k == 0 else 1)[:, 0, 0]
    assert rel(got, np.fft.fft(c_ls)) < 1e-13


def test_prec_fun_disk_idx_is_block_jacobi():
    # two disjoint sub-boxes of one grid: the composite applies the 1-level
    # form on the first and the 2-level form on the second, passes nothing else
    rng = np.random.default_rng(21)
    L, M, N = 6, 7, 2
    mask = np.zeros((L, M, N), bool)
    mask[:, 0:2, :] = True           # "bus" box: all of x, y in [0, 2)
    mask[1:5, 3:7, :] = rng.random((4, 4, N)) < 0.8   # "disk" box x in [1,5), y in [3,7)
    idxS3 = ovie.idx_from_mask(mask)
    V = L * M * N
    x = rng.standard_normal(idxS3.size) + 1j * rng.standard_normal(idxS3.size)
    # bus box: whole-box index set idx_T into the 3V field
    Lw, Mw, Nw = L, 2, N
    bl, bm, bn = np.meshgrid(np.arange(Lw), np.arange(Mw), np.arange(Nw), indexing='ij')
    bvox = (bl + L * (bm + M * bn)).reshape(-1, order='F')
    idx_T = np.concatenate([bvox + c * V for c in range(3)]) + 1
    # disk box: box-local index set idx_Sq of its unknowns and the field set idx_R
    Ld, Md, Nd = 4, 4, N
    dmask = mask[1:5, 3:7, :]
    idx_Sq = ovie.idx_from_mask(dmask)
    dl, dm, dn = np.meshgrid(np.arange(Ld), np.arange(Md), np.arange(Nd), indexing='ij')
    dvox = ((dl + 1) + L * ((dm + 3) + M * dn)).reshape(-1, order='F')
    dfield = np.concatenate([dvox + c * V for c in range(3)])
    idx_R = dfield[idx_Sq - 1] + 1
    nw = 3 * Mw * Nw
    Bw = rng.standard_normal((Lw, nw, nw)) + 1j * rng.standard_normal((Lw, nw, nw))
    Bd = rng.standard_normal((Ld, Md, 3 * Nd, 3 * Nd)) + 1j * rng.standard_normal((Ld, Md, 3 * Nd, 3 * Nd))
    y = opre.prec_fun_disk_idx(x, Bw, idx_T, idx_R, Lw, Mw, Nw, Bd, Ld, Md, Nd, L, M, N, idxS3, idx_Sq)
    full = np.zeros(3 * V, complex)
    full[idxS3 - 1] = x
    expect = full.copy()
    expect[idx_T - 1] = opre.chan_mvp(Bw, full[idx_T - 1], Lw, Mw, Nw)
    expect[idx_R - 1] = opre.chan_2_mvp_for_parallel_idx(Bd, full[idx_R - 1], Ld, Md, Nd, idx_Sq)
    assert rel(y, expect[idxS3 - 1]) < 1e-14
    # the bus apply is the dense Fourier form on its box (pinned above)
    assert not np.allclose(y, x)


def test_toeplitz_value_matches_dense_operator():
    rng = np.random.default_rng(31)
    L, M, N = 3, 3, 2
    G = random_toeplitz(L, M, N, rng)
    D = ovie.dense_N(G)
    V = L * M * N
    for (lt, mt, nt, ls, ms, ns) in [(0, 0, 0, 2, 1, 1), (2, 1, 0, 0, 2, 1), (1, 2, 1, 2, 0, 0)]:
        for c in range(3):
            for d in range(3):
                k = ovie.COMP[c, d]
                row = lt + L * (mt + M * nt) + c * V
                col = ls + L * (ms + M * ns) + d * V
                assert ovie.toeplitz_value(G, lt - ls, mt - ms, nt - ns, k) == D[row, col]


def test_lagrange_int_reproduces_polynomials():
    from oracle import mor as omor
    s = np.array([1.0, 1.3, 1.7, 2.2])
    coef = np.array([0.5, -1.0, 2.0, 0.25])
    F = [np.full((2, 2), np.polyval(coef, si)) for si in s]
    for S in (1.1, 1.5, 2.0):
        assert np.abs(omor.lagrange_int(F, S, s) - np.polyval(coef, S)).max() < 1e-12


def test_bicgstab_matches_direct_solve_and_true_residuals():
    rng = np.random.default_rng(12)
    n = 40
    A = np.eye(n) * 8 + rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    seen = []

    def Af(v):
        return A @ v
    x, flag, relres, it, rv = okry.bicgstab(Af, b, 1e-12, 200)
    assert flag == 0 and relres <= 1e-12
    assert rel(x, np.linalg.solve(A, b)) < 1e-10
    assert len(rv) == int(2 * it) + 1
    # preconditioned with the exact inverse: done after the first half step
    Ainv = np.linalg.inv(A)
    x, flag, relres, it, rv = okry.bicgstab(Af, b, 1e-12, 5, M=lambda v: Ainv @ v)
    assert flag == 0 and it == 0.5 and rel(x, np.linalg.solve(A, b)) < 1e-12


def test_bicgstab_recurrence_residual_is_the_true_residual():
    # in exact arithmetic r_k = b - A x_k: recorded norms equal the true ones
    rng = np.random.default_rng(13)
    n = 30
    A = np.eye(n) * 6 + rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    for k in range(1, 8):
        x, flag, relres, it, rv = okry.bicgstab(lambda v: A @ v, b, 1e-30, k)
        assert abs(rv[-1] - np.linalg.norm(b - A @ x) / np.linalg.norm(b)) <= 1e-12 * max(rv[-1], 1e-3)


def test_an_rows_equal_dense_product_rows():
    # the sampled-row product (direct Toeplitz sums) against the dense
    # operator, with Mr/Mc per voxel and per entry, and a tensor that holds
    # more offsets than the unknowns span
    from tests.helpers import random_problem
    G, Mr, Mc, Gram, idx, x = random_problem(6, 5, 4, seed=21, with_mr=True)
    A = ovie.dense_AN(G, Mr, Mc, Gram, idx)
    rows = [0, 3, idx.size // 2, idx.size - 1]
    y = A @ x
    assert np.abs(ovie.an_rows(G, Mr, Mc, Gram, idx, x, rows, (6, 5, 4)) - y[rows]).max() <= 1e-13 * np.abs(y).max()
    i0 = idx - 1
    mr_e = np.tile(Mr.reshape(-1, order='F'), 3)[i0]
    mc_e = np.tile(Mc.reshape(-1, order='F'), 3)[i0]
    assert np.abs(ovie.an_rows(G, mr_e, mc_e, Gram, idx, x, rows, (6, 5, 4)) - y[rows]).max() <= \
        1e-13 * np.abs(y).max()
    Gbig = np.concatenate([G, np.ones((3, 5, 4, 6), complex)], axis=0)
    assert np.abs(ovie.an_rows(Gbig, Mr, Mc, Gram, idx, x, rows, (6, 5, 4)) - y[rows]).max() <= \
        1e-13 * np.abs(y).max()
