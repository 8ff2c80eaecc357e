"""FP64 restatement of the Chan circulant preconditioners (test infrastructure only).

Block containers:
  * 1-level: ``blocks[i]`` is the 3MN x 3MN matrix for x-frequency i
    (MATLAB cell ``inv_blocks{i+1}``), array shape (L, 3MN, 3MN).
  * 2-level: ``blocks[i, j]`` is the 3N x 3N matrix for (kx, ky) = (i, j)
    (MATLAB cell ``inv_blocks{i+1, j+1}``), array shape (L, M, 3N, 3N).
Vectors inside a block follow the reference's reshape order: for a 1-level
block the index is m + M*(n + N*c); for a 2-level block it is n + N*c.
"""
import numpy as np
from scipy.linalg import toeplitz


def _scatter(JIn, idx, total):
    V = np.zeros(total, dtype=complex)
    V[np.asarray(idx, dtype=np.int64) - 1] = JIn
    return V


def chan_mvp(inv_blocks, Vrhs, L, M, N):
    """1-level apply on a full 3LMN vector (pc/src_preconditioning/chan_mvp.m:2-12)."""
    temp = np.asarray(Vrhs, dtype=complex).reshape((L, 3 * M * N), order='F')
    temp = np.fft.fft(temp, axis=0).T
    for i in range(L):
        temp[:, i] = inv_blocks[i] @ temp[:, i]
    temp = np.fft.ifft(temp.T, axis=0)
    return temp.reshape(-1, order='F')


def chan_mvp_idx(inv_blocks, JIn, L, M, N, idx):
    """1-level apply with scatter/gather (chan_mvp_idx.m:2-16)."""
    V = _scatter(JIn, idx, 3 * L * M * N)
    out = chan_mvp(inv_blocks, V, L, M, N)
    return out[np.asarray(idx, dtype=np.int64) - 1]


def chan_2_mvp(inv_blocks, Vrhs, L, M, N):
    """2-level apply on a full 3LMN vector (body of chan_2_mvp_for_parallel_idx.m:7-20)."""
    temp = np.asarray(Vrhs, dtype=complex).reshape((L, 3 * M * N), order='F')
    temp = np.fft.fft(temp, axis=0).T.copy()
    for i in range(L):
        T = temp[:, i].reshape((M, 3 * N), order='F')
        T = np.fft.fft(T, axis=0).T.copy()
        for j in range(M):
            T[:, j] = inv_blocks[i, j] @ T[:, j]
        T = np.fft.ifft(T.T, axis=0)
        temp[:, i] = T.reshape(-1, order='F')
    temp = np.fft.ifft(temp.T, axis=0)
    return temp.reshape(-1, order='F')


def chan_2_mvp_for_parallel_idx(inv_blocks, JIn, L, M, N, idx):
    """2-level apply with scatter/gather (chan_2_mvp_for_parallel_idx.m:2-23)."""
    V = _scatter(JIn, idx, 3 * L * M * N)
    out = chan_2_mvp(inv_blocks, V, L, M, N)
    return out[np.asarray(idx, dtype=np.int64) - 1]


def prec_fun_disk_idx(J, circ_WG_inv, idx_T, idx_R, L_WG, M_WG, N_WG,
                      circ_disk_inv, L_D, M_D, N_D, L, M, N, idxS3, idx_Sq):
    """Block-Jacobi composite (pc/src_preconditioning/prec_fun_disk_idx.m:1-23).

    1-level apply on the bus-waveguide index set ``idx_T`` and 2-level apply on
    the disk-box set ``idx_R``; every index argument is 1-based.
    """
    Jt = np.zeros(3 * L * M * N, dtype=complex)
    Jt[np.asarray(idxS3) - 1] = J
    JT = Jt[np.asarray(idx_T) - 1]
    JR = Jt[np.asarray(idx_R) - 1]
    JT = chan_mvp(circ_WG_inv, JT, L_WG, M_WG, N_WG)
    JR = chan_2_mvp_for_parallel_idx(circ_disk_inv, JR, L_D, M_D, N_D, idx_Sq)
    Jt[np.asarray(idx_T) - 1] = JT
    Jt[np.asarray(idx_R) - 1] = JR
    return Jt[np.asarray(idxS3) - 1]


# ---------------------------------------------------------------------------
# Setup (block producers)
# ---------------------------------------------------------------------------

_EVEN_X = (True, False, False, True, True, True)   # + or - reflected term in x


def chan_x(A, k):
    """Chan optimal circulant along x, then fft along x.

    Restates level_1_parallel_func.m:12-21 (and the per-component copies at
    :26-33, :38-44, ...): c_1 = A_1, c_i = (L+1-i)/L*A_i +/- (i-1)/L*A_{L-i+2};
    the minus sign is used for components odd in x (xy, xz).
    """
    L = A.shape[0]
    c = np.zeros_like(A, dtype=complex)
    c[0] = A[0]
    sgn = 1.0 if _EVEN_X[k] else -1.0
    for i in range(2, L + 1):
        c[i - 1] = (L + 1 - i) / L * A[i - 1] + sgn * (i - 1) / L * A[L - i + 1]
    return np.fft.fft(c, axis=0)


def folded_toeplitz(opToeplitz, Mc, Mr, Gram):
    """A_k = -Mc.*T_k, with Gram*Mr(1) added to the (1,1,1) entry of xx, yy, zz.

    Restates level_2_parallel_func.m:5-7,24,36,47-48,61,74-75 (Mc may be a
    scalar or an L x M x N array, used elementwise exactly as the reference).
    """
    A = -np.asarray(Mc)[..., None] * opToeplitz if np.ndim(Mc) else -Mc * opToeplitz
    A = np.array(A, dtype=complex)
    mr0 = np.asarray(Mr).reshape(-1, order='F')[0] if np.ndim(Mr) else Mr
    for k in (0, 3, 5):
        A[0, 0, 0, k] += Gram * mr0
    return A


def _component_blocks_literal(cks, i_loop, M, N):
    """3MN x 3MN matrix for one kx, literal restatement of
    level_1_parallel_func.m:113-188 / level_2_parallel_func.m:86-162."""
    cc, CC, C3, C4, C5, C6 = cks
    MN = M * N
    temp = np.zeros((3 * MN, 3 * MN), dtype=complex)
    idx = toeplitz(np.arange(N))

    def blockmat(chan):
        return np.block([[chan[idx[a, b]] for b in range(N)] for a in range(N)])

    chan = [toeplitz(cc[i_loop, :, i], cc[i_loop, :, i]) for i in range(N)]
    temp[:MN, :MN] = blockmat(chan)
    chan = [toeplitz(np.concatenate([[-CC[i_loop, 0, i]], CC[i_loop, 1:, i]]),
                     -CC[i_loop, :, i]) for i in range(N)]
    r = blockmat(chan)
    temp[:MN, MN:2 * MN] = r
    temp[MN:2 * MN, :MN] = r
    chan = [toeplitz(-C3[i_loop, :, i], -C3[i_loop, :, i]) for i in range(N)]
    r = blockmat(chan)
    U = np.triu(r)
    r = U - U.T + np.diag(np.diag(U))
    temp[:MN, 2 * MN:] = r
    temp[2 * MN:, :MN] = r
    chan = [toeplitz(C4[i_loop, :, i], C4[i_loop, :, i]) for i in range(N)]
    temp[MN:2 * MN, MN:2 * MN] = blockmat(chan)
    # MATLAB toeplitz([C5(1) -C5(2:M)], C5(1:M)): first column then first row
    chan = [toeplitz(np.concatenate([[C5[i_loop, 0, i]], -C5[i_loop, 1:, i]]),
                     C5[i_loop, :, i]) for i in range(N)]
    r = blockmat(chan)
    U = np.triu(r)
    r = U + U.T - np.diag(np.diag(U))
    temp[MN:2 * MN, 2 * MN:] = r
    temp[2 * MN:, MN:2 * MN] = r
    chan = [toeplitz(C6[i_loop, :, i], C6[i_loop, :, i]) for i in range(N)]
    temp[2 * MN:, 2 * MN:] = blockmat(chan)
    return temp


def level_1_blocks(A):
    """Per-kx 3MN x 3MN circulant blocks (no inversion) from a folded tensor.

    ``A`` is L x M x N x 6 (already material-folded or the bare Toeplitz tensor).
    Restates level_1_parallel_func.m:5-169 and level_1_parallel_func_N.m:5-192
    up to (not including) the inverse.
    """
    L, M, N, _ = A.shape
    cks = [chan_x(A[..., k], k) for k in range(6)]
    return np.stack([_component_blocks_literal(cks, i, M, N) for i in range(L)])


def level_1_parallel_func(opToeplitz, Mc, Mr, L, M, N, Gram):
    """Inverse 1-level blocks (level_1_parallel_func.m:1-169)."""
    A = folded_toeplitz(opToeplitz, Mc, Mr, Gram)
    return np.linalg.inv(level_1_blocks(A))


def level_1_parallel_func_N_face(opToeplitz, Mr_mid, Gram):
    """Face-averaged 1-level inverse blocks (level_1_parallel_func_N.m:1-192
    followed by Bragg_MOR_script.m:61-72): inv(Gram*I - diag(face3)*circ_N{i})
    with face = (x-mean(Mr) - 1)/x-mean(Mr) (real part)."""
    L, M, N, _ = opToeplitz.shape
    circ_N = level_1_blocks(np.asarray(opToeplitz, dtype=complex))
    hey = np.real(np.sum(Mr_mid, axis=0) / L)
    face = (hey - 1.0) / hey
    face3 = np.tile(face.reshape(-1, order='F'), 3)
    eye = np.eye(3 * M * N)
    return np.stack([np.linalg.inv(Gram * eye - face3[:, None] * circ_N[i])
                     for i in range(L)])


def level_2_blocks_literal(opToeplitz, Mc, Mr, L, M, N, Gram):
    """2-level blocks (before inversion), literal restatement of
    level_2_parallel_func.m:1-192 including the 3MN x 3MN intermediate
    (tiny grids only)."""
    A = folded_toeplitz(opToeplitz, Mc, Mr, Gram)
    cks = [chan_x(A[..., k], k) for k in range(6)]
    out = np.zeros((L, M, 3 * N, 3 * N), dtype=complex)
    for i_loop in range(L):
        temp = _component_blocks_literal(cks, i_loop, M, N)
        diag_t = np.zeros_like(temp)
        c = np.zeros(M, dtype=complex)
        for k in range(3 * N):
            for j in range(3 * N):
                row_vec = temp[j * M, k * M:(k + 1) * M]
                col_vec = temp[j * M:(j + 1) * M, k * M]
                c[0] = col_vec[0]
                for i in range(2, M + 1):
                    c[i - 1] = (M + 1 - i) / M * col_vec[i - 1] + \
                        (i - 1) / M * row_vec[M - i + 1]
                diag_t[j * M:(j + 1) * M, k * M:(k + 1) * M] = np.diag(np.fft.fft(c))
        X = diag_t.T.reshape((3 * M * M * N, 3 * N), order='F').T
        coll = X.reshape((3 * M * N, 3 * M * N), order='F')
        for m in range(M):
            out[i_loop, m] = coll[m * 3 * N:(m + 1) * 3 * N, m * 3 * N:(m + 1) * 3 * N]
    return out


def _y_first_row_col(ck_x, k, N, M):
    """First column f(+dy) and first row f(-dy) of every M x M block (n_t, n_s)
    of component k, for one kx; shape (N, N, M) each.  Encodes the toeplitz /
    triu sign conventions of level_2_parallel_func.m:94-162."""
    col = np.zeros((N, N, M), dtype=complex)
    row = np.zeros((N, N, M), dtype=complex)
    for nt in range(N):
        for ns in range(N):
            dz = abs(nt - ns)
            v = ck_x[:, dz]
            if k in (0, 3, 5):                    # symmetric toeplitz
                col[nt, ns] = v
                row[nt, ns] = v
            elif k == 1:                          # xy: odd in y, even in z
                col[nt, ns] = np.concatenate([[-v[0]], v[1:]])
                row[nt, ns] = -v
            elif k == 2:                          # xz: -C3 symmetric, triu(+lower=-upper^T)
                if nt < ns:
                    col[nt, ns] = -v
                    row[nt, ns] = -v
                elif nt > ns:
                    col[nt, ns] = v
                    row[nt, ns] = v
                else:
                    col[nt, ns] = np.concatenate([[-v[0]], v[1:]])
                    row[nt, ns] = -v
            elif k == 4:                          # yz: odd in y and z
                if nt < ns:
                    col[nt, ns] = np.concatenate([[v[0]], -v[1:]])
                    row[nt, ns] = v
                elif nt > ns:
                    col[nt, ns] = v
                    row[nt, ns] = np.concatenate([[v[0]], -v[1:]])
                else:
                    col[nt, ns] = v
                    row[nt, ns] = v
    return col, row


def level_2_blocks_direct(opToeplitz, Mc, Mr, L, M, N, Gram):
    """2-level 3N x 3N blocks built directly (no 3MN x 3MN intermediate).

    Same result as ``level_2_blocks_literal``; this is the construction a
    device setup kernel follows: x-Chan + fft per component, then for every
    (c, n_t, d, n_s) the y-Chan coefficients of the signed Toeplitz block and
    their fft along y.
    """
    A = folded_toeplitz(opToeplitz, Mc, Mr, Gram)
    cks = [chan_x(A[..., k], k) for k in range(6)]
    COMP = np.array([[0, 1, 2], [1, 3, 4], [2, 4, 5]])
    out = np.zeros((L, M, 3 * N, 3 * N), dtype=complex)
    wgt_a = (M - np.arange(M)) / M          # (M+1-i)/M for i = 1..M
    wgt_b = np.arange(M) / M                # (i-1)/M
    for i_loop in range(L):
        for c in range(3):
            for d in range(3):
                k = COMP[c, d]
                col, row = _y_first_row_col(cks[k][i_loop], k, N, M)
                # chan: c_1 = col_1, c_i = wa*col_i + wb*row_{M-i+2}
                rowrev = np.concatenate([row[..., :1], row[..., :0:-1]], axis=-1)
                cy = wgt_a * col + wgt_b * rowrev
                cy[..., 0] = col[..., 0]
                ey = np.fft.fft(cy, axis=-1)          # (N, N, M)
                for m in range(M):
                    out[i_loop, m, c * N:(c + 1) * N, d * N:(d + 1) * N] = ey[:, :, m]
    return out


def level_2_face_blocks(opToeplitz, face, Gram):
    """Face-averaged 2-level inverse blocks inv(Gram*I - diag(face3)*N2).

    The form of Bragg_MOR_script.m:61-72 (circ_N from level_1_parallel_func_N
    without material, then Gram*I - diag(face3)*circ_N) with the circulant
    approximation of level_2_parallel_func.m taken along the first two axes
    and ``face`` (length N) the contrast along the exact third axis.
    """
    L, M, N, _ = opToeplitz.shape
    N2 = level_2_blocks_direct(np.asarray(opToeplitz, dtype=complex), -1.0, 0.0, L, M, N, 0.0)
    face3 = np.tile(np.asarray(face, dtype=complex), 3)
    eye = np.eye(3 * N)
    return np.linalg.inv(Gram * eye - face3[None, None, :, None] * N2)


def level_2_parallel_func(opToeplitz, Mc, Mr, L, M, N, Gram):
    """Inverse 2-level blocks (level_2_parallel_func.m:1-192)."""
    return np.linalg.inv(level_2_blocks_direct(opToeplitz, Mc, Mr, L, M, N, Gram))
