"""FP64 restatement of the JVIE N-operator path (test infrastructure only).

Arrays follow MATLAB's column-major convention: a grid array of shape
(L, M, N[, k]) is stored with ``order='F'`` semantics, so the linear voxel
index is ``l + L*(m + M*n)`` (0-based) and component ``c`` of a 3-vector field
occupies ``[c*V, (c+1)*V)`` with ``V = L*M*N``.  ``idx`` holds 1-based linear
indices into that 3V array, exactly as the reference's ``idxS3``.
"""
import numpy as np

# Tensor component order xx, xy, xz, yy, yz, zz (pc/src_vie/src_operators/
# vector_projection.m:15-22) and the (c, d) -> component map used by the
# matvec (pc/src_vie/src_mvp/mv_AN_const.m:109-124).
COMP = np.array([[0, 1, 2], [1, 3, 4], [2, 4, 5]])

# Per-component parity (+1 even / -1 odd) in x, y, z; the products reproduce
# the sign table of pc/src_vie/src_operators/gperiodic_coeff_nop.m:4-31.
PARITY = np.array([
    [+1, +1, +1],   # xx
    [-1, -1, +1],   # xy
    [-1, +1, -1],   # xz
    [+1, +1, +1],   # yy
    [+1, -1, -1],   # yz
    [+1, +1, +1],   # zz
])

_CUBE_SIGNS = {  # gperiodic_coeff_nop.m:4-31, verbatim table of signs
    'L':   [+1, -1, -1, +1, +1, +1],
    'M':   [+1, -1, +1, +1, -1, +1],
    'N':   [+1, +1, -1, +1, -1, +1],
    'LM':  [+1, +1, -1, +1, -1, +1],
    'LN':  [+1, -1, +1, +1, -1, +1],
    'MN':  [+1, -1, -1, +1, +1, +1],
    'LMN': [+1, +1, +1, +1, +1, +1],
}


def gperiodic_coeff_nop(cube):
    """Sign vector for one reflected cube (gperiodic_coeff_nop.m:1-32)."""
    return np.array(_CUBE_SIGNS[cube], dtype=float)


def circulant_nop_const(G):
    """Embed the L x M x N x 6 Toeplitz tensor into its 2L x 2M x 2N circulant.

    Restates pc/src_vie/src_operators/circulant_nop_const.m:1-40: octant 1 holds
    G, plane index L (0-based) stays zero, reflected copies carry the parity
    signs of gperiodic_coeff_nop.
    """
    L, M, N, _ = G.shape
    Gp = np.zeros((2 * L, 2 * M, 2 * N, 6), dtype=complex)
    Gp[:L, :M, :N] = G
    rl = slice(L - 1, 0, -1)   # MATLAB L:-1:2
    rm = slice(M - 1, 0, -1)
    rn = slice(N - 1, 0, -1)
    sl, sm, sn = slice(L + 1, 2 * L), slice(M + 1, 2 * M), slice(N + 1, 2 * N)
    for ii in range(6):
        Gp[sl, :M, :N, ii] = G[rl, :, :, ii] * gperiodic_coeff_nop('L')[ii]
        Gp[:L, sm, :N, ii] = G[:, rm, :, ii] * gperiodic_coeff_nop('M')[ii]
        Gp[:L, :M, sn, ii] = G[:, :, rn, ii] * gperiodic_coeff_nop('N')[ii]
        Gp[sl, sm, :N, ii] = G[rl, rm, :, ii] * gperiodic_coeff_nop('LM')[ii]
        Gp[sl, :M, sn, ii] = G[rl, :, rn, ii] * gperiodic_coeff_nop('LN')[ii]
        Gp[:L, sm, sn, ii] = G[:, rm, rn, ii] * gperiodic_coeff_nop('MN')[ii]
    Gp[sl, sm, sn, :] = G[rl, rm, rn, :]
    return Gp


def fft_operator(A, workers=None):
    """Per-component unnormalised 3D FFT (pc/src_vie/src_operators/fft_operator.m:1-8)."""
    if workers and workers > 1:
        import scipy.fft as sfft
        return sfft.fftn(A, axes=(0, 1, 2), workers=workers)
    return np.fft.fftn(A, axes=(0, 1, 2))


def fN_from_toeplitz(G, workers=None):
    """fN = fft_operator(circulant_nop_const(G)) (getOPERATORS_SAM.m:89-91)."""
    return fft_operator(circulant_nop_const(G), workers)


def _scatter(JIn0, idx, L, M, N):
    JIn = np.zeros(3 * L * M * N, dtype=complex)
    JIn[np.asarray(idx, dtype=np.int64) - 1] = JIn0
    return JIn.reshape((L, M, N, 3), order='F')


def mv_AN_const(JIn0, fN, Mr, Mc, Gram, transp_flag, idx, workers=None):
    """JOut = [Gram*Mr.*J - Mc.*N(J)](idx) via zero-padded FFT convolution.

    Restates pc/src_vie/src_mvp/mv_AN_const.m:1-146 (CPU branch, GPU_flag=0):
    scatter (:52-58), 'notransp' (:102-132), 'transp' (:66-99), gather (:145).
    ``Mr`` and ``Mc`` are L x M x N complex arrays; ``Gram`` is a real scalar.
    ``workers`` > 1 runs the FFTs multi-threaded (scipy.fft; same pocketfft
    algorithm as numpy.fft).
    """
    if workers and workers > 1:
        import scipy.fft as sfft
        fftn = lambda a, s=None: sfft.fftn(a, s=s, axes=(0, 1, 2), workers=workers)   # noqa: E731
        ifftn = lambda a: sfft.ifftn(a, axes=(0, 1, 2), workers=workers)               # noqa: E731
    else:
        fftn = lambda a, s=None: np.fft.fftn(a, s=s, axes=(0, 1, 2))                  # noqa: E731
        ifftn = lambda a: np.fft.ifftn(a, axes=(0, 1, 2))                              # noqa: E731
    Lp, Mp, Np = fN.shape[:3]
    L, M, N = Mr.shape
    JIn = _scatter(JIn0, idx, L, M, N)
    JOut = np.zeros((L, M, N, 3), dtype=complex)
    if transp_flag == 'notransp':
        fJ = [fftn(JIn[..., d], s=(Lp, Mp, Np)) for d in range(3)]
        for c in range(3):
            acc = sum(fN[..., COMP[c, d]] * fJ[d] for d in range(3))
            Jc = ifftn(acc)[:L, :M, :N]
            JOut[..., c] = Gram * Mr * JIn[..., c] - Mc * Jc
    elif transp_flag == 'transp':
        Mcc = np.conj(Mc)
        fJ = [fftn(np.conj(Mcc * JIn[..., d]), s=(Lp, Mp, Np)) for d in range(3)]
        for c in range(3):
            acc = sum(fN[..., COMP[c, d]] * fJ[d] for d in range(3))
            Jc = ifftn(acc)[:L, :M, :N]
            JOut[..., c] = Gram * np.conj(Mr) * JIn[..., c] - np.conj(Jc)
    else:
        raise ValueError("transp_flag must be 'transp' or 'notransp'")
    return JOut.reshape(-1, order='F')[np.asarray(idx, dtype=np.int64) - 1]


def toeplitz_value(G, dx, dy, dz, k):
    """Signed Toeplitz tensor entry T_k(dx,dy,dz) for any integer offsets.

    Offsets are target minus source.  Negative offsets pick up the
    component's parity, which is the circulant embedding's convention
    (circulant_nop_const.m:27-40 with gperiodic_coeff_nop.m).
    """
    s = 1.0
    if dx < 0:
        s *= PARITY[k, 0]
    if dy < 0:
        s *= PARITY[k, 1]
    if dz < 0:
        s *= PARITY[k, 2]
    return s * G[abs(dx), abs(dy), abs(dz), k]


def dense_N(G, L=None, M=None, N=None):
    """Dense 3V x 3V matrix of the N operator on an L x M x N grid (tiny grids).

    Row (c, target voxel), column (d, source voxel); entry T_{cd}(r_t - r_s).
    This is the closed-form Toeplitz product that the FFT path must equal.
    """
    Lg, Mg, Ng, _ = G.shape
    L = L or Lg
    M = M or Mg
    N = N or Ng
    V = L * M * N
    ll, mm, nn = np.meshgrid(np.arange(L), np.arange(M), np.arange(N), indexing='ij')
    ll = ll.reshape(-1, order='F')
    mm = mm.reshape(-1, order='F')
    nn = nn.reshape(-1, order='F')
    D = np.zeros((3 * V, 3 * V), dtype=complex)
    dxs = ll[:, None] - ll[None, :]
    dys = mm[:, None] - mm[None, :]
    dzs = nn[:, None] - nn[None, :]
    for c in range(3):
        for d in range(3):
            k = COMP[c, d]
            sign = (np.where(dxs < 0, PARITY[k, 0], 1) *
                    np.where(dys < 0, PARITY[k, 1], 1) *
                    np.where(dzs < 0, PARITY[k, 2], 1))
            D[c * V:(c + 1) * V, d * V:(d + 1) * V] = \
                sign * G[np.abs(dxs), np.abs(dys), np.abs(dzs), k]
    return D


def dense_AN(G, Mr, Mc, Gram, idx):
    """Dense JVIE-II system matrix restricted to idx (tiny grids)."""
    L, M, N = Mr.shape
    D = dense_N(G, L, M, N)
    mr = np.tile(Mr.reshape(-1, order='F'), 3)
    mc = np.tile(Mc.reshape(-1, order='F'), 3)
    A = Gram * np.diag(mr) - mc[:, None] * D
    i0 = np.asarray(idx, dtype=np.int64) - 1
    return A[np.ix_(i0, i0)]


def an_rows(G, Mr, Mc, Gram, idx, x, rows, shape):
    """Selected entries of y = A x, A = Gram*Mr - Mc.*N on idx, by direct
    Toeplitz sums (the product mv_AN_const.m:102-146 evaluates by FFT).

    ``G`` holds the Toeplitz tensor at offsets 0 .. D-1 per axis with D at
    least the extent of the unknowns' voxels, ``Mr``/``Mc`` are per-entry of
    idx (length n) or per voxel (``shape``), ``rows`` 0-based positions in
    idx.  Cost O(n) per row: usable at full size for a handful of rows.
    Negative offsets take the parity signs of the dense product above.
    """
    L, M, N = shape
    V = L * M * N
    i0 = np.asarray(idx, dtype=np.int64) - 1
    comp, vox = np.divmod(i0, V)
    vi = vox % L
    vj = (vox // L) % M
    vk = vox // (L * M)
    x = np.asarray(x, dtype=complex)

    def per_entry(a, r):
        a = np.asarray(a)
        if a.ndim == 0:
            return complex(a)
        if a.size == i0.size and a.ndim == 1:
            return complex(a[r])
        return complex(a.reshape(-1, order='F')[vox[r]])

    out = np.zeros(len(rows), dtype=complex)
    for t, r in enumerate(rows):
        c = int(comp[r])
        dx = vi[r] - vi
        dy = vj[r] - vj
        dz = vk[r] - vk
        acc = 0j
        for d in range(3):
            sel = comp == d
            k = COMP[c, d]
            sx, sy, sz = dx[sel], dy[sel], dz[sel]
            sign = (np.where(sx < 0, PARITY[k, 0], 1) * np.where(sy < 0, PARITY[k, 1], 1) *
                    np.where(sz < 0, PARITY[k, 2], 1))
            acc += np.sum(sign * G[np.abs(sx), np.abs(sy), np.abs(sz), k] * x[sel])
        out[t] = Gram * per_entry(Mr, r) * x[r] - per_entry(Mc, r) * acc
    return out


def idx_from_mask(mask):
    """idxS3 construction (pc/src_vie/src_solver/JVIE_Solver_const.m:148-149).

    ``mask`` is an L x M x N boolean array of unknown voxels.  Returns 1-based
    int64 indices [idxS; V+idxS; 2V+idxS] with idxS sorted (MATLAB ``find``).
    """
    V = mask.size
    idxS = np.flatnonzero(mask.reshape(-1, order='F')) + 1
    return np.concatenate([idxS, V + idxS, 2 * V + idxS]).astype(np.int64)


# -- constants (pc/src_vie/src_solver/EMconstants.m:13-20) -----------------
MU0 = 4 * np.pi * 1e-7
C0 = 299792458.0
EPS0 = 1.0 / C0 ** 2 / MU0


def em_constants(freq):
    """omega, ko, lambda for a frequency (EMconstants.m:13-20)."""
    omega = 2 * np.pi * freq
    lam = C0 / freq
    ko = 2 * np.pi / lam
    return omega, ko, lam


def jvie2_rhs(Einc, Mc2, Gram, freq, idx):
    """Vrhs = Gram * (j*omega*eps0*Mc) .* Einc(idx) under JVIE-II scaling.

    Restates JVIE_Solver_const.m:151-157 followed by the JVIE-II division by
    Mr (:243-251), which turns Mc=e_r-1 into Mc2=(e_r-1)/e_r; also the form
    used directly by Bragg_MOR_script.m:20-23.  ``Einc`` is L x M x N x 3.
    """
    omega, _, _ = em_constants(freq)
    tau = 1j * omega * EPS0 * Mc2
    tau3 = np.concatenate([tau.reshape(-1, order='F')] * 3)
    i0 = np.asarray(idx, dtype=np.int64) - 1
    return Gram * tau3[i0] * Einc.reshape(-1, order='F')[i0]
