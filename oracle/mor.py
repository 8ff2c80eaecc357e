"""FP64 restatement of the wavelength sweep by model order reduction (test
infrastructure only): Lagrange interpolation of operators over the sample
wavelengths, the moment basis of Bragg_MOR_script.m and the reduced solve
of example_Bragg_1500nm_to_1600nm.m, on dense matrices of tiny grids.
"""
import numpy as np


def lagrange_weights(S, s):
    """L_i(S) of lagrange_int.m:1-13 (sum_i L_i(S) F_i interpolates F)."""
    s = np.asarray(s, dtype=float)
    w = np.ones(s.size)
    for i in range(s.size):
        for j in range(s.size):
            if j != i:
                w[i] *= (S - s[j]) / (s[i] - s[j])
    return w


def lagrange_deriv_weights(S, s):
    """L_i'(S) of lagrange_int_deriv.m:1-20 (sum over j != i of
    1/(s_i - s_j) prod_{k != i, j} (S - s_k)/(s_i - s_k))."""
    s = np.asarray(s, dtype=float)
    n = s.size
    w = np.zeros(n)
    for i in range(n):
        for j in range(n):
            if j == i:
                continue
            ld = 1.0
            for k in range(n):
                if k != i and k != j:
                    ld *= (S - s[k]) / (s[i] - s[k])
            w[i] += ld / (s[i] - s[j])
    return w


def lagrange_int(F, S, s):
    """sum_i L_i(S) F_i for a list of arrays F (lagrange_int.m:1-13)."""
    return sum(wi * Fi for wi, Fi in zip(lagrange_weights(S, s), F))


def mor_basis(A, E, b, order):
    """Moment basis of Bragg_MOR_script.m:97-143 with exact solves.

    A[i], E[i]: dense system and derivative operators at sample i (E is
    MVP_E: the operator of the differentiated tensor minus Gram I); b[i]:
    right-hand sides.  Per sample: x = A^-1 b, V = [x/|x|], then for
    n = 1..order-1: v = A^-1 (-+ E V_n) (minus for odd n), classical
    Gram-Schmidt against V, normalised.  Returns the orthonormal left
    singular vectors of [V_1 ... V_nG] (svd(P, 0))."""
    cols = []
    for Ai, Ei, bi in zip(A, E, b):
        x = np.linalg.solve(Ai, bi)
        V = [x / np.linalg.norm(x)]
        for nn in range(1, order):
            v = -(Ei @ V[-1]) if nn % 2 == 1 else Ei @ V[-1]
            v = np.linalg.solve(Ai, v)
            for q in V:
                v = v - (q.conj() @ v) * q
            V.append(v / np.linalg.norm(v))
        cols.extend(V)
    U, _, _ = np.linalg.svd(np.column_stack(cols), full_matrices=False)
    return U


def mor_project(A, U):
    """Reduced matrices Q_i = U' A_i U (Bragg_MOR_script.m:135-143)."""
    return [U.conj().T @ (Ai @ U) for Ai in A]


def mor_solve(Q, s, S, U, b):
    """x = U (lagrange_int(Q, S, s) \\ U' b) (example:97-100)."""
    return U @ np.linalg.solve(lagrange_int(Q, S, s), U.conj().T @ b)


def spectrum_point(Q, s, U, S, fN, Mc, idx, Gram, origin, dx, Eo, src, cutoff, x_thru, m_edge, n_ext=1.0,
                   lam_scale=1.0):
    """One wavelength of the reduced sweep (example_Bragg_1500nm_to_1600nm.m:85-114).

    S: wavelength in the units of s (lam_scale metres each); fN: the sample
    tensors (Lagrange-interpolated to S for the total field); Mc: L x M x N
    contrast; idx: 1-based unknowns; the incident field is the point current
    Eo at the complex point src observed at the voxel centres, each
    component divided by its largest modulus (max(max(max(abs(.)))) leaves
    one maximum per component).  Returns (power through x = x_thru,
    reduced solution x, total field)."""
    from .fields import E_field_DGF_mollified, E_field_Nop_const, grid_coords, power_through
    from .vie import C0, em_constants, jvie2_rhs
    L, M, N = Mc.shape
    freq = C0 * n_ext / (S * lam_scale)
    _, ko, _ = em_constants(freq)
    E = E_field_DGF_mollified(np.asarray(Eo, complex), np.atleast_2d(src), grid_coords(origin, dx, L, M, N), ko,
                              cutoff).reshape(L, M, N, 3, order='F')
    Einc = E / np.abs(E).max(axis=(0, 1, 2))
    b = jvie2_rhs(Einc, Mc, Gram, freq, idx)
    x = mor_solve(Q, s, S, U, b)
    J = np.zeros(3 * L * M * N, dtype=complex)
    J[np.asarray(idx, dtype=np.int64) - 1] = x
    Eout = E_field_Nop_const(J.reshape(L, M, N, 3, order='F'), lagrange_int(fN, S, s), Gram, freq, Einc)
    return power_through(Eout, x_thru, m_edge), x, Eout
