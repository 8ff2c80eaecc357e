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
