"""FP64 restatement of the steps either side of the solve (test
infrastructure only): the mollified dyadic Green's-function source field
that excites the configs, the total field through the L operator, and the
transmitted power through a cross-section.

Arrays follow the conventions of ``oracle.vie`` (MATLAB column-major).
"""
import numpy as np

from .vie import C0, EPS0, MU0, em_constants, mv_AN_const


def E_field_DGF_mollified(J, Icoord, Ocoord, ko, cutoff):
    """Field of point currents J at Icoord, observed at Ocoord.

    Restates pc/src_vie/src_fields/src_dgf/E_field_DGF_mollified.m:36-111:
    with R = Ocoord - Icoord (complex source coordinates allowed: the
    complex-point-source beam of example_Bragg_1500nm_to_1600nm.m:64-67),
    R2 = R.R, R = sqrt(R2) (principal branch), chi = j w mu0 e^{-jkR} /
    (4 pi k^2 R^3), Q = chi (k^2 R2 - (jkR + 1)), P = chi (k^2 R2 - 3 (jkR +
    1)) / R2, the field is E_i = sum_j (P X_i X_j - Q delta_ij) J_j, times
    the mollifier psi = exp(-cutoff / (cutoff - R2)) on Re(R) < cutoff (MATLAB
    compares complex numbers by their real parts) and 0 elsewhere.

    J: 3*Ni currents, column-major (Ni x 3); Icoord: Ni x 3; Ocoord: No x 3.
    Returns E as No x 3.  The reference applies psi to the summed field and is
    defined for one source (Ni = 1); here psi weighs each source's term, which
    is the same for Ni = 1.
    """
    Icoord = np.atleast_2d(np.asarray(Icoord, dtype=complex))
    Ocoord = np.atleast_2d(np.asarray(Ocoord, dtype=float))
    Ni = Icoord.shape[0]
    J = np.asarray(J, dtype=complex).reshape(-1).reshape(Ni, 3, order='F')
    omega = 2 * np.pi * (C0 / (2 * np.pi / ko))
    X = Ocoord[:, 0:1] - Icoord[None, :, 0]
    Y = Ocoord[:, 1:2] - Icoord[None, :, 1]
    Z = Ocoord[:, 2:3] - Icoord[None, :, 2]
    R2 = X * X + Y * Y + Z * Z
    R = np.sqrt(R2)
    R3 = R * R2
    with np.errstate(over='ignore', invalid='ignore'):
        psi = np.where(R.real < cutoff, np.exp(-cutoff / (cutoff - R2)), 0.0)
    chi = 1j * omega * MU0 * np.exp(-1j * ko * R) / (4 * np.pi * ko * ko * R3)
    P = 1j * ko * R + 1
    Q = ko * ko * R2 - P
    P = (Q - 2 * P) / R2
    Q = chi * Q * psi
    P = chi * P * psi
    E = np.empty((Ocoord.shape[0], 3), dtype=complex)
    E[:, 0] = (P * X * X - Q) @ J[:, 0] + (P * (X * Y)) @ J[:, 1] + (P * (X * Z)) @ J[:, 2]
    E[:, 1] = (P * (X * Y)) @ J[:, 0] + (P * (Y * Y) - Q) @ J[:, 1] + (P * (Y * Z)) @ J[:, 2]
    E[:, 2] = (P * (X * Z)) @ J[:, 0] + (P * (Y * Z)) @ J[:, 1] + (P * (Z * Z) - Q) @ J[:, 2]
    return E


def grid_coords(origin, dx, L, M, N):
    """Voxel centres origin + (l, m, n) dx as an (L*M*N) x 3 array in the
    column-major voxel order (Ocoord = reshape(r, L*M*N, 3) in the configs)."""
    l, m, n = np.meshgrid(np.arange(L), np.arange(M), np.arange(N), indexing='ij')
    pts = np.stack([origin[0] + l * dx, origin[1] + m * dx, origin[2] + n * dx], axis=-1)
    return pts.reshape(-1, 3, order='F')


def E_field_Nop_const(J, fN, Gram, freq, Einc=None):
    """Total field Einc + (N(J) - Gram J) / (Gram j w eps0) on the whole grid.

    Restates pc/src_vie/src_fields/src_op/E_field_Nop_const.m:76-84 over
    mv_L_const.m:61-85, whose output N(J) - Gram J is mv_AN_const with
    Mr = Mc = 1 and idx = every entry, negated.  J, Einc: L x M x N x 3.
    """
    L, M, N, _ = J.shape
    if Einc is None:
        Einc = np.zeros_like(J)
    if not np.any(J):
        return Einc.copy()
    omega, _, _ = em_constants(freq)
    idx = np.arange(1, 3 * L * M * N + 1)
    ones = np.ones((L, M, N), complex)
    jv = J.reshape(-1, order='F')
    lj = -mv_AN_const(jv, fN, ones, ones, Gram, 'notransp', idx)
    return Einc + lj.reshape(L, M, N, 3, order='F') / (Gram * 1j * omega * EPS0)


def power_through(E, x_thru, m_edge):
    """Transmitted power sum |E|^2 over the slice x = x_thru (1-based) and
    m_edge < m <= M - m_edge (example_Bragg_1500nm_to_1600nm.m:108-113)."""
    M = E.shape[1]
    s = E[x_thru - 1, m_edge:M - m_edge, :, :]
    return float(np.sum((s * np.conj(s)).real))
