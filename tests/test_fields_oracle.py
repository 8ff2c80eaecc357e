"""CPU tests pinning the field oracle (source field, total field, power)
against the textbook dyadic Green's function, hand-evaluated complex-source
points and the dense N operator."""
import cmath

import numpy as np

from oracle import fields as ofl, vie as ovie
from tests.helpers import random_toeplitz, rel


def _textbook(J, src, obs, ko, cutoff):
    """E = -j w mu0 G J with G = [(1 - j/kR - 1/(kR)^2) I + (-1 + 3j/kR +
    3/(kR)^2) RR/R^2] e^{-jkR} / (4 pi R) (real coordinates), times the
    mollifier exp(-cutoff / (cutoff - R^2)) inside R < cutoff."""
    omega = ko * ovie.C0
    d = obs - src
    R = np.linalg.norm(d)
    u = d / R
    kr = ko * R
    G = ((1 - 1j / kr - 1 / kr ** 2) * np.eye(3) + (-1 + 3j / kr + 3 / kr ** 2) * np.outer(u, u)) \
        * np.exp(-1j * kr) / (4 * np.pi * R)
    psi = np.exp(-cutoff / (cutoff - R * R)) if R < cutoff else 0.0
    return -1j * omega * ovie.MU0 * (G @ J) * psi


def test_dgf_matches_textbook_dyadic():
    rng = np.random.default_rng(3)
    ko = 2 * np.pi / 1.55e-6
    cutoff = 0.5 * 1.55e-6
    src = rng.standard_normal(3) * 1e-7
    obs = src + rng.standard_normal((20, 3)) * 2e-7
    J = rng.standard_normal(3) + 1j * rng.standard_normal(3)
    E = ofl.E_field_DGF_mollified(J, src[None, :], obs, ko, cutoff)
    for i in range(obs.shape[0]):
        ref = _textbook(J, src, obs[i], ko, cutoff)
        assert rel(E[i], ref) < 1e-12 or (np.abs(ref).max() == 0 and np.abs(E[i]).max() == 0)
    # beyond the cutoff radius the mollifier removes the field
    far = src + np.array([[cutoff * 1.01, 0, 0]])
    assert np.all(ofl.E_field_DGF_mollified(J, src[None, :], far, ko, cutoff) == 0)


def test_dgf_complex_source_by_hand():
    # complex-point-source beam: principal square root, mask on Re(R)
    lam = 440e-9
    ko = 2 * np.pi / 1.55e-6
    cutoff = 0.5 * lam
    src = np.array([[0.25 * lam - 1j * lam, 0, 0]])
    obs = np.array([[30e-9, 20e-9, -10e-9], [0.25 * lam + 10e-9, 0, 5e-9]])
    J = np.array([0, 1, 0])
    E = ofl.E_field_DGF_mollified(J, src, obs, ko, cutoff)
    omega = ko * ovie.C0
    for i in range(2):
        d = [complex(obs[i, c]) - complex(src[0, c]) for c in range(3)]
        R2 = sum(v * v for v in d)
        R = cmath.sqrt(R2)
        psi = cmath.exp(-cutoff / (cutoff - R2)) if R.real < cutoff else 0.0
        chi = 1j * omega * ovie.MU0 * cmath.exp(-1j * ko * R) / (4 * np.pi * ko * ko * R * R2)
        p = 1j * ko * R + 1
        q = ko * ko * R2 - p
        P = chi * (q - 2 * p) / R2
        Q = chi * q
        want = [P * d[0] * d[1], P * d[1] * d[1] - Q, P * d[2] * d[1]]
        assert rel(E[i], np.array(want) * psi) < 1e-13


def test_grid_coords_column_major():
    pts = ofl.grid_coords((1.0, 2.0, 3.0), 0.5, 3, 2, 2)
    assert pts.shape == (12, 3)
    v = 1 + 3 * (1 + 2 * 1)          # voxel (l, m, n) = (1, 1, 1)
    assert np.allclose(pts[v], [1.5, 2.5, 3.5])


def test_total_field_matches_dense_operator():
    rng = np.random.default_rng(11)
    L, M, N = 4, 3, 3
    G = random_toeplitz(L, M, N, rng)
    fN = ovie.fN_from_toeplitz(G)
    Gram, freq = 0.37, 1.9e14
    J = rng.standard_normal((L, M, N, 3)) + 1j * rng.standard_normal((L, M, N, 3))
    Einc = rng.standard_normal((L, M, N, 3)) + 1j * rng.standard_normal((L, M, N, 3))
    E = ofl.E_field_Nop_const(J, fN, Gram, freq, Einc)
    omega, _, _ = ovie.em_constants(freq)
    jv = J.reshape(-1, order='F')
    want = Einc.reshape(-1, order='F') + (ovie.dense_N(G) @ jv - Gram * jv) / (Gram * 1j * omega * ovie.EPS0)
    assert rel(E.reshape(-1, order='F'), want) < 1e-12
    assert np.array_equal(ofl.E_field_Nop_const(0 * J, fN, Gram, freq, Einc), Einc)


def test_power_through_slice():
    rng = np.random.default_rng(2)
    E = rng.standard_normal((5, 6, 2, 3)) + 1j * rng.standard_normal((5, 6, 2, 3))
    want = sum(abs(E[3, m, n, c]) ** 2 for m in range(1, 5) for n in range(2) for c in range(3))
    assert abs(ofl.power_through(E, 4, 1) - want) < 1e-12 * want
