"""CPU: the host eigensolver of the harmonic Ritz selection (csrc/eig.hpp),
compiled with g++, against numpy.linalg.eig and the eigen-equation."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from tests.conftest import ROOT


@pytest.fixture(scope='module')
def shim(tmp_path_factory):
    out = tmp_path_factory.mktemp('eig') / 'libeig_shim.so'
    src = os.path.join(ROOT, 'tests', 'native', 'eig_shim.cpp')
    subprocess.run(['g++', '-O2', '-std=c++17', '-shared', '-fPIC', '-I', os.path.join(ROOT, 'voxphoton_amd', 'csrc'),
                    src, '-o', str(out)], check=True)
    lib = ctypes.CDLL(str(out))
    lib.eig_shim.restype = ctypes.c_int
    lib.eig_shim.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    return lib


def _eig(lib, A):
    n = A.shape[0]
    a = np.asfortranarray(A.astype(np.complex128))
    w = np.zeros(n, np.complex128)
    X = np.zeros((n, n), np.complex128, order='F')
    rc = lib.eig_shim(n, a.ctypes.data, w.ctypes.data, X.ctypes.data)
    assert rc == 0
    return w, X


def _hess(rng, n, scale=1.0):
    A = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) * scale
    return np.triu(A, -1)


@pytest.mark.parametrize('n', [1, 2, 3, 7, 30, 100])
def test_eigenpairs_of_random_hessenberg(shim, n):
    rng = np.random.default_rng(n)
    A = _hess(rng, n)
    w, X = _eig(shim, A)
    ref = np.linalg.eigvals(A)
    # same spectrum (matched by sorting on a fixed key)
    key = lambda v: (np.round(v.real, 8), np.round(v.imag, 8))   # noqa: E731
    assert np.allclose(sorted(w, key=key), sorted(ref, key=key), atol=1e-9 * max(1, np.abs(ref).max()))
    # each column is a unit eigenvector of its eigenvalue
    res = np.linalg.norm(A @ X - X * w[None, :], axis=0)
    assert np.all(np.abs(np.linalg.norm(X, axis=0) - 1) < 1e-12)
    assert res.max() < 1e-9 * np.linalg.norm(A)


def test_harmonic_ritz_matrix_of_an_arnoldi_cycle(shim):
    # the matrix GMRES-DR diagonalises: H_m + h^2 H_m^-H e_m e_m^H from an
    # Arnoldi run on a matrix with a few small outlying eigenvalues; the
    # harmonic Ritz values of smallest modulus approximate them
    rng = np.random.default_rng(5)
    n, m = 300, 40
    lam = np.concatenate([[1e-3, 2e-3 + 1e-3j, -3e-3], 1 + 0.5 * rng.random(n - 3) + 0.3j * rng.random(n - 3)])
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    A = Q @ np.diag(lam) @ Q.conj().T
    V = np.zeros((n, m + 1), complex)
    Hb = np.zeros((m + 1, m), complex)
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    V[:, 0] = b / np.linalg.norm(b)
    for j in range(m):
        wv = A @ V[:, j]
        for i in range(j + 1):
            Hb[i, j] = V[:, i].conj() @ wv
            wv -= Hb[i, j] * V[:, i]
        Hb[j + 1, j] = np.linalg.norm(wv)
        V[:, j + 1] = wv / Hb[j + 1, j]
    Hm = Hb[:m, :m]
    em = np.zeros(m); em[-1] = 1
    f = np.linalg.solve(Hm.conj().T, em)
    G = Hm + abs(Hb[m, m - 1]) ** 2 * np.outer(f, em)
    w, X = _eig(shim, G)
    ref = np.linalg.eigvals(G)
    small = np.sort(np.abs(w))[:3]
    assert np.allclose(small, np.sort(np.abs(ref))[:3], rtol=1e-8)
    assert np.allclose(np.sort(np.abs(w))[:3], np.sort(np.abs(lam[:3])), rtol=0.05)
    assert np.linalg.norm(G @ X - X * w[None, :], axis=0).max() < 1e-9 * np.linalg.norm(G)


def test_repeated_and_defective_eigenvalues(shim):
    # diagonal with repeats, a Jordan-like block, and a zero matrix: finite output
    A = np.diag([2.0, 2.0, 1.0, 1.0]).astype(complex)
    w, X = _eig(shim, A)
    assert np.allclose(sorted(w.real), [1, 1, 2, 2])
    J = np.array([[1, 1, 0], [0, 1, 1], [0, 0, 1]], complex)
    w, X = _eig(shim, J)
    assert np.allclose(w, 1) and np.isfinite(X).all()
    w, X = _eig(shim, np.zeros((4, 4), complex))
    assert np.allclose(w, 0) and np.isfinite(X).all()


def test_harmonic_pair_gives_recycled_space(shim):
    # harmonic_pair (csrc/eig.hpp) on the Arnoldi relation A V_K = V_{K+1} Hb:
    # C = V_{K+1} sc is orthonormal, U = V_K su satisfies A U = C, and span(U)
    # holds the harmonic Ritz vectors of smallest modulus
    lib = shim
    lib.harm_shim.restype = ctypes.c_int
    lib.harm_shim.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
    rng = np.random.default_rng(9)
    n, K, kd = 200, 30, 6
    lam = np.concatenate([[2e-3, -1e-3 + 2e-3j, 4e-3j], 1 + 0.4 * rng.random(n - 3) + 0.2j * rng.random(n - 3)])
    Qm, _ = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    A = Qm @ np.diag(lam) @ np.linalg.inv(Qm)
    V = np.zeros((n, K + 1), complex)
    Hb = np.zeros((K + 1, K), complex)
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    V[:, 0] = b / np.linalg.norm(b)
    for j in range(K):
        wv = A @ V[:, j]
        for _ in range(2):
            h = V[:, :j + 1].conj().T @ wv
            wv -= V[:, :j + 1] @ h
            Hb[:j + 1, j] += h
        Hb[j + 1, j] = np.linalg.norm(wv)
        V[:, j + 1] = wv / Hb[j + 1, j]
    hb = np.asfortranarray(Hb)
    sc = np.zeros((K + 1, kd), complex, order='F')
    su = np.zeros((K, kd), complex, order='F')
    kept = lib.harm_shim(K, hb.ctypes.data, kd, sc.ctypes.data, su.ctypes.data)
    assert kept == kd
    C = V @ sc
    U = V[:, :K] @ su
    assert np.abs(C.conj().T @ C - np.eye(kd)).max() < 1e-10
    assert np.linalg.norm(A @ U - C) < 1e-9 * np.linalg.norm(C)
    # reference selection with numpy: eigenvectors of H_K + h^2 H_K^-H e_K e_K^H
    Hk = Hb[:K, :K]
    eK = np.zeros(K); eK[-1] = 1
    G = Hk + abs(Hb[K, K - 1]) ** 2 * np.outer(np.linalg.solve(Hk.conj().T, eK), eK)
    w, P = np.linalg.eig(G)
    Pref = P[:, np.argsort(np.abs(w))[:kd]]
    Y = V[:, :K] @ Pref
    # span(U) = span(V_K Pref): projecting Y onto span(U) leaves nothing
    Qu, _ = np.linalg.qr(U)
    assert np.linalg.norm(Y - Qu @ (Qu.conj().T @ Y)) < 1e-8 * np.linalg.norm(Y)
