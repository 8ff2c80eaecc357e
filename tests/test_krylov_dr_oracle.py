"""CPU tests pinning the GMRES-with-deflated-restarts oracle: its first cycle
is GMRES, the recycled pair keeps C = A U with orthonormal C, harmonic Ritz
vectors of an invariant Krylov space are eigenvectors, the recycled space
converges to the eigenvectors of the smallest eigenvalues, and deflation beats
plain restarts on a matrix with small outlying eigenvalues."""
import numpy as np
import pytest

from oracle import krylov as okry


def _outliers(n, rng, nsmall=5):
    # eigenvalues: nsmall near the origin, the rest in an annulus; mildly non-normal
    lam = np.concatenate([np.linspace(0.01, 0.05, nsmall) * np.exp(1j * rng.uniform(0, 2 * np.pi, nsmall)),
                          (1 + rng.uniform(0, 1, n - nsmall)) * np.exp(1j * rng.uniform(-0.5, 0.5, n - nsmall))])
    S = np.eye(n) + 0.1 * (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / np.sqrt(n)
    A = S @ np.diag(lam) @ np.linalg.inv(S)
    return A, S, lam


def test_first_cycle_is_gmres():
    rng = np.random.default_rng(1)
    n = 80
    A = np.eye(n) * 2 + (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / np.sqrt(n)
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    f = lambda v: A @ v      # noqa: E731
    x1, _, _, it1, rv1 = okry.pgmresDR(f, b, 20, 1e-14, 0)
    x2, _, _, it2, rv2 = okry.pgmres(f, b, 20, 1e-14, 1)
    assert len(rv1) == len(rv2) == 21 and np.allclose(rv1, rv2, rtol=1e-10, atol=0)
    assert np.linalg.norm(x1 - x2) < 1e-12 * np.linalg.norm(x2)


@pytest.mark.parametrize('side', ['none', 'right', 'left'])
def test_recycled_pair_and_convergence(side):
    rng = np.random.default_rng(2)
    n = 120
    A, S, lam = _outliers(n, rng)
    Mi = np.diag(1 + 0.2 * rng.uniform(size=n)).astype(complex)          # a mild preconditioner
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    f = lambda v: A @ v      # noqa: E731
    m = lambda v: Mi @ v     # noqa: E731
    kw = dict(R1=m) if side == 'right' else dict(L1=m) if side == 'left' else {}
    Aeff = A @ Mi if side == 'right' else Mi @ A if side == 'left' else A
    trace = []
    x, flag, relres, it, rv = okry.pgmresDR(f, b, 15, 1e-10, 60, 5, trace=trace, **kw)
    assert flag == 0 and relres <= 1e-10
    assert np.linalg.norm(x - np.linalg.solve(A, b)) < 1e-8 * np.linalg.norm(x)
    assert len(trace) >= 2
    for U, C in trace:
        assert np.abs(C.conj().T @ C - np.eye(C.shape[1])).max() < 1e-10
        assert np.linalg.norm(Aeff @ U - C) < 1e-9 * np.linalg.norm(C)
    # plain restarts need more iterations on the same problem
    _, flag2, _, _, rv2 = okry.pgmres(f, b, 15, 1e-10, 60, **kw)
    print(side, 'deflated', len(rv), 'plain', len(rv2), flag2)
    assert len(rv) < len(rv2)


def test_recycled_space_finds_small_eigenvectors():
    rng = np.random.default_rng(3)
    n = 120
    A, S, lam = _outliers(n, rng)
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    trace = []
    okry.pgmresDR(lambda v: A @ v, b, 15, 1e-12, 60, 5, trace=trace)
    E = S[:, np.argsort(np.abs(lam))[:5]]              # right eigenvectors of the small eigenvalues
    Qe, _ = np.linalg.qr(E)
    U, _ = np.linalg.qr(trace[-1][0])
    gap = np.linalg.norm(U - Qe @ (Qe.conj().T @ U), 2)   # sine of the largest principal angle
    print('principal angle sine', gap, 'cycles', len(trace))
    assert gap < 1e-3


def test_harmonic_ritz_of_invariant_space_are_eigenvectors():
    rng = np.random.default_rng(4)
    n = 30
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    lam = np.linspace(0.1, 3, n) * np.exp(1j * rng.uniform(0, 2 * np.pi, n))
    A = Q @ np.diag(lam) @ Q.conj().T
    r = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    V, H, _, k, _ = okry.pgmresDR_iter(lambda v: A @ v, r, n, None, None, None, None, None, 0.0)
    # n steps span the whole space: H(n+1, n) vanishes and the harmonic Ritz
    # problem is the eigenproblem of A in the Krylov basis
    assert k == n and abs(H[n, n - 1]) < 1e-12 * np.abs(H).max()
    P = okry.harm_vecs1(k, 4, H)
    Uq, _ = np.linalg.qr(V[:, :k] @ P)
    E = Q[:, np.argsort(np.abs(lam))[:4]]
    gap = np.linalg.norm(Uq - E @ (E.conj().T @ Uq), 2)
    assert gap < 1e-8
