"""CPU tests pinning the model-order-reduction oracle: Lagrange weights
against polynomials, and the reduced model against exact solves of a
polynomial operator family (exact at the samples, derivative matching)."""
import numpy as np

from oracle import mor as omor


def test_lagrange_weights_reproduce_polynomials():
    s = np.array([1.5, 1.52, 1.55, 1.57, 1.6])
    for S in (1.5, 1.513, 1.59, 1.6):
        w = omor.lagrange_weights(S, s)
        wd = omor.lagrange_deriv_weights(S, s)
        for p in range(5):
            assert abs(w @ s ** p - S ** p) < 1e-10 * max(1, S ** p)
            want = p * S ** (p - 1) if p else 0.0
            assert abs(wd @ s ** p - want) < 1e-7 * max(1, abs(want))
    # at a sample the weights pick that sample
    assert np.allclose(omor.lagrange_weights(s[2], s), np.eye(5)[2])


def _family(n, rng):
    A0 = np.eye(n) * 3 + 0.3 * (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / np.sqrt(n)
    A1 = 0.2 * (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / np.sqrt(n)
    A2 = 0.1 * (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / np.sqrt(n)
    A = lambda t: A0 + t * A1 + t * t * A2                      # noqa: E731
    dA = lambda t: A1 + 2 * t * A2                              # noqa: E731
    return A, dA


def test_reduced_model_exact_at_samples_and_matches_derivatives():
    rng = np.random.default_rng(3)
    n = 60
    A, dA = _family(n, rng)
    s = np.array([-1.0, -0.5, 0.0, 0.5, 1.0])
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    As = [A(t) for t in s]
    # the derivative operator from interpolated samples (lagrange_int_deriv)
    Es = [sum(w * Ai for w, Ai in zip(omor.lagrange_deriv_weights(t, s), As)) for t in s]
    for t, Ei in zip(s, Es):
        assert np.abs(Ei - dA(t)).max() < 1e-9
    U = omor.mor_basis(As, Es, [b] * 5, order=3)
    assert U.shape == (n, 15) and np.allclose(U.conj().T @ U, np.eye(15), atol=1e-10)
    Q = omor.mor_project(As, U)
    for t in s:
        x = omor.mor_solve(Q, s, t, U, b)
        assert np.linalg.norm(x - np.linalg.solve(A(t), b)) < 1e-9 * np.linalg.norm(x)
    # first moment: d x / d t at a sample from the reduced model
    h = 1e-5
    xd = (omor.mor_solve(Q, s, 0.5 + h, U, b) - omor.mor_solve(Q, s, 0.5 - h, U, b)) / (2 * h)
    x0 = np.linalg.solve(A(0.5), b)
    want = -np.linalg.solve(A(0.5), dA(0.5) @ x0)
    assert np.linalg.norm(xd - want) < 1e-6 * np.linalg.norm(want)
    # between samples the reduced solution stays close to the full one
    for t in (-0.8, 0.27, 0.9):
        x = omor.mor_solve(Q, s, t, U, b)
        r = np.linalg.norm(A(t) @ x - b) / np.linalg.norm(b)
        assert r < 1e-4, (t, r)
