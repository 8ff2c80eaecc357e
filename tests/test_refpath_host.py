"""CPU: the host Arnoldi loop of the reference-path baseline
(voxphoton_amd/refpath.py) against numpy.linalg.solve and the oracle's
pgmres restatement (same iterates: one classical Gram-Schmidt pass, right
preconditioning, restarts)."""
import numpy as np

from oracle import krylov as okry


def _system(n, seed):
    rng = np.random.default_rng(seed)
    A = np.eye(n) * 3 + (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / np.sqrt(n)
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    Mi = np.linalg.inv(A + 0.3 * (rng.standard_normal((n, n)) / np.sqrt(n)))   # rough right preconditioner
    return A, b, Mi


def test_pgmres_host_solves_and_matches_oracle():
    from voxphoton_amd.refpath import pgmres_host
    A, b, Mi = _system(60, 1)
    x, flag, relres, it, rv = pgmres_host(lambda v: A @ v, b, restart=7, tol=1e-10, maxit=40,
                                          apply_R=lambda v: Mi @ v)
    xs = np.linalg.solve(A, b)
    assert flag == 0 and relres <= 1e-10
    assert np.linalg.norm(x - xs) / np.linalg.norm(xs) < 1e-8
    xo, flag_o, relres_o, it_o, rv_o = okry.pgmres(lambda v: A @ v, b, 7, 1e-10, 40, R1=lambda v: Mi @ v)
    assert it[0] == len(rv_o) - 1
    assert np.allclose(rv, rv_o, rtol=1e-8, atol=1e-14)


def test_pgmres_host_unpreconditioned_and_maxit_flag():
    from voxphoton_amd.refpath import pgmres_host
    A, b, _ = _system(40, 2)
    x, flag, relres, it, rv = pgmres_host(lambda v: A @ v, b, restart=40, tol=1e-12, maxit=1)
    assert flag == 0 and np.linalg.norm(A @ x - b) / np.linalg.norm(b) <= 1e-12
    x, flag, relres, it, rv = pgmres_host(lambda v: A @ v, b, restart=2, tol=1e-14, maxit=1)
    assert flag == 1 and relres > 1e-14
    x, flag, relres, it, rv = pgmres_host(lambda v: A @ v, 0 * b, restart=5, tol=1e-6, maxit=3)
    assert flag == 0 and not np.any(x)
