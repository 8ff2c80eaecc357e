"""Shared fixtures for the parity tests (seeded inputs on tiny grids)."""
import numpy as np

from oracle import vie as ovie


def random_toeplitz(L, M, N, rng, scale=1.0):
    """Random complex opToeplitz with the physical parity zeros (odd
    components vanish on their zero-offset planes), so that N is complex
    symmetric as the assembled tensor is."""
    G = (rng.standard_normal((L, M, N, 6)) + 1j * rng.standard_normal((L, M, N, 6))) * scale
    for k in range(6):
        for ax in range(3):
            if ovie.PARITY[k, ax] < 0:
                sl = [slice(None)] * 3
                sl[ax] = 0
                G[tuple(sl) + (k,)] = 0.0
    return G


def random_problem(L, M, N, seed, frac=0.6, with_mr=False):
    rng = np.random.default_rng(seed)
    G = random_toeplitz(L, M, N, rng)
    mask = rng.random((L, M, N)) < frac
    mask.flat[0] = True
    idx = ovie.idx_from_mask(mask)
    Mc = (rng.standard_normal((L, M, N)) + 0.3j * rng.standard_normal((L, M, N)))
    Mr = (1.0 + 0.2 * rng.standard_normal((L, M, N)) + 0.1j) if with_mr else np.ones((L, M, N), complex)
    Gram = 0.9
    x = rng.standard_normal(idx.size) + 1j * rng.standard_normal(idx.size)
    return G, Mr, Mc, Gram, idx, x


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)
