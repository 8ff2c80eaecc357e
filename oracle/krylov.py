"""FP64 restatement of restarted GMRES (test infrastructure only).

Restates pc/src_numeric/src_iterative/pgmres.m:70-180 and pgmres_iter.m:22-150.
Operators and preconditioners are callables (the reference's function
handles, dispatched by iterapp.m:41-57).  ``orth='cgs'`` is the reference's
single-pass classical Gram-Schmidt (pgmres_iter.m:106-107); ``orth='cgs2'``
repeats the projection once (classical Gram-Schmidt with
reorthogonalisation), the variant the device solver runs.
"""
import numpy as np


def _apply(f, x):
    y = f(x)
    y = np.asarray(y)
    if y.ndim != 1:          # iterapp.m:65-69 requires a column vector
        raise ValueError('MATLAB:MustReturnColumn')
    return y


def pgmres_iter(A, x, r, m, L1, L2, R1, R2, tol, orth='cgs'):
    """One restart cycle (pgmres_iter.m:22-150)."""
    n = r.shape[0]
    V = np.zeros((n, m + 1), dtype=complex)
    beta = np.linalg.norm(r)
    V[:, 0] = r / beta
    H = np.zeros((m + 1, m), dtype=complex)
    resvec = np.zeros(m)
    y = np.zeros(0, dtype=complex)
    res = np.array([beta], dtype=complex)
    k = 0
    for k in range(1, m + 1):
        w = V[:, k - 1]
        if R1 is not None:
            w = _apply(R1, w)
        if R2 is not None:
            w = _apply(R2, w)
        w = _apply(A, w)
        if L1 is not None:
            w = _apply(L1, w)
        if L2 is not None:
            w = _apply(L2, w)
        h = V[:, :k].conj().T @ w                       # :106
        w = w - V[:, :k] @ h                            # :107
        if orth == 'cgs2':
            h2 = V[:, :k].conj().T @ w
            w = w - V[:, :k] @ h2
            h = h + h2
        H[:k, k - 1] = h
        H[k, k - 1] = np.linalg.norm(w)                 # :109
        V[:, k] = w / H[k, k - 1]                       # :110
        rhs = np.zeros(k + 1, dtype=complex)
        rhs[0] = beta                                   # :113-114
        Hc = H[:k + 1, :k]
        y = np.linalg.lstsq(Hc, rhs, rcond=None)[0]     # :117-118
        res = rhs - Hc @ y                              # :119
        resvec[k - 1] = np.linalg.norm(res)             # :120
        if resvec[k - 1] < tol:                         # :123-125
            break
    resvec = resvec[:k]
    xupd = V[:, :k] @ y                                 # :133
    if R1 is not None:
        xupd = _apply(R1, xupd)
    if R2 is not None:
        xupd = _apply(R2, xupd)
    x = x + xupd                                        # :149
    r = V[:, :k + 1] @ res                              # :150
    return x, r, k, resvec


def pgmres(A, b, restart=50, tol=1e-3, maxit=10, L1=None, L2=None, R1=None,
           R2=None, x0=None, orth='cgs'):
    """Restarted GMRES(m) (pgmres.m:70-180).

    Returns ``(x, flag, relres, iter, resvec)`` with ``iter = [p, outit]``;
    ``flag`` is 0 when the final unpreconditioned relative residual
    ``||b - A x|| / ||b||`` is <= tol, else 1 (non-convergence is not an error).
    """
    b = np.asarray(b, dtype=complex)
    x = np.zeros_like(b) if x0 is None else np.asarray(x0, dtype=complex).copy()
    r = b - _apply(A, x) if np.count_nonzero(x) else b.copy()     # :115-119
    bp = b.copy()
    if L1 is not None:
        r = _apply(L1, r)
        bp = _apply(L1, bp)
    if L2 is not None:
        r = _apply(L2, r)
        bp = _apply(L2, bp)
    bnorm = np.linalg.norm(bp)                                     # :142
    resvec = [np.linalg.norm(r) / bnorm]
    outit = 0
    p = 0
    while resvec[-1] > tol and outit < maxit:                      # :153
        outit += 1
        x, r, p, rin = pgmres_iter(A, x, r, restart, L1, L2, R1, R2,
                                   tol * bnorm, orth)
        resvec.extend(list(rin / bnorm))
    relres = np.linalg.norm(b - _apply(A, x)) / np.linalg.norm(b)  # :175
    flag = 0 if relres <= tol else 1
    return x, flag, relres, [p, outit], np.array(resvec)
