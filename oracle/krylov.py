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


# -- GMRES with deflated restarts (GCRO-DR) ------------------------------------

def pgmresDR_iter(A, r, m, L1, L2, R1, R2, C, tol):
    """One deflated cycle (pgmresDR_iter.m:23-142): builds
    (I - C C') A V(:, :k) = V(:, :k+1) H with B = C' A V(:, :k), stopping
    early once the least-squares residual of H alone falls below tol.
    Returns (V, H, B, k, resvec)."""
    n = r.shape[0]
    kc = 0 if C is None else C.shape[1]
    V = np.zeros((n, m + 1), dtype=complex)
    beta = np.linalg.norm(r)
    V[:, 0] = r / beta                                       # :63
    B = np.zeros((kc, m), dtype=complex)
    H = np.zeros((m + 1, m), dtype=complex)
    resvec = np.zeros(m)
    for k in range(1, m + 1):
        w = V[:, k - 1]
        if R2 is not None:                                   # :76-90 (R2 first, then R1)
            w = _apply(R2, w)
        if R1 is not None:
            w = _apply(R1, w)
        w = _apply(A, w)                                     # :94
        if L1 is not None:
            w = _apply(L1, w)
        if L2 is not None:
            w = _apply(L2, w)
        if kc:                                               # :113-117
            B[:, k - 1] = C.conj().T @ w
            w = w - C @ B[:, k - 1]
        H[:k, k - 1] = V[:, :k].conj().T @ w                 # :121-122
        w = w - V[:, :k] @ H[:k, k - 1]
        H[k, k - 1] = np.linalg.norm(w)                      # :124-125
        V[:, k] = w / H[k, k - 1]
        rhs = np.zeros(k + 1, dtype=complex)
        rhs[0] = beta                                        # :128-129
        Hk = H[:k + 1, :k]
        y = np.linalg.lstsq(Hk, rhs, rcond=None)[0]          # :132-134
        resvec[k - 1] = np.linalg.norm(rhs - Hk @ y)
        if resvec[k - 1] < tol:                              # :137-147
            return V[:, :k + 1], Hk.copy(), B[:, :k].copy(), k, resvec[:k]
    return V, H, B, m, resvec


def _by_modulus(vals):
    # sortrows([abs(vals) vecs']) orders by modulus (ties are measure-zero here)
    return np.argsort(np.abs(vals), kind='stable')


def harm_vecs1(m, k, H):
    """getHarmVecs1.m:14-28: eigenvectors of H_m + h_{m+1,m}^2 H_m^-H e_m e_m'
    (the harmonic Ritz problem of one GMRES cycle) for the k harmonic Ritz
    values of smallest modulus."""
    Hm = H[:m, :]
    X = np.linalg.solve(Hm.conj().T, np.eye(m))
    X[:, :m - 1] = 0
    T = Hm + H[m, m - 1] ** 2 * X
    vals, vecs = np.linalg.eig(T)
    return vecs[:, _by_modulus(vals)[:k]]


def harm_vecs2(m, k, H2, V, U, C):
    """getHarmVecs2.m:16-35: the generalised problem
    H2' [C'U 0; V'U I] z = theta H2' H2 z of a deflated cycle; returns the
    eigenvectors of the k largest |theta| (the k smallest harmonic Ritz
    values)."""
    import scipy.linalg
    Bm = H2.conj().T @ H2
    Am = np.zeros((m + 1, m), dtype=complex)
    Am[:k, :k] = C.conj().T @ U
    Am[k:m + 1, :k] = V.conj().T @ U
    Am[k:m, k:m] = np.eye(m - k)
    Am = H2.conj().T @ Am
    vals, vecs = scipy.linalg.eig(Am, Bm)
    return vecs[:, _by_modulus(vals)[m - k:m]]


def _right_div(U, R):
    """U / R (MATLAB right division) for a square R."""
    return np.linalg.solve(R.T, U.T).T


def pgmresDR(A, b, restart=50, tol=1e-3, maxit=10, ritz=None, L1=None, L2=None, R1=None, R2=None, x0=None,
             trace=None):
    """GMRES with deflated restarts (pgmresDR.m:73-281, GCRO-DR).

    The first cycle is plain GMRES(restart) from the preconditioned
    residual; the recycled pair (U, C = A U, C orthonormal) then comes from
    the ``ritz`` harmonic Ritz vectors of smallest modulus (getHarmVecs1) and
    is refreshed after every later cycle (getHarmVecs2); each later cycle runs
    ``restart`` Arnoldi steps orthogonal to C and minimises the residual over
    [U V].  The iterate lives in the right-preconditioned variables and is
    mapped back at the end.  maxit bounds the cycles after the first.  ``ritz``
    defaults to min(10, round(restart/2)) and is capped at restart-1.
    Returns (x, flag, relres, iter, resvec) with iter = [p, outit]; with a
    list ``trace`` every (U, C) pair formed is appended to it.

    Where the reference scales U with ritz columns (pgmresDR.m:223-228) the
    recycled space's actual width is used; the two agree unless the first
    cycle stopped with fewer than ritz+1 steps without converging.
    """
    b = np.asarray(b, dtype=complex)
    ritz = min(10, int(np.floor(restart / 2 + 0.5))) if ritz is None else min(ritz, restart - 1)
    x0 = np.zeros_like(b) if x0 is None else np.asarray(x0, dtype=complex).copy()
    r = b - _apply(A, x0)                                            # :129
    bp = b.copy()
    if L1 is not None:
        r = _apply(L1, r)
        bp = _apply(L1, bp)
    if L2 is not None:
        r = _apply(L2, r)
        bp = _apply(L2, bp)
    bnorm = np.linalg.norm(bp)                                       # :154
    resvec = [np.linalg.norm(r) / bnorm]
    outit = 0
    V, H, _, p, rin = pgmresDR_iter(A, r, restart, L1, L2, R1, R2, None, tol * bnorm)   # :168
    resvec.extend(list(rin / bnorm))
    y = np.linalg.lstsq(H, V.conj().T @ r, rcond=None)[0]            # :175-177
    x = V[:, :p] @ y
    r = r - V @ (H @ y)
    U = C = None
    if resvec[-1] > tol:                                             # :180-203
        P = harm_vecs1(p, min(ritz, p - 1), H)
        U = V[:, :p] @ P
        Q, R = np.linalg.qr(H @ P)
        C = V @ Q
        U = _right_div(U, R)
        if trace is not None:
            trace.append((U.copy(), C.copy()))
    while resvec[-1] > tol and outit < maxit:                        # :210
        outit += 1
        V, H, B, p, rin = pgmresDR_iter(A, r, restart, L1, L2, R1, R2, C, tol * bnorm)
        resvec.extend(list(rin / bnorm))
        kc = U.shape[1]
        d = np.linalg.norm(U, axis=0)                                # :223-228
        D = np.diag(1.0 / d)
        U = U @ D
        H2 = np.block([[D, B], [np.zeros((H.shape[0], kc)), H]])     # :231
        W = np.column_stack([C, V])
        y = np.linalg.lstsq(H2, W.conj().T @ r, rcond=None)[0]       # :234-236
        x = x + np.column_stack([U, V[:, :p]]) @ y
        r = r - W @ (H2 @ y)
        if resvec[-1] > tol:                                         # :239-250
            P = harm_vecs2(p + kc, kc, H2, V, U, C)
            U = np.column_stack([U, V[:, :p]]) @ P
            Q, R = np.linalg.qr(H2 @ P)
            C = W @ Q
            U = _right_div(U, R)
            if trace is not None:
                trace.append((U.copy(), C.copy()))
    if R2 is not None:                                               # :256-269
        x = _apply(R2, x)
    if R1 is not None:
        x = _apply(R1, x)
    x = x0 + x
    relres = np.linalg.norm(b - _apply(A, x)) / np.linalg.norm(b)    # :276
    flag = 0 if relres <= tol else 1
    return x, flag, relres, [p, outit], np.array(resvec)


# -- BiCGSTAB ------------------------------------------------------------------

def bicgstab(A, b, tol=1e-6, maxit=20, M=None, x0=None):
    """Preconditioned BiCGSTAB, the harness's 'B' solver
    (JVIE_Solver_const.m:469-494 calls MATLAB's bicgstab(A,b,tol,maxit,M),
    a closed builtin).  Restates the published iteration (van der Vorst,
    SIAM J. Sci. Stat. Comput. 13 (1992) 631-644) with M applied to the
    search directions, the stopping rule ||b - A x|| <= tol ||b|| confirmed
    on the true residual (the recurrence restarts from it when the
    confirmation fails), and the half-iteration residual history.

    Returns ``(x, flag, relres, iter, resvec)``: flag 0 converged, 1 not,
    4 breakdown; iter whole iterations (+0.5 at a half iteration); resvec
    the relative residual at the start and after every half iteration.
    """
    b = np.asarray(b, dtype=complex)
    Mf = (lambda v: v) if M is None else (lambda v: _apply(M, v))
    x = np.zeros_like(b) if x0 is None else np.asarray(x0, dtype=complex).copy()
    bnorm = np.linalg.norm(b)
    if bnorm == 0:
        return np.zeros_like(b), 0, 0.0, 0.0, np.array([0.0])
    r = b - _apply(A, x)
    rt = r.copy()
    resvec = [np.linalg.norm(r) / bnorm]
    flag = 0 if resvec[0] <= tol else 1
    half = 0
    rho = np.vdot(rt, r)
    rho_old = alpha = omega = 1.0
    p = v = None
    first = True
    it = 0
    while flag != 0 and it < maxit:
        it += 1
        if not first:
            rho = np.vdot(rt, r)
        if rho == 0:
            flag = 4
            break
        p = r.copy() if first else r + (rho / rho_old) * (alpha / omega) * (p - omega * v)
        first = False
        ph = Mf(p)
        v = _apply(A, ph)
        rtv = np.vdot(rt, v)
        if rtv == 0:
            flag = 4
            break
        alpha = rho / rtv
        x = x + alpha * ph
        s = r - alpha * v
        half += 1
        resvec.append(np.linalg.norm(s) / bnorm)
        if resvec[-1] <= tol:
            r = b - _apply(A, x)
            if np.linalg.norm(r) <= tol * bnorm:
                flag = 0
                break
            rho = np.vdot(rt, r)
            first = True
            continue
        sh = Mf(s)
        t = _apply(A, sh)
        tt = np.vdot(t, t)
        if tt == 0:
            flag = 4
            break
        omega = np.vdot(t, s) / tt
        if omega == 0:
            flag = 4
            break
        x = x + omega * sh
        r = s - omega * t
        half += 1
        resvec.append(np.linalg.norm(r) / bnorm)
        rho_old = rho
        rho = np.vdot(rt, r)
        if resvec[-1] <= tol:
            r = b - _apply(A, x)
            if np.linalg.norm(r) <= tol * bnorm:
                flag = 0
                break
            rho = np.vdot(rt, r)
            first = True
    relres = np.linalg.norm(b - _apply(A, x)) / bnorm
    if flag == 0 and relres > tol:
        flag = 1
    return x, flag, relres, half // 2 + 0.5 * (half % 2), np.array(resvec)
