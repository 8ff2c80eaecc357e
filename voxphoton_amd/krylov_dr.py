"""GMRES with deflated restarts (GCRO-DR) on the device.

Mirrors pc/src_numeric/src_iterative/pgmresDR.m:73-281 with
pgmresDR_iter.m:23-142, getHarmVecs1.m:14-28 and getHarmVecs2.m:16-35.  Every
tall operation runs in libvoxphoton.so on FP64 device vectors: operator and
preconditioner applies, the Gram-Schmidt products against the recycled space
C and the Arnoldi basis V (vp_vec_dots), and the updates of V, x, r, U and C
(vp_vec_lincomb).  The small dense algebra of a cycle (least squares of the
Hessenberg matrices, harmonic Ritz eigenproblems, QR of H P) runs on the host
in NumPy/SciPy, as the reference does in MATLAB.
"""
import ctypes
import os
import sys

import numpy as np

from . import _lib as L
from ._lib import check
from .mor import vec_dots, vec_lincomb
from .vie import DeviceArray


def _by_modulus(vals):
    return np.argsort(np.abs(vals), kind='stable')


def harm_vecs1(m, k, H):
    """getHarmVecs1.m:14-28: eigenvectors of H_m + h_{m+1,m}^2 H_m^-H e_m e_m'
    for the k harmonic Ritz values of smallest modulus."""
    Hm = H[:m, :]
    X = np.linalg.solve(Hm.conj().T, np.eye(m))
    X[:, :m - 1] = 0
    vals, vecs = np.linalg.eig(Hm + H[m, m - 1] ** 2 * X)
    return vecs[:, _by_modulus(vals)[:k]]


def harm_vecs2(m, k, H2, VU, CU):
    """getHarmVecs2.m:16-35 from the products VU = V'U and CU = C'U: the
    eigenvectors of H2'[CU 0; VU I] z = theta H2'H2 z for the k largest |theta|."""
    import scipy.linalg
    Am = np.zeros((m + 1, m), dtype=complex)
    Am[:k, :k] = CU
    Am[k:m + 1, :k] = VU
    Am[k:m, k:m] = np.eye(m - k)
    vals, vecs = scipy.linalg.eig(H2.conj().T @ Am, H2.conj().T @ H2)
    return vecs[:, _by_modulus(vals)[m - k:m]]


def _right_div(U, R):
    return np.linalg.solve(R.T, U.T).T


class _Work:
    """Device vectors of one solve: Arnoldi basis, two U and C buffers,
    three temporaries, two x and two r buffers (all complex128, length n)."""

    def __init__(self, ctx, n, m, kmax):
        self.n = n
        self.V = DeviceArray(ctx, (m + 1) * n * 16)
        self.U = [DeviceArray(ctx, max(kmax, 1) * n * 16) for _ in range(2)]
        self.C = [DeviceArray(ctx, max(kmax, 1) * n * 16) for _ in range(2)]
        self.t = [DeviceArray(ctx, n * 16) for _ in range(3)]
        self.x = [DeviceArray(ctx, n * 16) for _ in range(2)]
        self.r = [DeviceArray(ctx, n * 16) for _ in range(2)]

    def cols(self, buf, j0, j1):
        return [buf.ptr.value + j * self.n * 16 for j in range(j0, j1)]


def pgmresDR(A, b, restart=50, tol=1e-3, maxit=10, ritz=None, L1=None, L2=None, R1=None, R2=None, x0=None):
    """[x, flag, relres, iter, resvec] = pgmresDR(A, b, restart, tol, maxit, ritz, L1, L2, R1, R2, x0).

    A: NOperator; L1 / R1: Preconditioner (left and/or right; the second
    factors L2 / R2 are not supported); b, x0: host vectors.  The first cycle
    is GMRES(restart); later cycles recycle the span of the ``ritz``
    harmonic Ritz vectors of smallest modulus (U, with C = A U orthonormal),
    run ``restart`` Arnoldi steps orthogonal to C and minimise over [U V].
    ``maxit`` bounds the cycles after the first; ``ritz`` defaults to
    min(10, round(restart/2)) and is capped at restart-1.  flag = 0 when the
    true relative residual ||b - A x|| / ||b|| is <= tol.  FP64 throughout.
    """
    if L2 is not None or R2 is not None:
        raise ValueError('a second preconditioner factor is not supported')
    ctx, n = A.ctx, A.n
    b = np.ascontiguousarray(np.asarray(b, dtype=np.complex128).reshape(-1))
    if b.size != n:
        raise ValueError(f'vector length {b.size} != operator size {n}')
    ritz = min(10, int(np.floor(restart / 2 + 0.5))) if ritz is None else min(int(ritz), restart - 1)
    m = int(restart)
    w = _Work(ctx, n, m, m)
    db = DeviceArray(ctx, n * 16)
    db.upload(b)
    dx0 = None
    if x0 is not None and np.any(x0):
        dx0 = DeviceArray(ctx, n * 16)
        dx0.upload(np.ascontiguousarray(np.asarray(x0, dtype=np.complex128).reshape(-1)))

    def op(src, dst):
        A.apply_device(src, dst)

    def prec(P, src, dst):
        check(L.lib.vp_prec_apply(P.handle, ctypes.c_void_p(src), ctypes.c_void_p(dst), L.VP_MEM_DEVICE, L.VP_FP64))

    def addr(a):
        return a.ptr.value if isinstance(a, DeviceArray) else int(a)

    def nrm(v):
        return float(np.sqrt(vec_dots(ctx, [v], [v], n)[0, 0].real))

    def lincomb(X, c, out):
        vec_lincomb(ctx, X, np.asarray(c, dtype=complex).reshape(len(X), -1), n, out=out)

    # r = b - A x0 and the left-preconditioned residual (pgmresDR.m:129-151)
    ri = 0
    if dx0 is not None:
        op(addr(dx0), addr(w.t[0]))
        lincomb([db, w.t[0]], [1.0, -1.0], [w.r[ri]])
    else:
        lincomb([db], [1.0], [w.r[ri]])
    bp = db
    if L1 is not None:
        prec(L1, addr(w.r[ri]), addr(w.t[0]))
        lincomb([w.t[0]], [1.0], [w.r[ri]])
        prec(L1, addr(db), addr(w.t[1]))
        bp = w.t[1]
    bnorm = nrm(bp)
    resvec = [nrm(w.r[ri]) / bnorm]

    def cycle(r, Ccols):
        """pgmresDR_iter.m:23-142 with V in w.V; returns (H, B, k, resvec)."""
        kc = len(Ccols)
        beta = nrm(r)
        V = lambda j0, j1: w.cols(w.V, j0, j1)     # noqa: E731
        lincomb([r], [1.0 / beta], V(0, 1))
        H = np.zeros((m + 1, m), dtype=complex)
        B = np.zeros((kc, m), dtype=complex)
        rv = np.zeros(m)
        t = [addr(v) for v in w.t]
        for k in range(1, m + 1):
            src = V(k - 1, k)[0]
            if R1 is not None:
                prec(R1, src, t[0])
                src = t[0]
            op(src, t[1])
            src = t[1]
            if L1 is not None:
                prec(L1, src, t[0])
                src = t[0]
            if kc:
                bk = vec_dots(ctx, Ccols, [src], n)[:, 0]
                B[:, k - 1] = bk
                dst = t[2]
                lincomb(Ccols + [src], np.concatenate([-bk, [1.0]]), [dst])
                src = dst
            h = vec_dots(ctx, V(0, k), [src], n)[:, 0]
            H[:k, k - 1] = h
            dst = next(a for a in t if a != src)
            lincomb(V(0, k) + [src], np.concatenate([-h, [1.0]]), [dst])
            hk = nrm(dst)
            H[k, k - 1] = hk
            lincomb([dst], [1.0 / hk], V(k, k + 1))
            rhs = np.zeros(k + 1, dtype=complex)
            rhs[0] = beta
            Hk = H[:k + 1, :k]
            y = np.linalg.lstsq(Hk, rhs, rcond=None)[0]
            rv[k - 1] = np.linalg.norm(rhs - Hk @ y)
            if rv[k - 1] < tol * bnorm:
                return Hk.copy(), B[:, :k].copy(), k, rv[:k]
        return H, B, m, rv

    progress = os.environ.get('VP_PROGRESS') == '1'

    def report(outit):
        if progress:
            print(f'[vp] pgmresDR cycle {outit}: {len(resvec) - 1} iterations, estimated relres {resvec[-1]:.3e}',
                  file=sys.stderr, flush=True)

    # first cycle: GMRES(restart) (pgmresDR.m:165-203)
    H, _, p, rin = cycle(addr(w.r[ri]), [])
    resvec.extend(list(rin / bnorm))
    report(0)
    Vc = w.cols(w.V, 0, p + 1)
    y = np.linalg.lstsq(H, vec_dots(ctx, Vc, [w.r[ri]], n)[:, 0], rcond=None)[0]
    xi = 0
    lincomb(Vc[:p], y, [w.x[xi]])
    lincomb([w.r[ri]] + Vc, np.concatenate([[1.0], -(H @ y)]), [w.r[1 - ri]])
    ri = 1 - ri
    ui = ci = 0
    kc = 0
    if resvec[-1] > tol:
        kc = min(ritz, p - 1)
        P = harm_vecs1(p, kc, H)
        Q, R = np.linalg.qr(H @ P)
        lincomb(Vc, Q, w.cols(w.C[ci], 0, kc))
        lincomb(Vc[:p], _right_div(P, R), w.cols(w.U[ui], 0, kc))
    outit = 0
    while resvec[-1] > tol and outit < maxit:                   # pgmresDR.m:210-252
        outit += 1
        Cc = w.cols(w.C[ci], 0, kc)
        Uc = w.cols(w.U[ui], 0, kc)
        H, B, p, rin = cycle(addr(w.r[ri]), Cc)
        resvec.extend(list(rin / bnorm))
        report(outit)
        Vc = w.cols(w.V, 0, p + 1)
        d = np.sqrt(np.abs(np.diag(vec_dots(ctx, Uc, Uc, n)))) if kc else np.zeros(0)
        D = np.diag(1.0 / d)
        H2 = np.block([[D, B], [np.zeros((H.shape[0], kc)), H]]) if kc else H
        Wc = Cc + Vc
        y = np.linalg.lstsq(H2, vec_dots(ctx, Wc, [w.r[ri]], n)[:, 0], rcond=None)[0]
        # x += [U D, V] y and r -= [C V] H2 y (U keeps its unscaled columns)
        lincomb([w.x[xi]] + Uc + Vc[:p], np.concatenate([[1.0], d ** -1 * y[:kc], y[kc:]]), [w.x[1 - xi]])
        xi = 1 - xi
        lincomb([w.r[ri]] + Wc, np.concatenate([[1.0], -(H2 @ y)]), [w.r[1 - ri]])
        ri = 1 - ri
        if resvec[-1] > tol and kc:
            CU = vec_dots(ctx, Cc, Uc, n) @ D
            VU = vec_dots(ctx, Vc, Uc, n) @ D
            P = harm_vecs2(p + kc, kc, H2, VU, CU)
            Q, R = np.linalg.qr(H2 @ P)
            lincomb(Wc, Q, w.cols(w.C[1 - ci], 0, kc))
            G = _right_div(P, R)
            G[:kc, :] = D @ G[:kc, :]
            lincomb(Uc + Vc[:p], G, w.cols(w.U[1 - ui], 0, kc))
            ci, ui = 1 - ci, 1 - ui

    # back to the unknowns (pgmresDR.m:256-270) and the true residual
    xs = w.x[xi]
    if R1 is not None:
        prec(R1, addr(xs), addr(w.t[0]))
        xs = w.t[0]
    dx = w.x[1 - xi]
    if dx0 is not None:
        lincomb([dx0, xs], [1.0, 1.0], [dx])
    else:
        lincomb([xs], [1.0], [dx])
    op(addr(dx), addr(w.t[1]))
    lincomb([db, w.t[1]], [1.0, -1.0], [w.t[2]])
    relres = nrm(w.t[2]) / nrm(db)
    x = dx.download(np.empty(n, dtype=np.complex128))
    flag = 0 if relres <= tol else 1
    return x, flag, relres, [p, outit], np.array(resvec)
