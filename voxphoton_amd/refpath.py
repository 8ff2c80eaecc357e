"""The reference's own FP64 GPU solve path, run as a measured baseline.

With ``GPU_flag = 1`` the reference's driver (JVIE_Solver_const.m:253-295,
:333) keeps fN, Mr and Mc on the device but runs ``pgmres``
(pgmres.m:70-180, pgmres_iter.m:60-150) on host arrays.  Every operator
application uploads the current vector, builds zero arrays and runs
mv_AN_const on the device (mv_AN_const.m:46-58: zero-padded ``fftn`` at
2L x 2M x 2N, the tensor multiply, ``ifftn``, crop, contrast) and downloads
the result (:139-143).  Arnoldi, with one classical Gram-Schmidt pass per step
(pgmres_iter.m:106-107), and the least squares ``Hc \\ rhs`` re-solved every
step (:113-120) run on the host.

``pgmres_host`` keeps that data flow.  Operator and preconditioner apply
through this library's FP64 device kernels on host vectors (VP_MEM_HOST:
upload, apply, download).  The Arnoldi loop runs in numpy on the host cores.
``reference_operator`` builds the operator at the reference's FFT sizes on the
full grid.  It is the baseline of the "FP64 GPU plain path run serially"
comparison.  Nothing here is on the product's solve path.
"""
import os
import time

import numpy as np


def pgmres_host(apply_A, b, restart=50, tol=1e-3, maxit=10, apply_R=None):
    """[x, flag, relres, iter, resvec] of pgmres.m with a right preconditioner.

    ``apply_A`` and ``apply_R`` map host complex128 vectors to host vectors
    (each call crosses to the device and back, as mv_AN_const with GPU_flag =
    1 does).  Stops when the recurrence residual drops below tol * ||b||
    (pgmres.m:142,159) and reports the true relative residual
    ||b - A x|| / ||b|| (pgmres.m:175-180).
    """
    b = np.asarray(b, dtype=np.complex128).reshape(-1)
    n = b.size
    n2b = np.linalg.norm(b)
    x = np.zeros(n, np.complex128)
    if n2b == 0:
        return x, 0, 0.0, [0, 0], np.zeros(1)
    tolb = tol * n2b
    r = b.copy()
    resvec = [np.linalg.norm(r)]
    total = 0
    outer = 0
    flag = 1
    V = np.empty((n, restart + 1), np.complex128)
    for outer in range(1, maxit + 1):
        beta = np.linalg.norm(r)
        if beta <= tolb:
            flag = 0
            break
        V[:, 0] = r / beta
        H = np.zeros((restart + 1, restart), np.complex128)
        k_done = 0
        y = None
        res = None
        for k in range(restart):
            w = V[:, k]
            if apply_R is not None:
                w = apply_R(w)
            w = apply_A(w)
            h = V[:, :k + 1].conj().T @ w            # one classical Gram-Schmidt pass
            w = w - V[:, :k + 1] @ h
            H[:k + 1, k] = h
            H[k + 1, k] = np.linalg.norm(w)
            if H[k + 1, k] != 0:
                V[:, k + 1] = w / H[k + 1, k]
            else:
                V[:, k + 1] = 0
            rhs = np.zeros(k + 2, np.complex128)
            rhs[0] = beta
            Hc = H[:k + 2, :k + 1]
            y = np.linalg.lstsq(Hc, rhs, rcond=None)[0]   # Hc \ rhs, as MATLAB re-solves it
            res = rhs - Hc @ y
            resvec.append(np.linalg.norm(res))
            total += 1
            k_done = k + 1
            if resvec[-1] < tolb or H[k + 1, k] == 0:
                break
        xupd = V[:, :k_done] @ y
        if apply_R is not None:
            xupd = apply_R(xupd)
        x = x + xupd
        r = V[:, :k_done + 1] @ res                  # residual by recurrence (pgmres_iter.m:150)
        if resvec[-1] < tolb:
            flag = 0
            break
    relres = np.linalg.norm(b - apply_A(x)) / n2b
    flag = 0 if relres <= tol else 1
    return x, flag, relres, [total, outer], np.asarray(resvec) / n2b


def reference_operator(vie, prob, ctx=None):
    """FP64 operator at the reference's circulant sizes 2L x 2M x 2N on the
    full L x M x N grid (circulant_nop_const.m:5, mv_AN_const.m:102-132)."""
    saved = {k: os.environ.get(k) for k in ('VP_FFT_SIZES', 'VP_MATVEC_PRUNE')}
    os.environ['VP_FFT_SIZES'] = 'reference'
    os.environ['VP_MATVEC_PRUNE'] = '0'
    try:
        return vie.NOperator.assembled(prob.shape, prob.dx, prob.ko, None, prob.Mc, prob.idx,
                                       precisions=('fp64',), ctx=ctx)
    finally:
        for k, v in saved.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


def solve_reference_path(vie, prob, P, b, restart, tol, maxit, ctx=None):
    """Times one solve on the reference's FP64 GPU path: the operator at
    2L x 2M x 2N, host vectors through the device for every operator and
    preconditioner application, Arnoldi on the host.  Returns (x, info)."""
    op = reference_operator(vie, prob, ctx=ctx)
    t0 = time.perf_counter()
    x, flag, relres, it, rv = pgmres_host(lambda v: op.apply(v), b, restart, tol, maxit,
                                          apply_R=(lambda v: P.apply(v)) if P is not None else None)
    dt = time.perf_counter() - t0
    info = dict(seconds=dt, flag=int(flag), relres=float(relres), iterations=int(it[0]), cycles=int(it[1]),
                fft=tuple(op.info()['fft_fp64']))
    op.close()
    return x, info
