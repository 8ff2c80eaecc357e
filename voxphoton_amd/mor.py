"""Wavelength sweep by model order reduction on the GPU.

Mirrors Bragg_MOR_script.m:1-143 (moment basis at nG sample wavelengths,
orthonormalised, operators projected) and the reduced solve of
example_Bragg_1500nm_to_1600nm.m:85-100 (Lagrange interpolation of the
reduced matrices over the wavelength, a small dense solve, lift).  Solves,
operator applies, tensor interpolation and the tall vector products run in
libvoxphoton.so on device vectors; only the nG*order-sized reduced systems
are formed on the host, as the reference does in MATLAB.
"""
import ctypes

import numpy as np

from . import _lib as L
from ._lib import check, ptr
from .vie import DeviceArray, NOperator, _info


def lagrange_weights(S, s):
    """L_i(S) of lagrange_int.m:1-13."""
    s = np.asarray(s, dtype=float)
    w = np.ones(s.size)
    for i in range(s.size):
        for j in range(s.size):
            if j != i:
                w[i] *= (S - s[j]) / (s[i] - s[j])
    return w


def lagrange_deriv_weights(S, s):
    """L_i'(S) of lagrange_int_deriv.m:1-20."""
    s = np.asarray(s, dtype=float)
    w = np.zeros(s.size)
    for i in range(s.size):
        for j in range(s.size):
            if j == i:
                continue
            ld = 1.0
            for k in range(s.size):
                if k != i and k != j:
                    ld *= (S - s[k]) / (s[i] - s[k])
            w[i] += ld / (s[i] - s[j])
    return w


def op_lincomb(ops, w):
    """Operator whose tensor is sum_i w[i] fN_i (vp_op_create_lincomb)."""
    k = len(ops)
    hs = (ctypes.c_void_p * k)(*[o.handle.value for o in ops])
    wv = np.ascontiguousarray(np.asarray(w, dtype=np.complex128))
    h = ctypes.c_void_p()
    check(L.lib.vp_op_create_lincomb(hs, ptr(wv), k, ctypes.byref(h)))
    return NOperator(h, ops[0].ctx, ops[0].n, (ops[0].ctx,))


def vec_dots(ctx, X, Y, n):
    """X^H Y (len(X) x len(Y)) of device complex128 vectors (vp_vec_dots)."""
    k, m = len(X), len(Y)
    xs = (ctypes.c_void_p * k)(*[a.ptr.value for a in X])
    ys = (ctypes.c_void_p * m)(*[a.ptr.value for a in Y])
    out = np.empty(k * m, dtype=np.complex128)
    check(L.lib.vp_vec_dots(ctx.handle, xs, k, ys, m, int(n), ptr(out)))
    return out.reshape(k, m, order='F')


def vec_lincomb(ctx, X, C, n):
    """New device vectors Y_j = sum_i X_i C[i, j] (vp_vec_lincomb)."""
    C = np.asarray(C, dtype=np.complex128).reshape(len(X), -1)
    k, m = C.shape
    Y = [DeviceArray(ctx, n * 16) for _ in range(m)]
    xs = (ctypes.c_void_p * k)(*[a.ptr.value for a in X])
    ys = (ctypes.c_void_p * m)(*[a.ptr.value for a in Y])
    cf = np.ascontiguousarray(C.reshape(-1, order='F'))
    check(L.lib.vp_vec_lincomb(ctx.handle, xs, k, ptr(cf), m, int(n), ys))
    return Y


class ReducedModel:
    """Reduced model of A(s) x = b(s) over sample parameters s (wavelengths).

    ops[i]: the operator at s[i]; precs[i]: its preconditioner (or None);
    rhs[i]: host right-hand side at s[i].  Per sample the basis holds the
    solution and order-1 moments A^-1 (-+ E V_n), E = A'(s_i) the operator
    of the Lagrange-differentiated tensor minus Gram I (MVP_E), each
    orthogonalised against the sample's earlier columns by one classical
    Gram-Schmidt pass; the union is orthonormalised (two passes, columns
    whose remaining norm falls below drop times their norm are dropped: the
    span of svd(P, 0) up to numerical rank) and every A_i projected.
    Solves are mixed precision on the device to ``tol`` (the reference used
    GMRES to 1e-4).
    """

    def __init__(self, ops, s, rhs, order=5, precs=None, tol=1e-6, restart=100, inner_tol=1e-3, inner_maxit=6,
                 max_outer=12, side='right', drop=1e-10, log=None):
        self.ops = list(ops)
        self.s = np.asarray(s, dtype=float)
        self.ctx = ops[0].ctx
        self.n = ops[0].n
        self.precs = list(precs) if precs is not None else [None] * len(ops)
        self.solve_kw = dict(tol=tol, restart=restart, inner_tol=inner_tol, inner_maxit=inner_maxit,
                             max_outer=max_outer, side=side)
        self.solve_stats = []
        n, ctx = self.n, self.ctx
        gram = self._gram(ops[0])
        cols = []
        for i, (op, P) in enumerate(zip(self.ops, self.precs)):
            E = op_lincomb(self.ops, lagrange_deriv_weights(self.s[i], self.s))
            b = DeviceArray(ctx, n * 16)
            b.upload(np.ascontiguousarray(np.asarray(rhs[i], dtype=np.complex128)))
            V = [self._normalise(self._solve(op, P, b))]
            for nn in range(1, order):
                y = DeviceArray(ctx, n * 16)
                E.apply_device(V[-1].ptr.value, y.ptr.value)
                sg = -1.0 if nn % 2 == 1 else 1.0     # v = -+ (A'(V_n) - Gram V_n)
                v = vec_lincomb(ctx, [y, V[-1]], [[sg], [-sg * gram]], n)[0]
                w = self._solve(op, P, v)
                h = vec_dots(ctx, V, [w], n)[:, 0]
                w = vec_lincomb(ctx, V + [w], np.concatenate([-h, [1.0]])[:, None], n)[0]
                V.append(self._normalise(w))
            cols.extend(V)
            E.close()
            if log:
                log(f'  MOR sample {i}: {order} columns, solves {self.solve_stats[-order:]}')
        self.U = self._orthonormalise(cols, drop)
        self.K = len(self.U)
        self.Q = []
        for op in self.ops:
            AU = []
            for u in self.U:
                y = DeviceArray(ctx, n * 16)
                op.apply_device(u.ptr.value, y.ptr.value)
                AU.append(y)
            self.Q.append(vec_dots(ctx, self.U, AU, n))
            del AU

    @staticmethod
    def _gram(op):
        # an operator's FP32 copy is A / Gram, so its FP32 scale is Gram
        return op.fp32_scale

    def _solve(self, op, P, b):
        x = DeviceArray(self.ctx, self.n * 16)
        kw = self.solve_kw
        info, _ = _info(kw['max_outer'] + 2)
        check(L.lib.vp_solve_ir(op.handle, P.handle if P is not None else None,
                                L.VP_PREC_RIGHT if kw['side'] == 'right' else L.VP_PREC_LEFT, b.ptr, x.ptr,
                                L.VP_MEM_DEVICE, float(kw['tol']), int(kw['restart']), float(kw['inner_tol']),
                                int(kw['inner_maxit']), int(kw['max_outer']), ctypes.byref(info)))
        self.solve_stats.append((info.flag, info.total_iters, float(info.relres)))
        return x

    def _normalise(self, v):
        nr = np.sqrt(vec_dots(self.ctx, [v], [v], self.n)[0, 0].real)
        return vec_lincomb(self.ctx, [v], [[1.0 / nr]], self.n)[0]

    def _orthonormalise(self, cols, drop):
        U = []
        for c in cols:
            nr0 = np.sqrt(vec_dots(self.ctx, [c], [c], self.n)[0, 0].real)
            v = c
            for _ in range(2):
                if not U:
                    break
                h = vec_dots(self.ctx, U, [v], self.n)[:, 0]
                v = vec_lincomb(self.ctx, U + [v], np.concatenate([-h, [1.0]])[:, None], self.n)[0]
            nr = np.sqrt(vec_dots(self.ctx, [v], [v], self.n)[0, 0].real)
            if nr > drop * nr0:
                U.append(vec_lincomb(self.ctx, [v], [[1.0 / nr]], self.n)[0])
        return U

    def reduced(self, S):
        """lagrange_int(Q, S, s): the reduced matrix at S."""
        return sum(w * Q for w, Q in zip(lagrange_weights(S, self.s), self.Q))

    def solve(self, S, b):
        """Reduced solution at S for a device or host right-hand side:
        returns (device x, reduced coefficients)."""
        if not isinstance(b, DeviceArray):
            bd = DeviceArray(self.ctx, self.n * 16)
            bd.upload(np.ascontiguousarray(np.asarray(b, dtype=np.complex128)))
            b = bd
        c = vec_dots(self.ctx, self.U, [b], self.n)[:, 0]
        y = np.linalg.solve(self.reduced(S), c)
        return vec_lincomb(self.ctx, self.U, y[:, None], self.n)[0], y
