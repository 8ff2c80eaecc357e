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
import time

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


def _addr(a):
    """Device address of a DeviceArray or an int (offset pointers)."""
    return a.ptr.value if isinstance(a, DeviceArray) else int(a)


def _ops(ops):
    return (ctypes.c_void_p * len(ops))(*[o.handle.value for o in ops])


def op_lincomb(ops, w):
    """Operator whose tensor is sum_i w[i] fN_i (vp_op_create_lincomb)."""
    wv = np.ascontiguousarray(np.asarray(w, dtype=np.complex128))
    h = ctypes.c_void_p()
    check(L.lib.vp_op_create_lincomb(_ops(ops), ptr(wv), len(ops), ctypes.byref(h)))
    return NOperator(h, ops[0].ctx, ops[0].n, (ops[0].ctx,))


def op_lincomb_into(out, ops, w):
    """Overwrite out's tensor with sum_i w[i] fN_i (vp_op_lincomb_into)."""
    wv = np.ascontiguousarray(np.asarray(w, dtype=np.complex128))
    check(L.lib.vp_op_lincomb_into(out.handle, _ops(ops), ptr(wv), len(ops)))


def vec_dots(ctx, X, Y, n):
    """X^H Y (len(X) x len(Y)) of device complex128 vectors (vp_vec_dots)."""
    k, m = len(X), len(Y)
    xs = (ctypes.c_void_p * k)(*[_addr(a) for a in X])
    ys = (ctypes.c_void_p * m)(*[_addr(a) for a in Y])
    out = np.empty(k * m, dtype=np.complex128)
    check(L.lib.vp_vec_dots(ctx.handle, xs, k, ys, m, int(n), ptr(out)))
    return out.reshape(k, m, order='F')


def vec_lincomb(ctx, X, C, n, out=None):
    """Device vectors Y_j = sum_i X_i C[i, j] (vp_vec_lincomb), written to
    ``out`` (a list of device vectors) or to new DeviceArrays."""
    C = np.asarray(C, dtype=np.complex128).reshape(len(X), -1)
    k, m = C.shape
    Y = out if out is not None else [DeviceArray(ctx, n * 16) for _ in range(m)]
    xs = (ctypes.c_void_p * k)(*[_addr(a) for a in X])
    ys = (ctypes.c_void_p * m)(*[_addr(a) for a in Y])
    cf = np.ascontiguousarray(C.reshape(-1, order='F'))
    check(L.lib.vp_vec_lincomb(ctx.handle, xs, k, ptr(cf), m, int(n), ys))
    return Y


def vec_absmax(ctx, X, n):
    """max |X_i| of each device vector (vp_vec_absmax)."""
    xs = (ctypes.c_void_p * len(X))(*[_addr(a) for a in X])
    out = np.empty(len(X))
    check(L.lib.vp_vec_absmax(ctx.handle, xs, len(X), int(n), ptr(out)))
    return out


def vec_gather(ctx, X, nx, idx, n, Y, w=None, alpha=1.0):
    """Y[i] = alpha w[i] X[idx[i]] (idx: 0-based device int64; vp_vec_gather)."""
    a = complex(alpha)
    check(L.lib.vp_vec_gather(ctx.handle, ctypes.c_void_p(_addr(X)), int(nx), ctypes.c_void_p(_addr(idx)),
                              None if w is None else ctypes.c_void_p(_addr(w)), a.real, a.imag, int(n),
                              ctypes.c_void_p(_addr(Y))))


def vec_scatter(ctx, X, idx, n, Y, ny):
    """Y = 0; Y[idx[i]] = X[i] (vp_vec_scatter)."""
    check(L.lib.vp_vec_scatter(ctx.handle, ctypes.c_void_p(_addr(X)), ctypes.c_void_p(_addr(idx)), int(n),
                               ctypes.c_void_p(_addr(Y)), int(ny)))


class ReducedModel:
    """Reduced model of A(s) x = b(s) over sample parameters s (wavelengths).

    ops[i]: the operator at s[i] (no Mr, as MVP_E assumes); precs[i]: its
    preconditioner (or None); rhs[i]: right-hand side at s[i] (host array
    or DeviceArray).  Per sample the basis holds the
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
        self.times = dict(solves=0.0, moments=0.0, orthonormalise=0.0, project=0.0)
        n, ctx = self.n, self.ctx
        gram = self._gram(ops[0])
        cols = []
        for i, (op, P) in enumerate(zip(self.ops, self.precs)):
            t0 = time.perf_counter()
            E = op_lincomb(self.ops, lagrange_deriv_weights(self.s[i], self.s))
            if isinstance(rhs[i], DeviceArray):
                b = rhs[i]
            else:
                b = DeviceArray(ctx, n * 16)
                b.upload(np.ascontiguousarray(np.asarray(rhs[i], dtype=np.complex128)))
            self.times['moments'] += time.perf_counter() - t0
            V = [self._normalise(self._solve(op, P, b))]
            for nn in range(1, order):
                t0 = time.perf_counter()
                y = DeviceArray(ctx, n * 16)
                E.apply_device(V[-1].ptr.value, y.ptr.value)
                sg = -1.0 if nn % 2 == 1 else 1.0     # v = -+ (A'(V_n) - Gram V_n)
                v = vec_lincomb(ctx, [y, V[-1]], [[sg], [-sg * gram]], n)[0]
                self.times['moments'] += time.perf_counter() - t0
                w = self._solve(op, P, v)
                t0 = time.perf_counter()
                h = vec_dots(ctx, V, [w], n)[:, 0]
                w = vec_lincomb(ctx, V + [w], np.concatenate([-h, [1.0]])[:, None], n)[0]
                V.append(self._normalise(w))
                self.times['moments'] += time.perf_counter() - t0
            cols.extend(V)
            E.close()
            if log:
                log(f'  MOR sample {i}: {order} columns, solves (flag, inner its, relres) '
                    f'{self.solve_stats[-order:]}')
        t0 = time.perf_counter()
        self.U = self._orthonormalise(cols, drop)
        self.K = len(self.U)
        del cols
        self.times['orthonormalise'] = time.perf_counter() - t0
        t0 = time.perf_counter()
        self.Q = []
        y = DeviceArray(ctx, n * 16)
        for op in self.ops:
            Qi = np.empty((self.K, self.K), dtype=np.complex128)
            for j, u in enumerate(self.U):
                op.apply_device(u.ptr.value, y.ptr.value)
                Qi[:, j] = vec_dots(ctx, self.U, [y], n)[:, 0]
            self.Q.append(Qi)
        self.times['project'] = time.perf_counter() - t0

    @staticmethod
    def _gram(op):
        # an operator's FP32 copy is A / Gram, so its FP32 scale is Gram
        return op.fp32_scale

    def _solve(self, op, P, b):
        t0 = time.perf_counter()
        x = DeviceArray(self.ctx, self.n * 16)
        kw = self.solve_kw
        info, _ = _info(kw['max_outer'] + 2)
        check(L.lib.vp_solve_ir(op.handle, P.handle if P is not None else None,
                                L.VP_PREC_RIGHT if kw['side'] == 'right' else L.VP_PREC_LEFT, b.ptr, x.ptr,
                                L.VP_MEM_DEVICE, float(kw['tol']), int(kw['restart']), float(kw['inner_tol']),
                                int(kw['inner_maxit']), int(kw['max_outer']), ctypes.byref(info)))
        self.solve_stats.append((info.flag, info.total_iters, float(info.relres)))
        self.times['solves'] += time.perf_counter() - t0
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

    def solve(self, S, b, out=None):
        """Reduced solution at S for a device or host right-hand side:
        returns (device x, reduced coefficients); x goes to ``out`` if given."""
        if not isinstance(b, DeviceArray):
            bd = DeviceArray(self.ctx, self.n * 16)
            bd.upload(np.ascontiguousarray(np.asarray(b, dtype=np.complex128)))
            b = bd
        c = vec_dots(self.ctx, self.U, [b], self.n)[:, 0]
        y = np.linalg.solve(self.reduced(S), c)
        return vec_lincomb(self.ctx, self.U, y[:, None], self.n, out=None if out is None else [out])[0], y


class SpectrumSweep:
    """Per-wavelength evaluation of a reduced model on the device
    (example_Bragg_1500nm_to_1600nm.m:85-114).

    At wavelength lam (the units of s): the field of the point current Eo
    at the complex source point src observed at the voxel centres
    origin + (l, m, n) dx, each component divided by its largest modulus
    (max(max(max(abs(Einc)))) is per component); Vrhs = Gram tau3
    Einc(idxS3) with tau = j w eps0 Mc; the reduced solve and lift; J
    scattered onto the grid; the total field through the field operator
    interpolated to lam (fN_temp = lagrange_int(fN, S, s)); and the power
    through the cross-section x = x_thru (1-based), m_edge <= m < M - m_edge.

    s: sample wavelengths (those of the reduced models evaluated); fops[i]:
    field operators (NOperator.for_fields) at s[i]; Mc: L x M x N contrast;
    idx: 1-based unknown indices of the reduced models' operators;
    lam_scale: metres per unit of s; n_ext: background index.  rhs() forms
    the right-hand side the reduced models are built from at the samples.
    """

    def __init__(self, s, fops, Mc, idx, gram, origin, dx, Eo, src, cutoff, x_thru, m_edge=0, n_ext=1.0,
                 lam_scale=1.0):
        from .configs import C0, EPS0
        self.c0, self.eps0 = C0, EPS0
        self.s, self.fops = np.asarray(s, dtype=float), list(fops)
        self.ctx = self.fops[0].ctx
        Lg, Mg, Ng = np.shape(Mc)
        self.shape = (Lg, Mg, Ng)
        self.V = Lg * Mg * Ng
        self.n = int(np.size(idx))
        self.gram, self.dx, self.cutoff = float(gram), float(dx), float(cutoff)
        self.origin = np.ascontiguousarray(np.asarray(origin, dtype=float).reshape(3))
        self.Eo = np.ascontiguousarray(np.asarray(Eo, dtype=np.complex128).reshape(3))
        self.src = np.ascontiguousarray(np.asarray(src, dtype=np.complex128).reshape(3))
        self.n_ext, self.lam_scale = float(n_ext), float(lam_scale)
        i0 = np.ascontiguousarray(np.asarray(idx, dtype=np.int64).reshape(-1) - 1)
        mc3 = np.ascontiguousarray(np.tile(np.asarray(Mc, dtype=np.complex128).reshape(-1, order='F'), 3)[i0])
        ls = np.arange(m_edge, Mg - m_edge)
        ns = np.arange(Ng)
        cs = np.arange(3)
        l0 = x_thru - 1
        sl = (l0 + Lg * (ls[:, None, None] + Mg * (ns[None, :, None] + Ng * cs[None, None, :]))).reshape(-1)
        self.nsl = sl.size
        dev = lambda a: self._upload(a)      # noqa: E731
        self.d_idx, self.d_mc3, self.d_sl = dev(i0), dev(mc3), dev(np.ascontiguousarray(sl.astype(np.int64)))
        nb = 3 * self.V * 16
        self.d_E0, self.d_Einc, self.d_J, self.d_E = (DeviceArray(self.ctx, nb) for _ in range(4))
        self.d_b, self.d_x = DeviceArray(self.ctx, self.n * 16), DeviceArray(self.ctx, self.n * 16)
        self.d_slice = DeviceArray(self.ctx, self.nsl * 16)
        self.fop = op_lincomb(self.fops, lagrange_weights(self.s[0], self.s))

    def _upload(self, a):
        d = DeviceArray(self.ctx, a.nbytes)
        d.upload(a)
        return d

    def _comp(self, d, c):
        return d.ptr.value + c * self.V * 16

    def incident(self, lam):
        """Normalised incident field on the device (d_Einc); returns the
        per-component maxima it was divided by."""
        ko = 2 * np.pi * self.n_ext / (lam * self.lam_scale)
        Lg, Mg, Ng = self.shape
        check(L.lib.vp_field_dgf_grid(self.ctx.handle, ptr(self.Eo), ptr(self.src), 1, ptr(self.origin), self.dx,
                                      Lg, Mg, Ng, float(ko), self.cutoff, self.d_E0.ptr, L.VP_MEM_DEVICE))
        amax = vec_absmax(self.ctx, [self._comp(self.d_E0, c) for c in range(3)], self.V)
        for c in range(3):
            vec_lincomb(self.ctx, [self._comp(self.d_E0, c)], [[1.0 / amax[c]]], self.V,
                        out=[self._comp(self.d_Einc, c)])
        return amax

    def rhs(self, lam):
        """Vrhs = Gram tau3 Einc(idxS3) on the device (d_b)."""
        freq = self.c0 * self.n_ext / (lam * self.lam_scale)
        self.incident(lam)
        vec_gather(self.ctx, self.d_Einc, 3 * self.V, self.d_idx, self.n, self.d_b, w=self.d_mc3,
                   alpha=self.gram * 1j * 2 * np.pi * freq * self.eps0)
        return self.d_b

    def rhs_copy(self, lam):
        """rhs(lam) in a new DeviceArray (the sample right-hand sides)."""
        return vec_lincomb(self.ctx, [self.rhs(lam)], [[1.0]], self.n)[0]

    def point(self, lam, rm):
        """Transmitted power at lam from the reduced model rm (built over the
        same samples); the total field is left in d_E, the solution in d_x."""
        if rm.n != self.n or not np.array_equal(rm.s, self.s):
            raise ValueError('reduced model of another problem or other samples')
        self.rhs(lam)
        rm.solve(lam, self.d_b, out=self.d_x)
        self.field(lam)
        return self.power()

    def field(self, lam):
        """Total field at lam (d_E) of the current solution d_x."""
        freq = self.c0 * self.n_ext / (lam * self.lam_scale)
        vec_scatter(self.ctx, self.d_x, self.d_idx, self.n, self.d_J, 3 * self.V)
        op_lincomb_into(self.fop, self.fops, lagrange_weights(lam, self.s))
        check(L.lib.vp_field_total(self.fop.handle, self.d_J.ptr, self.d_Einc.ptr, float(freq), self.d_E.ptr,
                                   L.VP_MEM_DEVICE))

    def power(self):
        """Sum of |E|^2 over the cross-section of the field in d_E."""
        vec_gather(self.ctx, self.d_E, 3 * self.V, self.d_sl, self.nsl, self.d_slice)
        s = self.d_slice.download(np.empty(self.nsl, dtype=np.complex128))
        return float(np.sum(s.real ** 2 + s.imag ** 2))
