"""Host-side mirror of VoxPhoton's MATLAB call surface for the FFT-VIE solve.

Names, argument meaning and error behaviour follow the reference:
  * ``mv_AN_const``                  pc/src_vie/src_mvp/mv_AN_const.m:1
  * ``chan_2_mvp_for_parallel_idx``  pc/src_preconditioning/chan_2_mvp_for_parallel_idx.m:2
  * ``chan_mvp_idx``                 pc/src_preconditioning/chan_mvp_idx.m:2
  * ``pgmres``                       pc/src_numeric/src_iterative/pgmres.m:70
  * ``E_field_DGF_mollified``        pc/src_vie/src_fields/src_dgf/E_field_DGF_mollified.m:1
  * ``E_field_Nop_const``            pc/src_vie/src_fields/src_op/E_field_Nop_const.m:1
All arithmetic runs in libvoxphoton.so on the GPU; this module only moves
host arrays across the C ABI and keeps device-resident handles alive.
"""
import ctypes
import weakref

import numpy as np

from . import _lib as L
from ._lib import check, ptr, as_c128, as_idx


class Context:
    """One GPU and one HIP stream (vp_ctx)."""

    def __init__(self, device=0):
        h = ctypes.c_void_p()
        check(L.lib.vp_ctx_create(int(device), ctypes.byref(h)))
        self.handle = h
        self.device = device
        self._fin = weakref.finalize(self, L.lib.vp_ctx_destroy, h)

    def synchronize(self):
        check(L.lib.vp_ctx_synchronize(self.handle))

    def set_recycle_width(self, width):
        """Recycled Krylov space of the mixed-precision solves on this
        context (vp_ctx_set_recycle_width): 0 keeps the first inner solve's
        whole last cycle, k > 0 its k harmonic Ritz directions of smallest
        modulus, -1 the environment's VP_IR_RECYCLE_K (default)."""
        check(L.lib.vp_ctx_set_recycle_width(self.handle, int(width)))

    def mem_info(self):
        """(free, total) device memory in bytes (vp_mem_info)."""
        f, t = ctypes.c_int64(), ctypes.c_int64()
        check(L.lib.vp_mem_info(self.handle, ctypes.byref(f), ctypes.byref(t)))
        return f.value, t.value

    def close(self):
        self._fin()


_default_ctx = None


def default_context():
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


def _prec_mask(precisions):
    if isinstance(precisions, int):
        return precisions
    m = 0
    for p in precisions:
        m |= {'fp64': L.VP_FP64, 'fp32': L.VP_FP32}[p]
    return m


class NOperator:
    """Device-resident JVIE-II operator A = Gram*Mr - Mc*N restricted to idx.

    Create with one of the classmethods; ``apply`` mirrors mv_AN_const.
    """

    def __init__(self, handle, ctx, n, refs=()):
        self.handle = handle
        self.ctx = ctx
        self.n = int(n)
        self._refs = refs
        s = ctypes.c_double()
        check(L.lib.vp_op_fp32_scale(handle, ctypes.byref(s)))
        self.fp32_scale = s.value
        self._fin = weakref.finalize(self, L.lib.vp_op_destroy, handle)

    @staticmethod
    def _mats(Mr, Mc):
        mr = None if Mr is None else as_c128(np.asarray(Mr).reshape(-1, order='F'))
        mc = as_c128(np.asarray(Mc).reshape(-1, order='F'))
        return mr, mc

    @classmethod
    def from_fN(cls, fN, Mr, Mc, Gram, idx, precisions=('fp64', 'fp32'), ctx=None):
        """From the reference's circulant tensor fN (Lp x Mp x Np x 6)."""
        ctx = ctx or default_context()
        fN = np.asarray(fN)
        Lp, Mp, Np = fN.shape[:3]
        Lg, Mg, Ng = np.shape(Mc)
        f = as_c128(fN.reshape(-1, order='F'))
        mr, mc = cls._mats(Mr, Mc)
        ix = as_idx(idx)
        h = ctypes.c_void_p()
        check(L.lib.vp_op_create(ctx.handle, Lg, Mg, Ng, Lp, Mp, Np, ptr(f), ptr(mr), ptr(mc), float(Gram),
                                 ptr(ix), ix.size, _prec_mask(precisions), ctypes.byref(h)))
        return cls(h, ctx, ix.size, (ctx,))

    @classmethod
    def from_toeplitz(cls, G, Mr, Mc, Gram, idx, precisions=('fp64', 'fp32'), fft_shape=(0, 0, 0), ctx=None):
        """From opToeplitz (L x M x N x 6); the embedding and FFT run on the GPU.
        ``fft_shape`` (0, 0, 0) lets the autotuner choose the FFT sizes (at
        least 2B-1 for the unknowns' box B); any such sizes give the same operator."""
        ctx = ctx or default_context()
        G = np.asarray(G)
        Lg, Mg, Ng = G.shape[:3]
        g = as_c128(G.reshape(-1, order='F'))
        mr, mc = cls._mats(Mr, Mc)
        ix = as_idx(idx)
        h = ctypes.c_void_p()
        check(L.lib.vp_op_create_toeplitz(ctx.handle, Lg, Mg, Ng, *[int(v) for v in fft_shape], ptr(g), ptr(mr),
                                          ptr(mc), float(Gram), ptr(ix), ix.size, _prec_mask(precisions),
                                          ctypes.byref(h)))
        return cls(h, ctx, ix.size, (ctx,))

    @classmethod
    def assembled(cls, shape, dx, ko, Mr, Mc, idx, precisions=('fp64', 'fp32'), fft_shape=(0, 0, 0), ctx=None):
        """Green tensor assembled on the GPU for voxel size dx and wavenumber ko
        (FFT sizes as ``from_toeplitz``)."""
        ctx = ctx or default_context()
        Lg, Mg, Ng = shape
        mr, mc = cls._mats(Mr, Mc)
        ix = as_idx(idx)
        h = ctypes.c_void_p()
        check(L.lib.vp_op_create_assembled(ctx.handle, Lg, Mg, Ng, *[int(v) for v in fft_shape], float(dx),
                                           float(ko), ptr(mr), ptr(mc), ptr(ix), ix.size,
                                           _prec_mask(precisions), ctypes.byref(h)))
        return cls(h, ctx, ix.size, (ctx,))

    def apply(self, x, transp_flag='notransp', prec='fp64'):
        """A x (or A^H x) for a host complex128 vector (mv_AN_const.m:1-146)."""
        if transp_flag not in ('notransp', 'transp'):
            raise ValueError("transp_flag must be 'transp' or 'notransp'")
        x = as_c128(np.asarray(x).reshape(-1))
        if x.size != self.n:
            raise ValueError(f'vector length {x.size} != operator size {self.n}')
        y = np.empty_like(x)
        check(L.lib.vp_op_apply(self.handle, 1 if transp_flag == 'transp' else 0, ptr(x), ptr(y), L.VP_MEM_HOST,
                                L.VP_FP64 if prec == 'fp64' else L.VP_FP32))
        return y

    def info(self):
        """Grid, unknowns' bounding box and the FFT sizes / variant of each
        precision's plan (vp_op_info)."""
        v = (ctypes.c_int64 * 17)()
        check(L.lib.vp_op_info(self.handle, v))
        v = list(v)
        return dict(grid=tuple(v[0:3]), box_origin=tuple(v[3:6]), box=tuple(v[6:9]),
                    fft_fp64=tuple(v[9:12]), fft_fp32=tuple(v[12:15]), variant_fp64=v[15], variant_fp32=v[16])

    def apply_device(self, xdev, ydev, prec='fp64', transp=False):
        """Device pointers (ints) in the given precision; no host traffic."""
        check(L.lib.vp_op_apply(self.handle, int(transp), ctypes.c_void_p(xdev), ctypes.c_void_p(ydev),
                                L.VP_MEM_DEVICE, L.VP_FP64 if prec == 'fp64' else L.VP_FP32))

    @classmethod
    def for_fields(cls, shape, Gram=None, fN=None, G=None, dx=None, ko=None, precisions=('fp64',), ctx=None):
        """Field operator over every voxel with unit contrast (A = Gram I - N),
        the form E_field_Nop_const uses through mv_L_const: from fN or
        opToeplitz G (with Gram), or assembled on the GPU for dx and ko
        (Gram = dx^3)."""
        Lg, Mg, Ng = shape
        idx = np.arange(1, 3 * Lg * Mg * Ng + 1)
        ones = np.ones(shape, complex)
        if fN is not None:
            return cls.from_fN(fN, None, ones, Gram, idx, precisions=precisions, ctx=ctx)
        if G is not None:
            return cls.from_toeplitz(G, None, ones, Gram, idx, precisions=precisions, ctx=ctx)
        return cls.assembled(shape, dx, ko, None, ones, idx, precisions=precisions, ctx=ctx)

    def close(self):
        self._fin()


def assemble_toeplitz(shape, dx, ko, ctx=None):
    """opToeplitz (L x M x N x 6, complex128) assembled on the GPU."""
    ctx = ctx or default_context()
    Lg, Mg, Ng = shape
    G = np.empty(Lg * Mg * Ng * 6, dtype=np.complex128)
    check(L.lib.vp_assemble_toeplitz(ctx.handle, Lg, Mg, Ng, float(dx), float(ko), ptr(G), L.VP_MEM_HOST))
    return G.reshape((Lg, Mg, Ng, 6), order='F')


_op_cache = {}


def mv_AN_const(JIn0, fN, Mr, Mc, Gram, transp_flag, idx, GPU_flag=1):
    """JOut = mv_AN_const(JIn0, fN, Mr, Mc, Gram, transp_flag, idx, GPU_flag).

    Mirrors pc/src_vie/src_mvp/mv_AN_const.m:1.  The first call with a given
    (fN, Mr, Mc, Gram, idx) uploads the operator (as the reference's
    JVIE_Solver_const.m:277-279 does once); later calls reuse it.  Only the
    GPU path exists: ``GPU_flag`` is accepted for signature compatibility.
    """
    key = (id(fN), id(Mr), id(Mc), float(Gram), id(idx))
    ent = _op_cache.get(key)
    if ent is None:
        op = NOperator.from_fN(fN, Mr, Mc, Gram, idx, precisions=('fp64',))
        _op_cache[key] = ent = (op, (fN, Mr, Mc, idx))
    return ent[0].apply(JIn0, transp_flag, 'fp64')


class Preconditioner:
    """Device-resident Chan block-circulant preconditioner (handle returns M^-1 x)."""

    def __init__(self, handle, ctx, nbox):
        self.handle = handle
        self.ctx = ctx
        self.nbox = int(nbox)
        self.n = int(nbox)
        self._fin = weakref.finalize(self, L.lib.vp_prec_destroy, handle)

    @classmethod
    def chan2(cls, blocks, L_, M_, N_, idx, precisions=('fp64', 'fp32'), ctx=None):
        """From inverse blocks inv_blocks{i,j} given as an (L, M, 3N, 3N) array
        (chan_2_mvp_for_parallel_idx.m:2-23)."""
        ctx = ctx or default_context()
        B = np.asarray(blocks)
        # C ABI layout: cell {i,j} at (i + L*j), each column-major
        flat = as_c128(np.transpose(B, (1, 0, 3, 2)).reshape(-1))
        ix = as_idx(idx)
        h = ctypes.c_void_p()
        check(L.lib.vp_prec_create_chan2(ctx.handle, L_, M_, N_, ptr(flat), ptr(ix), ix.size,
                                         _prec_mask(precisions), ctypes.byref(h)))
        return cls(h, ctx, ix.size)

    @classmethod
    def chan1(cls, blocks, L_, M_, N_, idx, precisions=('fp64', 'fp32'), ctx=None):
        """From inverse blocks inv_blocks{i} given as an (L, 3MN, 3MN) array (chan_mvp_idx.m:2-16)."""
        ctx = ctx or default_context()
        B = np.asarray(blocks)
        flat = as_c128(np.transpose(B, (0, 2, 1)).reshape(-1))
        ix = as_idx(idx)
        h = ctypes.c_void_p()
        check(L.lib.vp_prec_create_chan1(ctx.handle, L_, M_, N_, ptr(flat), ptr(ix), ix.size,
                                         _prec_mask(precisions), ctypes.byref(h)))
        return cls(h, ctx, ix.size)

    @classmethod
    def level1(cls, G, Mc, Mr0, Gram, idx, precisions=('fp64', 'fp32'), ctx=None):
        """level_1_parallel_func(opToeplitz, Mc, Mr, L, M, N, Gram), built and inverted on the GPU; Mc a
        scalar or an L x M x N array (vp_prec_create_level1)."""
        ctx = ctx or default_context()
        G = np.asarray(G)
        Lg, Mg, Ng = G.shape[:3]
        g = as_c128(G.reshape(-1, order='F'))
        ix = as_idx(idx)
        h = ctypes.c_void_p()
        mcf, mc = (_contrast_field(Mc, (Lg, Mg, Ng)), 0j) if np.ndim(Mc) else (None, complex(Mc))
        check(L.lib.vp_prec_create_level1(ctx.handle, Lg, Mg, Ng, ptr(g), mc.real, mc.imag,
                                          ptr(mcf) if mcf is not None else None, float(Mr0), float(Gram), ptr(ix),
                                          ix.size, _prec_mask(precisions), ctypes.byref(h)))
        return cls(h, ctx, ix.size)

    @classmethod
    def level1_assembled(cls, shape, dx, ko, Mc, Mr0, idx, precisions=('fp64', 'fp32'), ctx=None, damping=0.0):
        """1-level preconditioner from a Green tensor assembled on the GPU (vp_prec_create_level1_assembled)."""
        ctx = ctx or default_context()
        Lg, Mg, Ng = shape
        ix = as_idx(idx)
        h = ctypes.c_void_p()
        mcf, mc = (_contrast_field(Mc, (Lg, Mg, Ng)), 0j) if np.ndim(Mc) else (None, complex(Mc))
        check(L.lib.vp_prec_create_level1_assembled(ctx.handle, Lg, Mg, Ng, float(dx), float(ko), float(damping),
                                                    mc.real, mc.imag, ptr(mcf) if mcf is not None else None,
                                                    float(Mr0), ptr(ix), ix.size, _prec_mask(precisions),
                                                    ctypes.byref(h)))
        return cls(h, ctx, ix.size)

    @classmethod
    def level1_face(cls, shape, dx, ko, face, idx, precisions=('fp64', 'fp32'), ctx=None, damping=0.0):
        """Face-averaged 1-level form of Bragg_MOR_script.m:61-72 (level_1_parallel_func_N): blocks
        inv(Gram I - diag(face3) circ_N); face is M x N (vp_prec_create_level1_face)."""
        ctx = ctx or default_context()
        Lg, Mg, Ng = shape
        fc = _contrast_field(face, (Mg, Ng))
        ix = as_idx(idx)
        h = ctypes.c_void_p()
        check(L.lib.vp_prec_create_level1_face(ctx.handle, Lg, Mg, Ng, float(dx), float(ko), float(damping), ptr(fc),
                                               ptr(ix), ix.size, _prec_mask(precisions), ctypes.byref(h)))
        return cls(h, ctx, ix.size)

    @classmethod
    def level2(cls, G, Mc, Mr0, Gram, idx, precisions=('fp64', 'fp32'), ctx=None):
        """level_2_parallel_func(opToeplitz, Mc, Mr, L, M, N, Gram), built on the GPU; Mc a scalar or an
        L x M x N array (vp_prec_create_level2 / vp_prec_create_level2_field)."""
        ctx = ctx or default_context()
        G = np.asarray(G)
        Lg, Mg, Ng = G.shape[:3]
        g = as_c128(G.reshape(-1, order='F'))
        ix = as_idx(idx)
        h = ctypes.c_void_p()
        if np.ndim(Mc):
            mcf = _contrast_field(Mc, (Lg, Mg, Ng))
            check(L.lib.vp_prec_create_level2_field(ctx.handle, Lg, Mg, Ng, ptr(g), ptr(mcf), float(Mr0), float(Gram),
                                                    ptr(ix), ix.size, _prec_mask(precisions), ctypes.byref(h)))
            return cls(h, ctx, ix.size)
        mc = complex(Mc)
        check(L.lib.vp_prec_create_level2(ctx.handle, Lg, Mg, Ng, ptr(g), mc.real, mc.imag, float(Mr0),
                                          float(Gram), ptr(ix), ix.size, _prec_mask(precisions),
                                          ctypes.byref(h)))
        return cls(h, ctx, ix.size)

    @classmethod
    def level2_assembled(cls, shape, dx, ko, Mc, Mr0, idx, precisions=('fp64', 'fp32'), ctx=None, damping=0.0):
        """2-level preconditioner from a Green tensor assembled on the GPU."""
        ctx = ctx or default_context()
        Lg, Mg, Ng = shape
        ix = as_idx(idx)
        h = ctypes.c_void_p()
        if np.ndim(Mc):
            mcf = _contrast_field(Mc, (Lg, Mg, Ng))
            check(L.lib.vp_prec_create_level2_assembled_field(ctx.handle, Lg, Mg, Ng, float(dx), float(ko),
                                                              float(damping), ptr(mcf), float(Mr0), ptr(ix), ix.size,
                                                              _prec_mask(precisions), ctypes.byref(h)))
            return cls(h, ctx, ix.size)
        mc = complex(Mc)
        check(L.lib.vp_prec_create_level2_assembled(ctx.handle, Lg, Mg, Ng, float(dx), float(ko), float(damping), mc.real,
                                                    mc.imag, float(Mr0), ptr(ix), ix.size,
                                                    _prec_mask(precisions), ctypes.byref(h)))
        return cls(h, ctx, ix.size)

    @classmethod
    def level2_face(cls, shape, dx, ko, face, idx, precisions=('fp64', 'fp32'), ctx=None, damping=0.0):
        """Face-averaged 2-level preconditioner, exact along the last axis of
        ``shape`` with contrast profile ``face`` (complex, length shape[2])."""
        ctx = ctx or default_context()
        Lg, Mg, Ng = shape
        f = as_c128(np.asarray(face).reshape(-1))
        if f.size != Ng:
            raise ValueError('face profile length must equal the exact-axis extent')
        ix = as_idx(idx)
        h = ctypes.c_void_p()
        check(L.lib.vp_prec_create_level2_face(ctx.handle, Lg, Mg, Ng, float(dx), float(ko), float(damping), ptr(f), ptr(ix),
                                               ix.size, _prec_mask(precisions), ctypes.byref(h)))
        return cls(h, ctx, ix.size)

    @classmethod
    def composite(cls, parts, ctx=None):
        """Block-Jacobi composite of embedded preconditioners on disjoint
        unknowns (prec_fun_disk_idx.m:1-23); keeps the parts alive."""
        ctx = ctx or parts[0].ctx
        hs = (ctypes.c_void_p * len(parts))(*[q.handle.value for q in parts])
        h = ctypes.c_void_p()
        check(L.lib.vp_prec_create_composite(ctx.handle, ctypes.cast(hs, ctypes.c_void_p), len(parts),
                                             ctypes.byref(h)))
        out = cls(h, ctx, parts[0].n)
        out.parts = list(parts)
        return out

    def set_embedding(self, n_outer, sub_of_op):
        """Act on an outer vector of length n_outer; box unknown i sits at
        sub_of_op[i] (1-based); other entries pass through (block Jacobi)."""
        s = as_idx(sub_of_op)
        check(L.lib.vp_prec_set_embedding(self.handle, int(n_outer), ptr(s), s.size))
        self.n = int(n_outer)
        return self

    def apply(self, x, prec='fp64'):
        x = as_c128(np.asarray(x).reshape(-1))
        if x.size != self.n:
            raise ValueError(f'vector length {x.size} != preconditioner size {self.n}')
        y = np.empty_like(x)
        check(L.lib.vp_prec_apply(self.handle, ptr(x), ptr(y), L.VP_MEM_HOST,
                                  L.VP_FP64 if prec == 'fp64' else L.VP_FP32))
        return y

    __call__ = apply


def _contrast_field(Mc, shape):
    """A contrast (or face) array of the given shape as interleaved complex128, column-major."""
    Mc = np.asarray(Mc)
    if Mc.shape != tuple(shape):
        raise ValueError(f'contrast field of shape {Mc.shape}, expected {tuple(shape)}')
    return as_c128(Mc.reshape(-1, order='F'))


def _info(cap):
    buf = np.zeros(max(int(cap), 1), dtype=np.float64)
    info = L.GmresInfo()
    info.resvec = buf.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    info.resvec_cap = buf.size
    return info, buf


def pgmres(A, b, restart=50, tol=1e-3, maxit=10, L1=None, L2=None, R1=None, R2=None, x0=None, prec='fp64'):
    """[x, flag, relres, iter, resvec] = pgmres(A, b, restart, tol, maxit, L1, L2, R1, R2, x0).

    Mirrors pc/src_numeric/src_iterative/pgmres.m:70.  ``A`` is an NOperator,
    L1 or R1 a Preconditioner (one side; L2/R2 are not supported).  The
    Krylov loop runs on the GPU in ``prec``; non-convergence gives flag = 1.
    """
    if L2 is not None or R2 is not None:
        raise ValueError('a second preconditioner factor is not supported')
    if L1 is not None and R1 is not None:
        raise ValueError('use either a left or a right preconditioner')
    Mh = (L1 or R1)
    side = L.VP_PREC_RIGHT if R1 is not None else L.VP_PREC_LEFT
    b = as_c128(np.asarray(b).reshape(-1))
    x = np.empty_like(b)
    x0a = None if x0 is None or not np.any(x0) else as_c128(np.asarray(x0).reshape(-1))
    info, buf = _info(restart * maxit + 1)
    check(L.lib.vp_gmres(A.handle, Mh.handle if Mh is not None else None, side, ptr(b), ptr(x0a), ptr(x),
                         L.VP_MEM_HOST, L.VP_FP64 if prec == 'fp64' else L.VP_FP32, int(restart), float(tol),
                         int(maxit), ctypes.byref(info)))
    return x, info.flag, info.relres, [info.iter_inner, info.iter_outer], buf[:info.resvec_len].copy()


def bicgstab(A, b, tol=1e-6, maxit=20, M=None, x0=None, prec='fp64'):
    """[x, flag, relres, iter, resvec] = bicgstab(A, b, tol, maxit, M, [], x0).

    MATLAB's call as the harness makes it (JVIE_Solver_const.m:469-494);
    ``M`` a Preconditioner applied as M^-1 (or None).  ``iter`` counts whole
    iterations, plus 0.5 when it stopped at a half iteration; ``resvec`` the
    relative residual at the start and after every half iteration."""
    b = as_c128(np.asarray(b).reshape(-1))
    x = np.empty_like(b)
    x0a = None if x0 is None or not np.any(x0) else as_c128(np.asarray(x0).reshape(-1))
    info, buf = _info(2 * int(maxit) + 2)
    check(L.lib.vp_bicgstab(A.handle, M.handle if M is not None else None, ptr(b), ptr(x0a), ptr(x), L.VP_MEM_HOST,
                            L.VP_FP64 if prec == 'fp64' else L.VP_FP32, float(tol), int(maxit), ctypes.byref(info)))
    it = info.iter_inner + 0.5 * info.iter_outer
    return x, info.flag, info.relres, it, buf[:info.resvec_len].copy()


def solve_ir(A, b, tol=1e-6, restart=50, inner_tol=1e-3, inner_maxit=4, max_outer=20, M=None, side='left',
             outer_restart=0):
    """Mixed precision: FP64 refinement around FP32 GMRES(restart) on the GPU.

    ``outer_restart`` > 0 runs the flexible form (vp_solve_fgmres_ir: FGMRES
    in FP64 over the inner solves, cycles of that many steps); 0 the plain
    refinement (vp_solve_ir).  Returns (x, info dict) with the true FP64
    relative residual ``relres``.
    """
    b = as_c128(np.asarray(b).reshape(-1))
    x = np.empty_like(b)
    info, buf = _info(max_outer + 2)
    args = [A.handle, M.handle if M is not None else None, L.VP_PREC_RIGHT if side == 'right' else L.VP_PREC_LEFT,
            ptr(b), ptr(x), L.VP_MEM_HOST, float(tol), int(restart), float(inner_tol), int(inner_maxit),
            int(max_outer)]
    if outer_restart:
        check(L.lib.vp_solve_fgmres_ir(*args, int(outer_restart), ctypes.byref(info)))
    else:
        check(L.lib.vp_solve_ir(*args, ctypes.byref(info)))
    return x, dict(flag=info.flag, relres=info.relres, ir_steps=info.ir_steps, inner_iters=info.total_iters,
                   seconds=info.seconds, history=buf[:info.resvec_len].copy())


class DeviceArray:
    """Minimal owned device buffer (vp_malloc) for batched calls."""

    def __init__(self, ctx, nbytes):
        p = ctypes.c_void_p()
        check(L.lib.vp_malloc(ctx.handle, int(nbytes), ctypes.byref(p)))
        self.ptr = p
        self.ctx = ctx
        self.nbytes = int(nbytes)
        self._fin = weakref.finalize(self, L.lib.vp_free, ctx.handle, p)

    def upload(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        check(L.lib.vp_memcpy_h2d(self.ctx.handle, self.ptr, ptr(a), a.nbytes))

    def download(self, a):
        assert a.flags['C_CONTIGUOUS'] and a.nbytes <= self.nbytes
        check(L.lib.vp_memcpy_d2h(self.ctx.handle, ptr(a), self.ptr, a.nbytes))
        return a


def solve_ir_batched(ops, precs, B, tol=1e-6, restart=50, inner_tol=1e-3, inner_maxit=4, max_outer=20,
                     side='left', outer_restart=0):
    """Batched mixed-precision solves: row i of B through ops[i] (and
    precs[i]); ``outer_restart`` as in solve_ir."""
    ctx = ops[0].ctx
    B = as_c128(np.asarray(B))
    ns, n = B.shape
    db = DeviceArray(ctx, B.nbytes)
    dx = DeviceArray(ctx, B.nbytes)
    db.upload(B)
    hops = (ctypes.c_void_p * ns)(*[o.handle.value for o in ops])
    hpre = None
    if precs is not None:
        hpre = (ctypes.c_void_p * ns)(*[p.handle.value for p in precs])
    infos = (L.GmresInfo * ns)()
    bufs = []
    for i in range(ns):
        buf = np.zeros(max_outer + 2)
        bufs.append(buf)
        infos[i].resvec = buf.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        infos[i].resvec_cap = buf.size
    args = [ctypes.cast(hops, ctypes.c_void_p), ctypes.cast(hpre, ctypes.c_void_p) if hpre is not None else None, ns,
            L.VP_PREC_RIGHT if side == 'right' else L.VP_PREC_LEFT, db.ptr, dx.ptr, float(tol), int(restart),
            float(inner_tol), int(inner_maxit), int(max_outer)]
    if outer_restart:
        check(L.lib.vp_solve_fgmres_ir_batched(*args, int(outer_restart),
                                               ctypes.cast(infos, ctypes.POINTER(L.GmresInfo))))
    else:
        check(L.lib.vp_solve_ir_batched(*args, ctypes.cast(infos, ctypes.POINTER(L.GmresInfo))))
    X = np.empty_like(B)
    dx.download(X)
    out = [dict(flag=infos[i].flag, relres=infos[i].relres, ir_steps=infos[i].ir_steps,
                inner_iters=infos[i].total_iters, seconds=infos[i].seconds,
                history=bufs[i][:infos[i].resvec_len].copy()) for i in range(ns)]
    return X, out


class Server:
    """Request server with continuous batching (vp_server_*).

    ``submit`` queues one mixed-precision solve of device vectors (b, x are
    DeviceArray objects of n complex128) and returns a ticket; ``wait``
    returns its info dict, with ``seconds`` = submit-to-finish latency.
    Operators, preconditioners and the vectors must stay alive and unused by
    the caller until the request has finished.
    """

    def __init__(self, hbm_budget_bytes, max_active=64, ctx=None):
        self.ctx = ctx or default_context()
        h = ctypes.c_void_p()
        check(L.lib.vp_server_create(self.ctx.handle, int(hbm_budget_bytes), int(max_active), ctypes.byref(h)))
        self.handle = h
        self._keep = {}
        self._fin = weakref.finalize(self, L.lib.vp_server_destroy, h)

    def _renew_finalizer(self):
        # At interpreter exit finalizers run newest first.  The server's own
        # is made newer than everything a request uses (all created before
        # its submit), so the server stops and drains its executors before
        # those operators and vectors are freed, also when the caller leaves
        # on an exception with requests in flight.
        if self._fin.detach() is not None:
            self._fin = weakref.finalize(self, L.lib.vp_server_destroy, self.handle)

    def submit(self, op, M, b, x, tol=1e-6, restart=50, inner_tol=1e-3, inner_maxit=4, max_outer=20, side='right'):
        prm = L.SolveParams(float(tol), int(restart), float(inner_tol), int(inner_maxit), int(max_outer),
                            L.VP_PREC_RIGHT if side == 'right' else L.VP_PREC_LEFT)
        t = ctypes.c_int64()
        self._renew_finalizer()
        check(L.lib.vp_server_submit(self.handle, op.handle, M.handle if M is not None else None, b.ptr, x.ptr,
                                     ctypes.byref(prm), ctypes.byref(t)))
        self._keep[t.value] = (op, M, b, x, max_outer)
        return t.value

    def _info(self, ticket, fn, *args):
        cap = self._keep[ticket][4] + 2
        info, buf = _info(cap)
        check(fn(self.handle, ticket, *args, ctypes.byref(info)))
        return dict(flag=info.flag, relres=info.relres, ir_steps=info.ir_steps, inner_iters=info.total_iters,
                    seconds=info.seconds, history=buf[:info.resvec_len].copy())

    def poll(self, ticket):
        done = ctypes.c_int()
        out = self._info(ticket, L.lib.vp_server_poll, ctypes.byref(done))
        return bool(done.value), (out if done.value else None)

    def wait(self, ticket):
        out = self._info(ticket, L.lib.vp_server_wait)
        self._keep.pop(ticket, None)
        return out

    def stats(self):
        v = [ctypes.c_int64() for _ in range(4)]
        check(L.lib.vp_server_stats(self.handle, *[ctypes.byref(a) for a in v]))
        return dict(pending=v[0].value, active=v[1].value, bytes_in_use=v[2].value, rounds=v[3].value)

    def close(self):
        if self.handle:
            self._fin.detach()
            h, self.handle = self.handle, None
            try:
                check(L.lib.vp_server_destroy(h))
            finally:
                self._keep.clear()


def E_field_DGF_mollified(J, Icoord, Ocoord, ko, cutoff, ctx=None):
    """Field of point currents J (3*Ni, column-major Ni x 3) at Icoord
    (Ni x 3, complex coordinates allowed) observed at Ocoord (No x 3);
    returns No x 3 (E_field_DGF_mollified.m:1-111)."""
    ctx = ctx or default_context()
    ic = np.atleast_2d(np.asarray(Icoord, dtype=complex))
    ni = ic.shape[0]
    j = as_c128(np.asarray(J).reshape(-1))
    if j.size != 3 * ni:
        raise ValueError('J must hold 3 components per source')
    icf = as_c128(ic.reshape(-1, order='F'))
    oc = np.atleast_2d(np.asarray(Ocoord, dtype=float))
    no = oc.shape[0]
    ocf = np.ascontiguousarray(oc.reshape(-1, order='F'))
    E = np.empty(3 * no, dtype=complex)
    check(L.lib.vp_field_dgf(ctx.handle, ptr(j), ptr(icf), ni, ptr(ocf), no, float(ko), float(cutoff), ptr(E),
                             L.VP_MEM_HOST))
    return E.reshape(no, 3, order='F')


def E_field_DGF_grid(J, Icoord, origin, dx, shape, ko, cutoff, ctx=None):
    """E_field_DGF_mollified at the voxel centres origin + (l, m, n) dx of a
    grid (the configs' Ocoord = reshape(r, L*M*N, 3)); returns L x M x N x 3."""
    ctx = ctx or default_context()
    ic = np.atleast_2d(np.asarray(Icoord, dtype=complex))
    ni = ic.shape[0]
    j = as_c128(np.asarray(J).reshape(-1))
    if j.size != 3 * ni:
        raise ValueError('J must hold 3 components per source')
    icf = as_c128(ic.reshape(-1, order='F'))
    o = np.ascontiguousarray(np.asarray(origin, dtype=float).reshape(3))
    Lg, Mg, Ng = shape
    E = np.empty(3 * Lg * Mg * Ng, dtype=complex)
    check(L.lib.vp_field_dgf_grid(ctx.handle, ptr(j), ptr(icf), ni, ptr(o), float(dx), Lg, Mg, Ng, float(ko),
                                  float(cutoff), ptr(E), L.VP_MEM_HOST))
    return E.reshape(Lg, Mg, Ng, 3, order='F')


def E_field_Nop_const(J, op, freq, Einc=None):
    """Total field Einc + (N(J) - Gram J) / (Gram j w eps0) (E_field_Nop_const.m:1-84).

    ``op`` is a field operator (``NOperator.for_fields``) of the grid, the
    counterpart of the reference's (fN, Gram) pair; J, Einc: L x M x N x 3.
    A zero J returns Einc, as the reference."""
    J = np.asarray(J)
    if Einc is None:
        Einc = np.zeros(J.shape, dtype=complex)
    if not np.any(J):
        return np.array(Einc, dtype=complex)
    j = as_c128(J.reshape(-1, order='F'))
    if j.size != op.n:
        raise ValueError(f'field of {j.size} entries for an operator over {op.n}')
    e0 = as_c128(np.asarray(Einc).reshape(-1, order='F'))
    E = np.empty_like(j)
    check(L.lib.vp_field_total(op.handle, ptr(j), ptr(e0), float(freq), ptr(E), L.VP_MEM_HOST))
    return E.reshape(J.shape, order='F')


def power_through(E, x_thru, m_edge):
    """Transmitted power: sum of |E|^2 over the cross-section x = x_thru
    (1-based), m_edge < m <= M - m_edge, all n and components
    (example_Bragg_1500nm_to_1600nm.m:108-113), on the host field."""
    M = E.shape[1]
    s = E[x_thru - 1, m_edge:M - m_edge, :, :]
    return float(np.sum((s * np.conj(s)).real))


def JVIE_Solver_const(Einc, e_r, r, freq, fN=None, tol=1e-3, form=2, solver='B', precond='', max_it=1000,
                      inner_it=50, outer_it=200, ritz=5, ctx=None):
    """[Je, flag, relres, iter, resvec, solvert, vsol] = JVIE_Solver_const(Einc, e_r, r, freq, fN, ...).

    The reference's one-call harness (pc/src_vie/src_solver/JVIE_Solver_const.m:1-521)
    over the device solve path:

    * Mr = e_r, Mc = e_r - 1, unknowns where |Mc| > 1e-12 (:133-149);
    * Vrhs = Gram tau3(idxS3) .* Einc(idxS3), tau = j w eps0 Mc, Gram = dx^3
      with dx the z spacing of ``r`` (:137-157);
    * form 2 (JVIE-II, default): Mc = Mc ./ Mr, Vrhs ./= Mr, Mr = 1 (:243-251);
      form 1 keeps Mr (JVIE-I);
    * solver 'G' pgmres(inner_it, tol, outer_it), 'D' pgmresDR(inner_it, tol,
      outer_it, ritz), 'B' bicgstab(tol, max_it) (:312-508);
    * Je = zeros(L,M,N,3), Je(idxS3) = vsol (:520-521).

    ``fN`` is the reference's circulant tensor; None assembles the Green
    tensor on the device for the grid spacing of ``r`` and the wavenumber
    of ``freq``.  The reference's "preconditioner" ('L'/'R') is the operator
    mv_AN_const with Mr' = 1/Mr and Mc' = 1/Mr - 1 (:301-303), which under
    JVIE-II is Gram I: a scalar, which leaves every GMRES and BiCGSTAB
    iterate unchanged, so it is accepted and not applied; under JVIE-I it is
    not available here (ValueError).  'F' (the FFTW BiCGSTAB) and 'M'
    (automatic) run 'B'.
    """
    import time
    from .krylov_dr import pgmresDR
    ctx = ctx or default_context()
    Einc = np.asarray(Einc, dtype=complex)
    e_r = np.asarray(e_r, dtype=complex)
    r = np.asarray(r, dtype=float)
    Lg, Mg, Ng = e_r.shape
    omega = 2 * np.pi * freq
    c0 = 299792458.0
    mu0 = 4e-7 * np.pi
    eps0 = 1.0 / (c0 ** 2 * mu0)
    ko = omega / c0
    dx = r[0, 0, 1, 2] - r[0, 0, 0, 2] if Ng > 1 else (r[1, 0, 0, 0] - r[0, 0, 0, 0] if Lg > 1 else
                                                       r[0, 1, 0, 1] - r[0, 0, 0, 1])
    Gram = dx ** 3
    Mr = e_r.copy()
    Mc = e_r - 1.0
    nD = Lg * Mg * Ng
    idxS = np.flatnonzero(np.abs(Mc.reshape(-1, order='F')) > 1e-12) + 1
    idxS3 = np.concatenate([idxS, nD + idxS, 2 * nD + idxS]).astype(np.int64)
    tau3 = np.tile((1j * omega * eps0 * Mc).reshape(-1, order='F'), 3)
    i0 = idxS3 - 1
    Vrhs = Gram * tau3[i0] * Einc.reshape(-1, order='F')[i0]
    if form == 2:
        mr = Mr.reshape(-1, order='F')[idxS - 1]
        Mc = Mc / Mr
        Vrhs = Vrhs / np.concatenate([mr, mr, mr])
        Mr_op = None
    elif form == 1:
        if precond in ('L', 'R'):
            raise ValueError('the JVIE-I preconditioner operator is not available')
        Mr_op = Mr
    else:
        raise ValueError('form must be 1 or 2')
    if fN is not None:
        op = NOperator.from_fN(fN, Mr_op, Mc, Gram, idxS3, precisions=('fp64',), ctx=ctx)
    else:
        # the harness's Green tensor is in the exterior medium of the grid (ko = w / c0)
        op = NOperator.assembled((Lg, Mg, Ng), dx, ko, Mr_op, Mc, idxS3, precisions=('fp64',), ctx=ctx)
    t0 = time.perf_counter()
    if solver == 'G':
        vsol, flag, relres, it, resvec = pgmres(op, Vrhs, inner_it, tol, outer_it)
    elif solver == 'D':
        vsol, flag, relres, it, resvec = pgmresDR(op, Vrhs, inner_it, tol, outer_it, ritz)
    elif solver in ('B', 'F', 'M'):
        vsol, flag, relres, it, resvec = bicgstab(op, Vrhs, tol, max_it)
    else:
        raise ValueError("solver must be 'G', 'D', 'B', 'F' or 'M'")
    solvert = time.perf_counter() - t0
    Je = np.zeros(3 * nD, dtype=complex)
    Je[i0] = vsol
    return Je.reshape(Lg, Mg, Ng, 3, order='F'), flag, relres, it, resvec, solvert, vsol
