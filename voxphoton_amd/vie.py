"""Host-side mirror of VoxPhoton's MATLAB call surface for the FFT-VIE solve.

Names, argument meaning and error behaviour follow the reference:
  * ``mv_AN_const``                  pc/src_vie/src_mvp/mv_AN_const.m:1
  * ``chan_2_mvp_for_parallel_idx``  pc/src_preconditioning/chan_2_mvp_for_parallel_idx.m:2
  * ``chan_mvp_idx``                 pc/src_preconditioning/chan_mvp_idx.m:2
  * ``pgmres``                       pc/src_numeric/src_iterative/pgmres.m:70
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
        """From opToeplitz (L x M x N x 6); the embedding and FFT run on the GPU."""
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
        """Green tensor assembled on the GPU for voxel size dx and wavenumber ko."""
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

    def apply_device(self, xdev, ydev, prec='fp64', transp=False):
        """Device pointers (ints) in the given precision; no host traffic."""
        check(L.lib.vp_op_apply(self.handle, int(transp), ctypes.c_void_p(xdev), ctypes.c_void_p(ydev),
                                L.VP_MEM_DEVICE, L.VP_FP64 if prec == 'fp64' else L.VP_FP32))

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
