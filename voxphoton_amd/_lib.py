"""ctypes binding of libvoxphoton.so (the C ABI declared in include/voxphoton.h).

The library is loaded from this package directory only; there is no fallback.
Importing this module without the built library raises ImportError.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, 'libvoxphoton.so')

VP_OK = 0
VP_ERR_INVALID = -1
VP_ERR_OOM = -2
VP_ERR_DEVICE = -3
VP_ERR_UNSUPPORTED = -4
VP_MEM_HOST = 0
VP_MEM_DEVICE = 1
VP_FP64 = 1
VP_FP32 = 2
VP_NOTRANSP = 0
VP_TRANSP = 1
VP_PREC_LEFT = 0
VP_PREC_RIGHT = 1

if not os.path.exists(LIB_PATH):
    raise ImportError(f'{LIB_PATH} is missing: build it with voxphoton_amd.build.build_library() '
                      '(there is no CPU fallback)')

lib = ctypes.CDLL(LIB_PATH, mode=ctypes.RTLD_GLOBAL)

c_int = ctypes.c_int
c_i64 = ctypes.c_int64
c_dbl = ctypes.c_double
c_vp = ctypes.c_void_p
P = ctypes.POINTER


class GmresInfo(ctypes.Structure):
    _fields_ = [('flag', c_int), ('iter_inner', c_int), ('iter_outer', c_int), ('total_iters', c_int),
                ('relres', c_dbl), ('resvec', P(c_dbl)), ('resvec_cap', c_int), ('resvec_len', c_int),
                ('ir_steps', c_int), ('seconds', c_dbl)]


class SolveParams(ctypes.Structure):
    _fields_ = [('tol', c_dbl), ('restart', c_int), ('inner_tol', c_dbl), ('inner_maxit', c_int),
                ('max_outer', c_int), ('side', c_int)]


# name -> (restype, argtypes)
SIGNATURES = {
    'vp_last_error': (ctypes.c_char_p, []),
    'vp_version': (ctypes.c_char_p, []),
    'vp_device_count': (c_int, [P(c_int)]),
    'vp_ctx_create': (c_int, [c_int, P(c_vp)]),
    'vp_ctx_destroy': (c_int, [c_vp]),
    'vp_ctx_synchronize': (c_int, [c_vp]),
    'vp_ctx_set_profiling': (c_int, [c_vp, c_int]),
    'vp_ctx_set_recycle_width': (c_int, [c_vp, c_int]),
    'vp_ctx_kernel_stats': (c_int, [c_vp, ctypes.c_char_p, P(c_i64), P(c_dbl), c_int]),
    'vp_mem_info': (c_int, [c_vp, P(c_i64), P(c_i64)]),
    'vp_malloc': (c_int, [c_vp, c_i64, P(c_vp)]),
    'vp_free': (c_int, [c_vp, c_vp]),
    'vp_memcpy_h2d': (c_int, [c_vp, c_vp, c_vp, c_i64]),
    'vp_memcpy_d2h': (c_int, [c_vp, c_vp, c_vp, c_i64]),
    'vp_op_create': (c_int, [c_vp, c_i64, c_i64, c_i64, c_i64, c_i64, c_i64, c_vp, c_vp, c_vp, c_dbl, c_vp,
                             c_i64, c_int, P(c_vp)]),
    'vp_op_create_toeplitz': (c_int, [c_vp, c_i64, c_i64, c_i64, c_i64, c_i64, c_i64, c_vp, c_vp, c_vp, c_dbl,
                                      c_vp, c_i64, c_int, P(c_vp)]),
    'vp_op_create_assembled': (c_int, [c_vp, c_i64, c_i64, c_i64, c_i64, c_i64, c_i64, c_dbl, c_dbl, c_vp, c_vp,
                                       c_vp, c_i64, c_int, P(c_vp)]),
    'vp_assemble_toeplitz': (c_int, [c_vp, c_i64, c_i64, c_i64, c_dbl, c_dbl, c_vp, c_int]),
    'vp_op_apply': (c_int, [c_vp, c_int, c_vp, c_vp, c_int, c_int]),
    'vp_op_apply_batched': (c_int, [c_vp, c_int, c_int, c_vp, c_vp, c_int]),
    'vp_op_fp32_scale': (c_int, [c_vp, P(c_dbl)]),
    'vp_op_size': (c_int, [c_vp, P(c_i64)]),
    'vp_op_info': (c_int, [c_vp, P(c_i64)]),
    'vp_op_destroy': (c_int, [c_vp]),
    'vp_prec_create_chan2': (c_int, [c_vp, c_i64, c_i64, c_i64, c_vp, c_vp, c_i64, c_int, P(c_vp)]),
    'vp_prec_create_chan1': (c_int, [c_vp, c_i64, c_i64, c_i64, c_vp, c_vp, c_i64, c_int, P(c_vp)]),
    'vp_prec_create_level2': (c_int, [c_vp, c_i64, c_i64, c_i64, c_vp, c_dbl, c_dbl, c_dbl, c_dbl, c_vp, c_i64,
                                      c_int, P(c_vp)]),
    'vp_prec_create_level2_assembled': (c_int, [c_vp, c_i64, c_i64, c_i64, c_dbl, c_dbl, c_dbl, c_dbl, c_dbl, c_dbl,
                                                c_vp, c_i64, c_int, P(c_vp)]),
    'vp_prec_create_level1': (c_int, [c_vp, c_i64, c_i64, c_i64, c_vp, c_dbl, c_dbl, c_vp, c_dbl, c_dbl, c_vp, c_i64,
                                      c_int, c_vp]),
    'vp_prec_create_level1_assembled': (c_int, [c_vp, c_i64, c_i64, c_i64, c_dbl, c_dbl, c_dbl, c_dbl, c_dbl, c_vp,
                                                c_dbl, c_vp, c_i64, c_int, c_vp]),
    'vp_prec_create_level1_face': (c_int, [c_vp, c_i64, c_i64, c_i64, c_dbl, c_dbl, c_dbl, c_vp, c_vp, c_i64, c_int,
                                           c_vp]),
    'vp_prec_create_level2_field': (c_int, [c_vp, c_i64, c_i64, c_i64, c_vp, c_vp, c_dbl, c_dbl, c_vp, c_i64, c_int,
                                            c_vp]),
    'vp_prec_create_level2_assembled_field': (c_int, [c_vp, c_i64, c_i64, c_i64, c_dbl, c_dbl, c_dbl, c_vp, c_dbl,
                                                      c_vp, c_i64, c_int, c_vp]),
    'vp_prec_create_level2_face': (c_int, [c_vp, c_i64, c_i64, c_i64, c_dbl, c_dbl, c_dbl, c_vp, c_vp, c_i64, c_int,
                                           P(c_vp)]),
    'vp_prec_set_embedding': (c_int, [c_vp, c_i64, c_vp, c_i64]),
    'vp_prec_create_composite': (c_int, [c_vp, c_vp, c_int, P(c_vp)]),
    'vp_prec_apply': (c_int, [c_vp, c_vp, c_vp, c_int, c_int]),
    'vp_prec_set_fp32_scale': (c_int, [c_vp, c_dbl]),
    'vp_prec_destroy': (c_int, [c_vp]),
    'vp_gmres': (c_int, [c_vp, c_vp, c_int, c_vp, c_vp, c_vp, c_int, c_int, c_int, c_dbl, c_int, P(GmresInfo)]),
    'vp_bicgstab': (c_int, [c_vp, c_vp, c_vp, c_vp, c_vp, c_int, c_int, c_dbl, c_int, P(GmresInfo)]),
    'vp_solve_ir': (c_int, [c_vp, c_vp, c_int, c_vp, c_vp, c_int, c_dbl, c_int, c_dbl, c_int, c_int,
                            P(GmresInfo)]),
    'vp_solve_ir_batched': (c_int, [c_vp, c_vp, c_int, c_int, c_vp, c_vp, c_dbl, c_int, c_dbl, c_int, c_int,
                                    P(GmresInfo)]),
    'vp_solve_fgmres_ir': (c_int, [c_vp, c_vp, c_int, c_vp, c_vp, c_int, c_dbl, c_int, c_dbl, c_int, c_int, c_int,
                                   P(GmresInfo)]),
    'vp_solve_fgmres_ir_batched': (c_int, [c_vp, c_vp, c_int, c_int, c_vp, c_vp, c_dbl, c_int, c_dbl, c_int, c_int,
                                           c_int, P(GmresInfo)]),
    'vp_server_create': (c_int, [c_vp, c_i64, c_int, P(c_vp)]),
    'vp_server_submit': (c_int, [c_vp, c_vp, c_vp, c_vp, c_vp, P(SolveParams), P(c_i64)]),
    'vp_server_poll': (c_int, [c_vp, c_i64, P(c_int), P(GmresInfo)]),
    'vp_server_wait': (c_int, [c_vp, c_i64, P(GmresInfo)]),
    'vp_server_stats': (c_int, [c_vp, P(c_i64), P(c_i64), P(c_i64), P(c_i64)]),
    'vp_server_destroy': (c_int, [c_vp]),
    'vp_field_dgf': (c_int, [c_vp, c_vp, c_vp, c_i64, c_vp, c_i64, c_dbl, c_dbl, c_vp, c_int]),
    'vp_field_dgf_grid': (c_int, [c_vp, c_vp, c_vp, c_i64, c_vp, c_dbl, c_i64, c_i64, c_i64, c_dbl, c_dbl, c_vp,
                                  c_int]),
    'vp_field_total': (c_int, [c_vp, c_vp, c_vp, c_dbl, c_vp, c_int]),
    'vp_op_create_lincomb': (c_int, [c_vp, c_vp, c_int, P(c_vp)]),
    'vp_vec_dots': (c_int, [c_vp, c_vp, c_int, c_vp, c_int, c_i64, c_vp]),
    'vp_vec_dots_noalloc': (c_int, [c_vp, c_vp, c_int, c_vp, c_int, c_i64, c_vp]),
    'vp_vec_lincomb': (c_int, [c_vp, c_vp, c_int, c_vp, c_int, c_i64, c_vp]),
    'vp_op_lincomb_into': (c_int, [c_vp, c_vp, c_vp, c_int]),
    'vp_vec_absmax': (c_int, [c_vp, c_vp, c_int, c_i64, c_vp]),
    'vp_vec_gather': (c_int, [c_vp, c_vp, c_i64, c_vp, c_vp, c_dbl, c_dbl, c_i64, c_vp]),
    'vp_vec_scatter': (c_int, [c_vp, c_vp, c_vp, c_i64, c_vp, c_i64]),
}

missing = []
for _name, (_res, _args) in SIGNATURES.items():
    try:
        _f = getattr(lib, _name)
    except AttributeError:
        missing.append(_name)
        continue
    _f.restype = _res
    _f.argtypes = _args


class VoxPhotonError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f'[{code}] {msg}')
        self.code = code


def check(rc):
    if rc != VP_OK:
        raise VoxPhotonError(rc, lib.vp_last_error().decode(errors='replace'))
    return rc


def ptr(a):
    """Address of a C-contiguous numpy array (no copy)."""
    if a is None:
        return None
    assert a.flags['C_CONTIGUOUS']
    return ctypes.c_void_p(a.ctypes.data)


def as_c128(a):
    return np.ascontiguousarray(a, dtype=np.complex128)


def as_idx(idx):
    return np.ascontiguousarray(np.asarray(idx).astype(np.int64))
