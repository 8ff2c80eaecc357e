"""Green-tensor entries by the reference's own rules (test infrastructure only).

The face-pair integrals come from the reference's C++ cores, compiled from
the reference sources by ``oracle/build_ref.sh`` into ``oracle/_ref/libvpref.so``:
the 4D tensor-Gauss cubature of ``cubature_nop_const.cpp:8-125`` and the
DEMCEM singular triangle-pair integrals of ``direct_ws_st_const.cpp:53-182``,
``direct_ws_ea_const.cpp:58-460`` and ``direct_ws_va_const.cpp:62-237``.
This module restates the MATLAB orchestration around them:

* ``gauss_2d`` (``cubatures/gauss_2d.m:1-19``) and ``cube_faces``
  (``cube_faces.m:1-41``);
* the square-pair singular values of ``cubatures/singular_st_const.m``,
  ``singular_ea_const.m`` and ``singular_va_const.m`` (each square split in
  two triangles, four triangle pairs summed);
* the rules of ``assembly_nop_const.m:16-168``: a 4-point rule everywhere,
  10 points for offsets below ``n_medium = 5``, 20 points for the cells
  touching the source cell (offsets 0 or 1), whose singular face pairs take
  the DEMCEM values (``nearby_correct.m``: J = P .* I + C);
* ``G_k(m) = 1/(4 pi) sum_pq J_pq(m) P(p, q, k)`` with ``vector_projection.m``.

``nearby_correct.m`` lists the singular pairs per touching cell as tables;
here each face pair is classified from its geometry instead (identical
squares: ST; a shared edge, coplanar or orthogonal: EAc / EAo; a shared
vertex only: VAc / VAo), which gives the same substitution.
"""
import ctypes
import os

import numpy as np

from .greens import vector_projection

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, '_ref', 'libvpref.so')
N_MEDIUM = 5
NP_DEMCEM, NP_NEAR, NP_MEDIUM, NP_FAR = 20, 20, 10, 4    # assembly_nop_const.m:16-19

_lib = None


def available():
    return os.path.exists(LIB_PATH)


def lib():
    global _lib
    if _lib is None:
        if not available():
            raise ImportError(f'{LIB_PATH} is not built (oracle/build_ref.sh)')
        _lib = ctypes.CDLL(LIB_PATH)
    return _lib


def _p(a):
    return np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def gauss_1d(n):
    """(w, x) of the n-point Gauss-Legendre rule on [-1, 1]."""
    x, w = np.polynomial.legendre.leggauss(n)
    return w, x


def gauss_2d(n):
    """gauss_2d.m: n^2 points on the unit square, weights summing to one."""
    w1, x1 = gauss_1d(n)
    W = np.outer(w1, w1).reshape(-1) / 4.0
    U = np.repeat((x1 + 1) / 2, n)
    V = np.tile((x1 + 1) / 2, n)
    return n * n, W, U, V


def cube_faces(dx):
    """R_faces[coord, corner, face] of the cube [0, dx]^3 (cube_faces.m)."""
    c = {
        0: [(0, 0, 0), (0, 1, 0), (0, 1, 1), (0, 0, 1)],
        1: [(1, 0, 0), (1, 1, 0), (1, 1, 1), (1, 0, 1)],
        2: [(0, 0, 0), (1, 0, 0), (1, 0, 1), (0, 0, 1)],
        3: [(0, 1, 0), (1, 1, 0), (1, 1, 1), (0, 1, 1)],
        4: [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0)],
        5: [(0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)],
    }
    R = np.zeros((3, 4, 6))
    for f, corners in c.items():
        for k, pt in enumerate(corners):
            R[:, k, f] = np.array(pt, float) * dx
    return R


def cubature_nop(n1d, m, ko, dx):
    """6 x 6 J_pq(m) by the reference's tensor cubature (n1d points per axis)."""
    Np, W, U, V = gauss_2d(n1d)
    r_m = np.array(m, float) * dx
    r_n = np.zeros(3)
    out = np.zeros(72)
    lib().vpref_cubature_nop(ctypes.c_int(Np), _p(W), _p(U), _p(V), _p(r_m), _p(r_n), _p(cube_faces(dx)),
                             ctypes.c_double(ko), ctypes.c_double(dx), _p_out(out))
    return (out[0::2] + 1j * out[1::2]).reshape(6, 6)


def _p_out(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _st(r1, r2, r3, ko, n):
    w, z = gauss_1d(n)
    out = np.zeros(2)
    lib().vpref_direct_ws_st(_p(r1), _p(r2), _p(r3), ctypes.c_double(ko), ctypes.c_int(n), _p(z), _p(w),
                             _p_out(out))
    return out[0] + 1j * out[1]


def _ea(r1, r2, r3, r4, ko, n):
    w, z = gauss_1d(n)
    out = np.zeros(2)
    lib().vpref_direct_ws_ea(_p(r1), _p(r2), _p(r3), _p(r4), ctypes.c_double(ko), ctypes.c_int(n), ctypes.c_int(n),
                             _p(w), _p(z), _p(w), _p(z), _p_out(out))
    return out[0] + 1j * out[1]


def _va(r1, r2, r3, r4, r5, ko, n):
    w, z = gauss_1d(n)
    out = np.zeros(2)
    lib().vpref_direct_ws_va(_p(r1), _p(r2), _p(r3), _p(r4), _p(r5), ctypes.c_double(ko), ctypes.c_int(n),
                             ctypes.c_int(n), ctypes.c_int(n), _p(w), _p(z), _p(w), _p(z), _p(w), _p(z),
                             _p_out(out))
    return out[0] + 1j * out[1]


def singular_values(ko, dx, n=NP_DEMCEM):
    """{'ST', 'EAc', 'EAo', 'VAc', 'VAo'}: square-pair integrals of the
    singular configurations (singular_st_const.m, singular_ea_const.m,
    singular_va_const.m), DEMCEM method."""
    d = dx
    r1, r2, r3, r4 = [np.array(v, float) * d for v in ((0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0))]
    out = {}
    out['ST'] = (_st(r1, r2, r3, ko, n) + _ea(r1, r3, r2, r4, ko, n) + _ea(r1, r3, r4, r2, ko, n) +
                 _st(r1, r3, r4, ko, n))
    r5, r6 = np.array((2, 0, 0), float) * d, np.array((2, 1, 0), float) * d
    r5o, r6o = np.array((1, 0, 1), float) * d, np.array((1, 1, 1), float) * d
    out['EAc'] = (_ea(r2, r3, r1, r5, ko, n) + _va(r3, r1, r2, r5, r6, ko, n) + _va(r3, r4, r1, r2, r5, ko, n) +
                  _va(r3, r4, r1, r5, r6, ko, n))
    out['EAo'] = (_ea(r2, r3, r1, r5o, ko, n) + _va(r3, r1, r2, r5o, r6o, ko, n) +
                  _va(r3, r4, r1, r2, r5o, ko, n) + _va(r3, r4, r1, r5o, r6o, ko, n))
    r5, r6, r7 = (np.array(v, float) * d for v in ((2, 1, 0), (2, 2, 0), (1, 2, 0)))
    r5o, r6o, r7o = (np.array(v, float) * d for v in ((1, 1, 1), (1, 2, 1), (1, 2, 0)))
    out['VAc'] = (_va(r3, r1, r2, r5, r6, ko, n) + _va(r3, r1, r2, r6, r7, ko, n) + _va(r3, r4, r1, r5, r6, ko, n) +
                  _va(r3, r4, r1, r6, r7, ko, n))
    out['VAo'] = (_va(r3, r1, r2, r5o, r6o, ko, n) + _va(r3, r1, r2, r6o, r7o, ko, n) +
                  _va(r3, r4, r1, r5o, r6o, ko, n) + _va(r3, r4, r1, r6o, r7o, ko, n))
    return out


def _face(m, p):
    """Face p (-x,+x,-y,+y,-z,+z) of the unit cell at integer offset m:
    (normal axis, coordinate, {axis: (lo, hi)})."""
    a, side = divmod(p, 2)
    rng = {b: (m[b], m[b] + 1) for b in range(3) if b != a}
    return a, m[a] + side, rng


def contact(m, p, q):
    """Singular type of face p of cell m against face q of cell 0, or None."""
    a1, c1, r1 = _face(m, p)
    a2, c2, r2 = _face((0, 0, 0), q)
    if a1 == a2:
        if c1 != c2:
            return None
        ov = [min(r1[b][1], r2[b][1]) - max(r1[b][0], r2[b][0]) for b in r1]
        if min(ov) < 0:
            return None
        return {2: 'ST', 1: 'EAc', 0: 'VAc'}[sum(ov)]
    # orthogonal: the line x_a1 = c1, x_a2 = c2 must meet both squares
    if not (r2[a1][0] <= c1 <= r2[a1][1] and r1[a2][0] <= c2 <= r1[a2][1]):
        return None
    t = 3 - a1 - a2
    ov = min(r1[t][1], r2[t][1]) - max(r1[t][0], r2[t][0])
    if ov < 0:
        return None
    return 'EAo' if ov == 1 else 'VAo'


def J_ref(m, ko, dx, sing=None):
    """6 x 6 J_pq(m) by the reference's rules (assembly_nop_const.m)."""
    m = tuple(int(v) for v in m)
    if max(m) <= 1:
        J = cubature_nop(NP_NEAR, m, ko, dx)
        sing = sing if sing is not None else singular_values(ko, dx)
        for p in range(6):
            for q in range(6):
                t = contact(m, p, q)
                if t is not None:
                    J[p, q] = sing[t]
        return J
    if max(m) < N_MEDIUM:
        return cubature_nop(NP_MEDIUM, m, ko, dx)
    return cubature_nop(NP_FAR, m, ko, dx)


def G_ref(offsets, ko, dx):
    """G_k(m) (6 components) at the given cell offsets by the reference's rules."""
    P = vector_projection()
    sing = singular_values(ko, dx)
    out = []
    for m in offsets:
        J = J_ref(m, ko, dx, sing)
        out.append(np.einsum('pq,pqk->k', J, P) / (4 * np.pi))
    return np.array(out)
