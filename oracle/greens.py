"""Green-tensor (opToeplitz) face integrals in FP64 (test infrastructure only).

Defines the quantity the reference assembles in
pc/src_vie/src_operators/assembly_nop_const.m:16-168:

    J_pq(m) = dx^4 * <<face p of cell m | e^{-j ko R}/R | face q of cell 0>>
    G_k(m)  = 1/(4 pi) * sum_{p,q} J_pq(m) * P(p, q, k)

with faces ordered -x,+x,-y,+y,-z,+z (cube_faces.m), unit-square
parameterisation (gauss_2d.m: weights sum to one, Jacobian dx^4 applied at
the end, cubature_nop_const.cpp:115-123) and the projection P of
vector_projection.m:7-22.

The cubature is this build's own (the reference uses 4/10/20-point tensor
Gauss rules plus DEMCEM singular corrections, which are upstream of the solve
path): every face pair integral reduces exactly, through the difference
variables of the two unit squares, to

  * parallel faces:      int_{[-1,1]^2} Lam(s1) Lam(s2) g(h, s1-m_b, s2-m_c)
  * perpendicular faces: int_{[0,1]^2 x [-1,1]} Lam(s) g(u'-m_a-s_p, s_q-m_a'-u, s-m_c)

with Lam(s) = 1-|s|.  The domains are split at the kinks of Lam; a 1/R
singularity can then only sit at a vertex of a sub-box, where a Duffy map
(triangles in 2D, pyramids in 3D) removes it.  The device assembly kernel
evaluates the same rules in the same order.
"""
import numpy as np

FACES = [(0, 0), (0, 1), (1, 0), (1, 1), (2, 0), (2, 1)]   # (axis, side)
COMP_AXES = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]

# Gauss orders per sub-box: singular (Duffy), touching cells, medium, far.
ORDER_SINGULAR = 12
ORDER_NEAR = 10
ORDER_MEDIUM = 6
ORDER_FAR = 3
N_MEDIUM = 5          # assembly_nop_const.m:22 (cells with all offsets < 5)


def gauss01(n):
    x, w = np.polynomial.legendre.leggauss(n)
    return 0.5 * (x + 1.0), 0.5 * w


def vector_projection():
    """P(p, q, k) = dot(n_p x e_a, n_q x e_b) (vector_projection.m:7-22)."""
    e = np.eye(3)
    normals = [-e[0], e[0], -e[1], e[1], -e[2], e[2]]
    P = np.zeros((6, 6, 6))
    for p in range(6):
        for q in range(6):
            for k, (a, b) in enumerate(COMP_AXES):
                P[p, q, k] = np.dot(np.cross(normals[p], e[a]), np.cross(normals[q], e[b]))
    return P


def _g(kappa, rho):
    return np.exp(-1j * kappa * rho) / rho


def _lam(s):
    return 1.0 - np.abs(s)


def _parallel(m, a, sp, sq, kappa, order):
    b, c = [ax for ax in range(3) if ax != a]
    h = sq - (m[a] + sp)
    mb, mc = m[b], m[c]
    total = 0.0 + 0.0j
    for al in (-1.0, 0.0):
        for be in (-1.0, 0.0):
            sing = (h == 0 and mb in (al, al + 1) and mc in (be, be + 1))
            if sing:
                v1, v2 = float(mb), float(mc)
                sg1 = 1.0 if v1 == al else -1.0
                sg2 = 1.0 if v2 == be else -1.0
                t, w = gauss01(ORDER_SINGULAR)
                XI, ETA = np.meshgrid(t, t, indexing='ij')
                W = np.outer(w, w)
                for t1, t2 in ((XI, XI * ETA), (XI * ETA, XI)):
                    s1 = v1 + sg1 * t1
                    s2 = v2 + sg2 * t2
                    rho = np.sqrt(t1 * t1 + t2 * t2)
                    f = _lam(s1) * _lam(s2) * np.exp(-1j * kappa * rho) * (XI / rho)
                    total += np.sum(W * f)
            else:
                t, w = gauss01(order)
                S1, S2 = np.meshgrid(al + t, be + t, indexing='ij')
                W = np.outer(w, w)
                rho = np.sqrt(h * h + (S1 - mb) ** 2 + (S2 - mc) ** 2)
                total += np.sum(W * _lam(S1) * _lam(S2) * _g(kappa, rho))
    return total


def _perpendicular(m, ap, sp, aq, sq, kappa, order):
    c = 3 - ap - aq
    ustar = m[ap] + sp            # u' (coordinate of face q along ap) at the singular point
    vstar = sq - m[aq]            # u  (coordinate of face p along aq)
    sstar = m[c]
    total = 0.0 + 0.0j
    for ga in (-1.0, 0.0):
        sing = (ustar in (0, 1) and vstar in (0, 1) and sstar in (ga, ga + 1))
        if sing:
            sgu = 1.0 if ustar == 0 else -1.0
            sgv = 1.0 if vstar == 0 else -1.0
            sgs = 1.0 if sstar == ga else -1.0
            t, w = gauss01(ORDER_SINGULAR)
            XI, ETA, ZETA = np.meshgrid(t, t, t, indexing='ij')
            W = w[:, None, None] * w[None, :, None] * w[None, None, :]
            for (tu, tv, ts) in ((XI, XI * ETA, XI * ZETA),
                                 (XI * ETA, XI, XI * ZETA),
                                 (XI * ETA, XI * ZETA, XI)):
                s = sstar + sgs * ts
                rho = np.sqrt(tu * tu + tv * tv + ts * ts)
                f = _lam(s) * np.exp(-1j * kappa * rho) * (XI * XI / rho)
                total += np.sum(W * f)
        else:
            t, w = gauss01(order)
            U, Vv, S = np.meshgrid(t, t, ga + t, indexing='ij')
            W = w[:, None, None] * w[None, :, None] * w[None, None, :]
            d1 = U - m[ap] - sp
            d2 = sq - m[aq] - Vv
            d3 = S - m[c]
            rho = np.sqrt(d1 * d1 + d2 * d2 + d3 * d3)
            total += np.sum(W * _lam(S) * _g(kappa, rho))
    return total


def rule_order(m):
    mx = max(m)
    if mx <= 1:
        return ORDER_NEAR
    if mx < N_MEDIUM:
        return ORDER_MEDIUM
    return ORDER_FAR


def face_pair_integrals(m, kappa):
    """6 x 6 dimensionless integrals <<face p of cell m | e^{-j kappa rho}/rho | face q of cell 0>>."""
    order = rule_order(m)
    I = np.zeros((6, 6), dtype=complex)
    for p, (ap, sp) in enumerate(FACES):
        for q, (aq, sq) in enumerate(FACES):
            if ap == aq:
                I[p, q] = _parallel(m, ap, sp, sq, kappa, order)
            else:
                I[p, q] = _perpendicular(m, ap, sp, aq, sq, kappa, order)
    return I


def assemble_G(L, M, N, dx, ko):
    """opToeplitz (L x M x N x 6) for a uniform grid (tiny grids only)."""
    P = vector_projection()
    kappa = ko * dx
    G = np.zeros((L, M, N, 6), dtype=complex)
    for i in range(L):
        for j in range(M):
            for k in range(N):
                I = face_pair_integrals((i, j, k), kappa)
                G[i, j, k] = np.einsum('pq,pqk->k', I, P)
    G *= dx ** 3 / (4.0 * np.pi)
    # Odd components vanish exactly on their symmetry planes.
    from .vie import PARITY
    for k in range(6):
        for ax in range(3):
            if PARITY[k, ax] < 0:
                sl = [slice(None)] * 3
                sl[ax] = 0
                G[tuple(sl) + (k,)] = 0.0
    return G
