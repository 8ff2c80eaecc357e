"""Synthetic silicon-photonics workloads for the solve path.

Each builder returns a ``Problem``: voxel grid, unknown mask (core voxels;
background voxels have Mc = 0 and are not unknowns, as in
JVIE_Solver_const.m:148-149), JVIE-II contrast Mc = (er-1)/er with
er = eps_core / eps_background (Si_in_SiO2_parameters.m / Bragg_params*:
permittivity relative to the exterior medium), wavenumber in the background
ko = 2*pi*n_ext/lambda, frequency freq = c*n_ext/lambda (EMconstants.m), the
right-hand side Vrhs = Gram * j*omega*eps0 * Mc .* Einc(idx)
(JVIE_Solver_const.m:151-157 with the JVIE-II scaling of :243-251), and the
core bounding box on which the 2-level circulant preconditioner is built
(resonator_MOR_scripts.m:45-61 builds preconditioners on sub-boxes).
"""
from dataclasses import dataclass, field

import numpy as np

C0 = 299792458.0
MU0 = 4e-7 * np.pi
EPS0 = 1.0 / (C0 ** 2 * MU0)
EPS_SI = 12.11
EPS_SIO2 = 2.085


@dataclass
class Problem:
    name: str
    L: int
    M: int
    N: int
    dx: float
    lam: float
    mask: np.ndarray            # (L, M, N) bool, unknown voxels
    eps: np.ndarray             # (L, M, N) complex relative permittivity of the core (1 outside)
    n_ext: float = float(np.sqrt(EPS_SIO2))
    box: tuple = None           # (slice_x, slice_y, slice_z) preconditioner box
    extra: dict = field(default_factory=dict)

    @property
    def ko(self):
        return 2 * np.pi * self.n_ext / self.lam

    @property
    def freq(self):
        return C0 * self.n_ext / self.lam

    @property
    def omega(self):
        return 2 * np.pi * self.freq

    @property
    def gram(self):
        return self.dx ** 3

    @property
    def shape(self):
        return (self.L, self.M, self.N)

    @property
    def Mc(self):
        er = self.eps / EPS_SIO2
        mc = (er - 1.0) / er
        return np.where(self.mask, mc, 0.0).astype(np.complex128)

    @property
    def idx(self):
        return idx_from_mask(self.mask)

    @property
    def n(self):
        return 3 * int(self.mask.sum())

    def core_contrast(self):
        er = EPS_SI / EPS_SIO2
        return (er - 1.0) / er

    def gaussian_field(self, pol='y', w0=300e-9):
        """Incident field (L x M x N x 3) of an x-propagating Gaussian-profile
        wave polarised along ``pol``, at the voxel centres (l + 1/2) dx."""
        dx = self.dx
        y = (np.arange(self.M) + 0.5) * dx
        z = (np.arange(self.N) + 0.5) * dx
        x = (np.arange(self.L) + 0.5) * dx
        # centred in the grid, or on the excited guide (extra['rhs_center'], in voxels)
        yc, zc = y.mean(), z.mean()
        if 'rhs_center' in self.extra:
            yc, zc = ((np.asarray(self.extra['rhs_center'], float) + 0.5) * dx)
        prof = np.exp(-(((y[None, :, None] - yc) ** 2) + ((z[None, None, :] - zc) ** 2)) / w0 ** 2)
        E = prof * np.exp(-1j * self.ko * x)[:, None, None]
        comp = {'x': 0, 'y': 1, 'z': 2}[pol]
        Einc = np.zeros((self.L, self.M, self.N, 3), dtype=np.complex128)
        Einc[..., comp] = E
        return Einc

    def gaussian_rhs(self, pol='y', w0=300e-9):
        """Vrhs for an x-propagating Gaussian-profile wave polarised along ``pol``."""
        Einc = self.gaussian_field(pol, w0)
        tau = 1j * self.omega * EPS0 * self.Mc
        tau3 = np.concatenate([tau.reshape(-1, order='F')] * 3)
        i0 = self.idx - 1
        return self.gram * tau3[i0] * Einc.reshape(-1, order='F')[i0]

    def box_problem(self, box=None):
        """Preconditioner box (default ``self.box``): (shape, box idx
        (1-based), sub_of_op (1-based))."""
        sx, sy, sz = box if box is not None else self.box
        bmask = self.mask[sx, sy, sz]
        bidx = idx_from_mask(bmask)
        # map box unknowns to operator unknowns
        Lb, Mb, Nb = bmask.shape
        full_vox = np.flatnonzero(self.mask.reshape(-1, order='F'))
        nS = full_vox.size
        bl, bm, bn = np.unravel_index(np.flatnonzero(bmask.reshape(-1, order='F')), bmask.shape, order='F')
        gvox = np.ravel_multi_index((bl + sx.start, bm + sy.start, bn + sz.start), self.mask.shape, order='F')
        pos = np.searchsorted(full_vox, gvox)
        assert np.array_equal(full_vox[pos], gvox)
        sub = np.concatenate([pos, pos + nS, pos + 2 * nS]) + 1
        return (Lb, Mb, Nb), bidx, sub.astype(np.int64)


def face_box(prob, exact_axis=1, box=None):
    """Face-averaged preconditioner box (default ``prob.box``) with
    ``exact_axis`` (1 = y, 2 = z) last.

    Returns (shape, idx, sub_of_op, face): the box permuted so that the
    exact axis is the third one, its unknown indices, the 1-based positions
    of those unknowns in the operator's vector (components permuted with the
    axes), and face = (e - 1)/e of the permittivity averaged over the two
    circulant axes (real part, as Bragg_MOR_script.m:64-65 takes it).
    """
    sx, sy, sz = box if box is not None else prob.box
    bmask = prob.mask[sx, sy, sz]
    beps = (prob.eps / EPS_SIO2)[sx, sy, sz]
    beps = np.where(bmask, beps, 1.0)
    if exact_axis == 1:
        axes = (0, 2, 1)
        cperm = [0, 2, 1]
    elif exact_axis == 2:
        axes = (0, 1, 2)
        cperm = [0, 1, 2]
    else:
        raise ValueError('exact_axis must be 1 or 2')
    pmask = np.transpose(bmask, axes)
    peps = np.transpose(beps, axes)
    ebar = np.real(peps.mean(axis=(0, 1)))
    face = ((ebar - 1.0) / ebar).astype(np.complex128)
    pidx = idx_from_mask(pmask)
    # permuted-box voxel -> full-grid voxel -> operator unknown position
    full_vox = np.flatnonzero(prob.mask.reshape(-1, order='F'))
    nS = full_vox.size
    pl, pa, pb = np.unravel_index(np.flatnonzero(pmask.reshape(-1, order='F')), pmask.shape, order='F')
    orig = [None, None, None]
    orig[0] = pl
    inv = np.argsort(axes)
    comps = [pl, pa, pb]
    orig = [comps[inv[0]], comps[inv[1]], comps[inv[2]]]
    gvox = np.ravel_multi_index((orig[0] + sx.start, orig[1] + sy.start, orig[2] + sz.start),
                                prob.mask.shape, order='F')
    pos = np.searchsorted(full_vox, gvox)
    assert np.array_equal(full_vox[pos], gvox)
    sub = np.concatenate([pos + cperm[c] * nS for c in range(3)]) + 1
    return pmask.shape, pidx, sub.astype(np.int64), face


def idx_from_mask(mask):
    """[idxS; V+idxS; 2V+idxS], 1-based, idxS sorted (JVIE_Solver_const.m:148-149)."""
    V = mask.size
    s = np.flatnonzero(mask.reshape(-1, order='F')) + 1
    return np.concatenate([s, V + s, 2 * V + s]).astype(np.int64)


def _bbox(mask):
    nz = np.nonzero(mask)
    return tuple(slice(int(a.min()), int(a.max()) + 1) for a in nz)


def straight_waveguide(L=200, M=48, N=24, dx=20e-9, lam=1550e-9, width=500e-9, thick=220e-9):
    """Si strip waveguide in SiO2 along x (full length), centred in y and z.

    Defaults: the 200 x 48 x 24 grid at 20 nm with the 25 x 11 voxel core at
    y in [11, 36), z in [6, 17).
    """
    wy = int(round(width / dx))
    tz = int(round(thick / dx))
    y0 = (M - wy) // 2
    z0 = (N - tz) // 2
    mask = np.zeros((L, M, N), bool)
    mask[:, y0:y0 + wy, z0:z0 + tz] = True
    eps = np.where(mask, EPS_SI, 1.0).astype(np.complex128)
    p = Problem('straight_wg', L, M, N, dx, lam, mask, eps)
    p.box = _bbox(mask)
    return p


def bragg_grating(periods=100, period=320e-9, dx=10e-9, lam=1550e-9, width=500e-9, corr=50e-9, thick=220e-9,
                  len_abs=2e-6, abs_im=1.0, M=None, N=None):
    """Bragg grating with quadratic-ramp absorbers at both ends.

    Geometry follows Bragg_geometry.m:20-42: each period is a wide half
    (width + 2*corr) followed by a narrow half (width - 2*corr); straight
    sections of the nominal width carry the absorbers, whose permittivity
    ramps as eps_Si - j*abs_im*(d/len_abs)^2 with depth d.
    Defaults: 3600 x 100 x 44 voxels at 10 nm.
    """
    na = int(round(len_abs / dx))
    npd = int(round(period / dx))
    L = 2 * na + periods * npd
    M = M or int(round(2 * (width + 2 * corr) / dx)) - 20
    N = N or int(round(2 * thick / dx))
    wide = int(round((width + 2 * corr) / dx))
    narrow = int(round((width - 2 * corr) / dx))
    nom = int(round(width / dx))
    tz = int(round(thick / dx))
    z0 = (N - tz) // 2
    mask = np.zeros((L, M, N), bool)
    ylo = lambda w: (M - w) // 2      # noqa: E731
    for l in range(L):
        if l < na or l >= L - na:
            w = nom
        else:
            ph = (l - na) % npd
            w = wide if ph < npd // 2 else narrow
        mask[l, ylo(w):ylo(w) + w, z0:z0 + tz] = True
    eps = np.where(mask, EPS_SI, 1.0).astype(np.complex128)
    x = (np.arange(L) + 0.5) * dx
    d = np.maximum(np.maximum(len_abs - x, x - (L * dx - len_abs)), 0.0)
    ramp = abs_im * (d / len_abs) ** 2
    eps = np.where(mask, EPS_SI - 1j * ramp[:, None, None], eps)
    p = Problem('bragg', L, M, N, dx, lam, mask, eps)
    p.box = _bbox(mask)
    p.extra = dict(periods=periods, period=period, len_abs=len_abs)
    return p


def disk_resonator(radius=2e-6, dx=20e-9, lam=1550e-9, thick=220e-9, margin=2):
    """Si microdisk (220 nm thick) in SiO2: a cylinder of ``radius`` centred
    in an (2R/dx + 2*margin)^2 x (thick/dx + 2*margin) grid."""
    d = int(round(2 * radius / dx))
    tz = int(round(thick / dx))
    L = M = d + 2 * margin
    N = tz + 2 * margin
    c = (L - 1) / 2.0
    ii, jj = np.meshgrid(np.arange(L), np.arange(M), indexing='ij')
    disk = ((ii - c) ** 2 + (jj - c) ** 2) <= (radius / dx) ** 2
    mask = np.zeros((L, M, N), bool)
    mask[:, :, margin:margin + tz] = disk[:, :, None]
    eps = np.where(mask, EPS_SI, 1.0).astype(np.complex128)
    p = Problem('disk', L, M, N, dx, lam, mask, eps)
    p.box = _bbox(mask)
    p.extra = dict(radius=radius)
    return p


def microdisk_bus(radius=10e-6, gap=200e-9, width=500e-9, thick=220e-9, dx=15e-9, lam=1550e-9, L=1400, M=1500,
                  N=30, len_abs=2e-6, abs_im=1.0):
    """Si microdisk with a straight bus waveguide in SiO2 (the resonator of
    resonator_MOR_scripts.m).  The disk (``radius``, ``thick``) and the bus
    (``width`` wide, running along x over the whole grid, ``gap`` from the
    disk edge in +y) share the z rows of the core; the pair is centred in y
    and z.  Defaults: radius 10 um, gap 200 nm, 500 nm bus, 220 nm thick,
    15 nm voxels on 1400 x 1500 x 30.

    The bus carries quadratic-ramp absorbers at both ends, as the resonator
    script gives its waveguide (resonator_MOR_scripts.m:66-77,
    sigma0*fun_absorb on the waveguide voxels only): the bus permittivity is
    eps_Si - j*abs_im*(d/len_abs)^2 at depth d < len_abs from either end, the
    form of the grating's absorbers.  Without them the cut bus ends reflect
    and, with the lossless disk, form a coupled cavity.  ``len_abs = 0``
    leaves the bus lossless.

    ``extra['parts']`` holds the two preconditioner boxes of the composite
    (prec_fun_disk_idx.m): the bus box (all of x, the bus rows) and the disk's
    bounding box, which are disjoint; ``extra['part_forms']`` the form of
    each part's 2-level circulant: face-averaged and exact across the bus
    ('face_y', the waveguide form of Bragg_MOR_script.m:61-72) and the core
    contrast on the disk box ('core'); ``extra['rhs_center']`` centres the
    Gaussian excitation on the bus."""
    R = radius / dx
    wb = int(round(width / dx))
    gp = max(int(round(gap / dx)), 1)
    tz = int(round(thick / dx))
    d = int(round(2 * R))
    span = d + gp + wb
    if d + 2 > L or span + 2 > M or tz + 2 > N:
        raise ValueError('grid too small for the resonator')
    y0 = (M - span) // 2                  # first disk row
    yc = y0 + (d - 1) / 2.0
    xc = (L - 1) / 2.0
    z0 = (N - tz) // 2
    ii, jj = np.meshgrid(np.arange(L), np.arange(M), indexing='ij')
    disk = ((ii - xc) ** 2 + (jj - yc) ** 2) <= R ** 2
    yb = y0 + d + gp                       # first bus row
    mask2 = disk.copy()
    mask2[:, yb:yb + wb] = True
    mask = np.zeros((L, M, N), bool)
    mask[:, :, z0:z0 + tz] = mask2[:, :, None]
    eps = np.where(mask, EPS_SI, 1.0).astype(np.complex128)
    if len_abs > 0:
        x = (np.arange(L) + 0.5) * dx
        depth = np.maximum(np.maximum(len_abs - x, x - (L * dx - len_abs)), 0.0)
        ramp = abs_im * (depth / len_abs) ** 2
        bus = np.zeros((L, M, N), bool)
        bus[:, yb:yb + wb, z0:z0 + tz] = True
        eps = np.where(bus, EPS_SI - 1j * ramp[:, None, None], eps)
    p = Problem('microdisk_bus', L, M, N, dx, lam, mask, eps)
    p.box = _bbox(mask)
    dbox = _bbox(np.where(np.arange(M)[None, :, None] < yb, mask, False))
    bbox = (slice(0, L), slice(yb, yb + wb), slice(z0, z0 + tz))
    # preconditioner boxes of lengths the device FFTs plan, widened inside
    # the grid without overlapping each other
    bbox = _smooth_box(bbox, p.shape, [dbox])
    dbox = _smooth_box(dbox, p.shape, [bbox])
    p.extra = dict(radius=radius, gap=gap, width=width, len_abs=len_abs, parts=[bbox, dbox],
                   part_forms=('face_y', 'core'),
                   rhs_center=(yb + (wb - 1) / 2.0, z0 + (tz - 1) / 2.0))
    return p


def _overlap(a, b):
    return all(x.start < y.stop and y.start < x.stop for x, y in zip(a, b))


def _smooth_box(box, shape, avoid):
    """``box`` widened per axis to a length with small prime factors (up or
    down), inside the grid and disjoint from every box in ``avoid``."""
    out = list(box)
    for ax, (sl, ext) in enumerate(zip(box, shape)):
        w = min(smooth_length(sl.stop - sl.start), ext)
        for start in (sl.start, sl.stop - w, ext - w, 0):
            cand = list(out)
            cand[ax] = slice(max(0, min(start, ext - w)), max(0, min(start, ext - w)) + w)
            if not any(_overlap(cand, a) for a in avoid):
                out = cand
                break
    return tuple(out)


def smooth_length(n, primes=(2, 3, 5, 7, 11)):
    """Smallest m >= n whose prime factors are all in ``primes`` (lengths the
    device FFT plans in every precision)."""
    m = max(int(n), 1)
    while True:
        r = m
        for q in primes:
            while r % q == 0:
                r //= q
        if r == 1:
            return m
        m += 1


def smoothen(prob):
    """Pads the grid with empty voxels (high end of every axis) to lengths
    with small prime factors, so that the 2x FFT sizes plan, and widens the
    preconditioner box likewise inside the padded grid.  The unknowns and
    their values are unchanged."""
    L, M, N = (smooth_length(v) for v in prob.shape)
    if (L, M, N) != prob.shape:
        mask = np.zeros((L, M, N), bool)
        eps = np.ones((L, M, N), np.complex128)
        mask[:prob.L, :prob.M, :prob.N] = prob.mask
        eps[:prob.L, :prob.M, :prob.N] = prob.eps
        prob = Problem(prob.name, L, M, N, prob.dx, prob.lam, mask, eps, prob.n_ext, prob.box, dict(prob.extra))
    box = []
    for sl, ext in zip(prob.box, prob.shape):
        w = smooth_length(sl.stop - sl.start)
        start = min(sl.start, ext - w)
        box.append(slice(start, start + w))
    prob.box = tuple(box)
    return prob


def config_problem(name):
    """The workloads named by the benchmark configurations."""
    if name == 'straight_wg_20nm':
        return straight_waveguide()
    if name == 'bragg_10nm':
        return bragg_grating()
    if name == 'bragg_20nm':
        return bragg_grating(dx=20e-9, M=50, N=22)
    if name == 'microdisk_15nm':
        return microdisk_bus()
    raise KeyError(name)
