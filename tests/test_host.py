"""CPU: host-side logic around the solve path (workload generators, grid
padding, benchmark byte models, argument validation of the C ABI)."""
import ctypes

import numpy as np

from oracle import vie as ovie
from tests.helpers import rel
import pytest


def _prime_factors(n):
    out, p = set(), 2
    while p * p <= n:
        while n % p == 0:
            out.add(p)
            n //= p
        p += 1
    if n > 1:
        out.add(n)
    return out


def test_smooth_length_is_minimal_and_plannable():
    from voxphoton_amd.configs import smooth_length
    for n in list(range(1, 300)) + [1793, 2601, 4177]:
        m = smooth_length(n)
        assert m >= n and _prime_factors(m) <= {2, 3, 5, 7, 11}
        assert all(not (_prime_factors(k) <= {2, 3, 5, 7, 11}) for k in range(n, m))


def test_smoothen_keeps_the_unknowns():
    from voxphoton_amd import configs
    from voxphoton_amd.configs import smooth_length
    for prob in (configs.bragg_grating(periods=23, dx=20e-9, M=50, N=22), configs.disk_resonator(2.2e-6),
                 configs.straight_waveguide(L=137)):
        q = configs.smoothen(prob)
        assert all(smooth_length(v) == v for v in q.shape)
        assert q.mask.sum() == prob.mask.sum()
        assert np.array_equal(q.mask[:prob.L, :prob.M, :prob.N], prob.mask)
        assert np.array_equal(q.eps[:prob.L, :prob.M, :prob.N][prob.mask], prob.eps[prob.mask])
        for sl, ext in zip(q.box, q.shape):
            assert 0 <= sl.start < sl.stop <= ext and smooth_length(sl.stop - sl.start) == sl.stop - sl.start
        # the box still holds every unknown voxel
        sx, sy, sz = q.box
        assert q.mask[sx, sy, sz].sum() == q.mask.sum()


def test_disk_volume_matches_cylinder():
    from voxphoton_amd import configs
    r, dx, t = 2.5e-6, 20e-9, 220e-9
    p = configs.disk_resonator(r, dx=dx, thick=t)
    expected = np.pi * (r / dx) ** 2 * round(t / dx)
    assert abs(p.mask.sum() - expected) / expected < 0.01


def test_serving_stream_is_seeded_and_mixed():
    import bench
    wl = bench.WORKLOADS['serve_mix']
    a, b = bench.serve_requests(wl), bench.serve_requests(wl)
    assert a == b and len(a) == wl['nreq']
    t = [r[0] for r in a]
    assert all(x < y for x, y in zip(t, t[1:]))
    kinds = [r[1] for r in a]
    assert {'waveguide', 'bragg', 'disk'} == set(kinds)
    for _, kind, size, lam in a:
        assert 1500e-9 <= lam <= 1600e-9
        if kind == 'waveguide':
            assert 2e-6 <= size <= 8e-6
        elif kind == 'bragg':
            assert 20 <= size <= 150
        else:
            assert 2e-6 <= size <= 5e-6


def test_byte_models():
    import bench
    from voxphoton_amd import configs
    prob = configs.config_problem('bragg_10nm')
    # canonical model of the survey at the Bragg 10 nm workload
    assert abs(bench.mv_model_bytes(prob, 16) - 18.91e9) / 18.91e9 < 2e-3
    assert bench.unknown_box(prob) == (3600, 60, 22)
    # split passes: tensor octant plus y-spectra read and written on the box
    z = bench.kernel_model_bytes(prob, 'zpass', 8)
    Pc = 3601 * 101 * 45
    assert z == 8 * (6 * Pc + 6 * 7200 * 200 * 22)
    # quadrant blocks of the face-averaged preconditioner, binary16 in FP32 applies
    pb = bench.kernel_model_bytes(prob, 'pc_block', 8, pc_box=(3600, 22, 60))
    assert pb == 4 * 1801 * 12 * 180 ** 2 + 8 * 2 * 3 * 3600 * 22 * 60
    # the same pass at sizes embedded on the unknowns' box (7200 x 120 x 48)
    zb = bench.kernel_model_bytes(prob, 'zpass', 8, fft=(7200, 120, 48))
    assert zb == 8 * (6 * 3601 * 61 * 25 + 6 * 7200 * 120 * 22)


def test_server_arguments_fail_cleanly():
    from voxphoton_amd import _lib
    h = ctypes.c_void_p()
    assert _lib.lib.vp_server_create(None, 1 << 30, 4, ctypes.byref(h)) == _lib.VP_ERR_INVALID
    t = ctypes.c_int64()
    prm = _lib.SolveParams(1e-6, 50, 1e-3, 4, 10, _lib.VP_PREC_RIGHT)
    assert _lib.lib.vp_server_submit(None, None, None, None, None, ctypes.byref(prm), ctypes.byref(t)) \
        == _lib.VP_ERR_INVALID
    assert _lib.lib.vp_server_destroy(None) == _lib.VP_OK
    assert _lib.lib.vp_last_error() is not None


@pytest.mark.parametrize('kind,size', [('waveguide', 3e-6), ('bragg', 40), ('disk', 2.1e-6)])
def test_serving_problems_plan(kind, size):
    # every generated request has FFT and preconditioner-box lengths the device plans
    import bench
    from voxphoton_amd import configs
    prob, pc = bench.serve_problem(kind, size, 1550e-9, configs)
    for v in prob.shape:
        assert _prime_factors(2 * v) <= {2, 3, 5, 7, 11}
    for sl in prob.box:
        assert _prime_factors(sl.stop - sl.start) <= {2, 3, 5, 7, 11}
    assert pc['prec'] in ('face', 'scalar')


def test_config_rhs_matches_oracle_rhs():
    # the configurations' JVIE-II right-hand side equals the oracle's
    # Gram * j w eps0 Mc .* Einc(idx) (JVIE_Solver_const.m:151-157, :243-251)
    from voxphoton_amd import configs
    prob = configs.straight_waveguide()
    for pol in ('y', 'z'):
        b = prob.gaussian_rhs(pol)
        ref = ovie.jvie2_rhs(prob.gaussian_field(pol), prob.Mc, prob.gram, prob.freq, prob.idx)
        assert rel(b, ref) < 1e-15
