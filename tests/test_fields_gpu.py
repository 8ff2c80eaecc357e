"""GPU: excitation and post-processing through the C ABI against the field
oracle (source field of a complex-point-source beam, total field through
the L operator, transmitted power)."""
import numpy as np
import pytest

from oracle import fields as ofl, vie as ovie
from tests.helpers import random_toeplitz, rel

pytestmark = pytest.mark.gpu


def _beam():
    lam = 440e-9
    ko = 2 * np.pi / 1.55e-6
    cutoff = 0.5 * lam
    src = np.array([[0.25 * lam + 0.4e-6 - 1j * lam, 0.0, 0.0]])
    return lam, ko, cutoff, src


def test_dgf_points_match_oracle(vp_ctx):
    from voxphoton_amd.vie import E_field_DGF_mollified
    rng = np.random.default_rng(5)
    lam, ko, cutoff, src = _beam()
    obs = np.column_stack([0.4e-6 + rng.uniform(0, 0.3e-6, 500), rng.uniform(-0.2e-6, 0.2e-6, 500),
                           rng.uniform(-0.1e-6, 0.1e-6, 500)])
    for J, ic in (([0, 1, 0], src),
                  (rng.standard_normal(6) + 1j * rng.standard_normal(6),
                   np.vstack([src, src + np.array([[30e-9, -20e-9, 10e-9]])]))):
        E = E_field_DGF_mollified(J, ic, obs, ko, cutoff, ctx=vp_ctx)
        ref = ofl.E_field_DGF_mollified(J, ic, obs, ko, cutoff)
        print('dgf points', np.count_nonzero(ref[:, 1]), 'nonzero, rel', rel(E, ref), flush=True)
        assert np.count_nonzero(ref[:, 1]) > 50
        assert rel(E, ref) < 1e-12


def test_dgf_grid_matches_oracle(vp_ctx):
    from voxphoton_amd.vie import E_field_DGF_grid
    lam, ko, cutoff, src = _beam()
    dx = 20e-9
    shape = (40, 12, 7)
    origin = (0.3e-6, -5.5 * dx, -3 * dx)
    E = E_field_DGF_grid([0, 1, 0], src, origin, dx, shape, ko, cutoff, ctx=vp_ctx)
    pts = ofl.grid_coords(origin, dx, *shape)
    ref = ofl.E_field_DGF_mollified([0, 1, 0], src, pts, ko, cutoff).reshape(*shape, 3, order='F')
    assert E.shape == shape + (3,)
    assert rel(E, ref) < 1e-12


@pytest.mark.parametrize('shape', [(6, 5, 4), (9, 4, 3)])
def test_total_field_matches_oracle(vp_ctx, shape):
    from voxphoton_amd.vie import NOperator, E_field_Nop_const, power_through
    rng = np.random.default_rng(sum(shape))
    G = random_toeplitz(*shape, rng)
    Gram, freq = 0.41, 1.93e14
    op = NOperator.for_fields(shape, Gram=Gram, G=G, ctx=vp_ctx)
    J = rng.standard_normal(shape + (3,)) + 1j * rng.standard_normal(shape + (3,))
    Einc = rng.standard_normal(shape + (3,)) + 1j * rng.standard_normal(shape + (3,))
    E = E_field_Nop_const(J, op, freq, Einc)
    ref = ofl.E_field_Nop_const(J, ovie.fN_from_toeplitz(G), Gram, freq, Einc)
    print('total field rel', rel(E, ref), flush=True)
    assert rel(E, ref) < 1e-12
    assert np.array_equal(E_field_Nop_const(0 * J, op, freq, Einc), Einc)
    assert abs(power_through(E, 3, 1) - ofl.power_through(ref, 3, 1)) <= 1e-12 * ofl.power_through(ref, 3, 1)


def test_total_field_rejects_solve_operators(vp_ctx):
    from voxphoton_amd._lib import VoxPhotonError
    from voxphoton_amd.vie import NOperator, E_field_Nop_const
    from tests.helpers import random_problem
    G, Mr, Mc, Gram, idx, x = random_problem(5, 4, 3, seed=2)
    J = np.ones((5, 4, 3, 3), complex)
    # every voxel but a non-unit contrast: refused by the C ABI
    all_idx = np.arange(1, 3 * 60 + 1)
    op = NOperator.from_toeplitz(G, None, 0.5 + 0 * Mc, Gram, all_idx, ctx=vp_ctx)
    with pytest.raises(VoxPhotonError):
        E_field_Nop_const(J, op, 1.9e14)
    # unit contrast on a subset of voxels: the sizes do not match the field
    op2 = NOperator.from_toeplitz(G, None, 1 + 0 * Mc, Gram, idx, ctx=vp_ctx)
    with pytest.raises(ValueError):
        E_field_Nop_const(J, op2, 1.9e14)
