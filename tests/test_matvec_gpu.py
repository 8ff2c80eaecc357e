"""GPU parity of the N-operator matvec (C ABI) against the FP64 oracle."""
import numpy as np
import pytest

from oracle import vie as ovie, greens as ogr
from tests.helpers import random_problem, rel

pytestmark = pytest.mark.gpu

SHAPES = [(5, 4, 3), (8, 6, 4), (12, 10, 7), (3, 1, 2), (1, 1, 1), (16, 9, 5), (20, 12, 6)]


@pytest.mark.parametrize('shape', SHAPES)
def test_matvec_fp64_from_fN(vp_ctx, shape):
    from voxphoton_amd.vie import NOperator
    G, Mr, Mc, Gram, idx, x = random_problem(*shape, seed=hash(shape) % 1000)
    fN = ovie.fN_from_toeplitz(G)
    ref = ovie.mv_AN_const(x, fN, Mr, Mc, Gram, 'notransp', idx)
    op = NOperator.from_fN(fN, None, Mc, Gram, idx, precisions=('fp64', 'fp32'), ctx=vp_ctx)
    y = op.apply(x, 'notransp', 'fp64')
    print(shape, 'fp64 rel', rel(y, ref), flush=True)
    assert rel(y, ref) < 1e-12
    y32 = op.apply(x, 'notransp', 'fp32')
    print(shape, 'fp32 rel', rel(y32, ref), flush=True)
    assert rel(y32, ref) < 1e-5


@pytest.mark.parametrize('shape', [(6, 5, 4), (10, 8, 6)])
def test_matvec_transp_and_mr(vp_ctx, shape):
    from voxphoton_amd.vie import NOperator
    G, Mr, Mc, Gram, idx, x = random_problem(*shape, seed=7, with_mr=True)
    fN = ovie.fN_from_toeplitz(G)
    op = NOperator.from_fN(fN, Mr, Mc, Gram, idx, ctx=vp_ctx)
    for flag in ('notransp', 'transp'):
        ref = ovie.mv_AN_const(x, fN, Mr, Mc, Gram, flag, idx)
        assert rel(op.apply(x, flag, 'fp64'), ref) < 1e-12
        assert rel(op.apply(x, flag, 'fp32'), ref) < 1e-5


def test_matvec_generic_index_order(vp_ctx):
    # idx not of the [idxS; V+idxS; 2V+idxS] form: per-entry map path
    from voxphoton_amd.vie import NOperator
    rng = np.random.default_rng(3)
    G, Mr, Mc, Gram, idx, x = random_problem(7, 5, 4, seed=3)
    idx2 = rng.permutation(idx)[: idx.size - 5]
    x2 = x[: idx2.size]
    fN = ovie.fN_from_toeplitz(G)
    ref = ovie.mv_AN_const(x2, fN, Mr, Mc, Gram, 'notransp', idx2)
    op = NOperator.from_fN(fN, None, Mc, Gram, idx2, ctx=vp_ctx)
    assert rel(op.apply(x2), ref) < 1e-12


@pytest.mark.parametrize('shape,fft', [((6, 5, 4), (0, 0, 0)), ((7, 5, 3), (16, 10, 6)), ((9, 4, 5), (21, 9, 9)),
                                       ((9, 4, 5), (18, 8, 12))])
def test_matvec_from_toeplitz(vp_ctx, shape, fft):
    from voxphoton_amd.vie import NOperator
    G, Mr, Mc, Gram, idx, x = random_problem(*shape, seed=11)
    ref = ovie.dense_AN(G, Mr, Mc, Gram, idx) @ x
    op = NOperator.from_toeplitz(G, None, Mc, Gram, idx, fft_shape=fft, ctx=vp_ctx)
    assert rel(op.apply(x), ref) < 1e-12
    assert rel(op.apply(x, prec='fp32'), ref) < 1e-5


def test_matvec_linearity_and_zero(vp_ctx):
    from voxphoton_amd.vie import NOperator
    G, Mr, Mc, Gram, idx, x = random_problem(12, 8, 6, seed=5)
    op = NOperator.from_toeplitz(G, None, Mc, Gram, idx, ctx=vp_ctx)
    rng = np.random.default_rng(0)
    z = rng.standard_normal(x.size) + 1j * rng.standard_normal(x.size)
    a = 0.3 - 1.7j
    assert rel(op.apply(a * x + z), a * op.apply(x) + op.apply(z)) < 1e-13
    assert np.all(op.apply(np.zeros_like(x)) == 0)


def test_assembly_matches_oracle(vp_ctx):
    from voxphoton_amd.vie import assemble_toeplitz
    dx, ko = 20e-9, 2 * np.pi * np.sqrt(2.085) / 1550e-9
    Gd = assemble_toeplitz((7, 6, 5), dx, ko)
    Go = ogr.assemble_G(7, 6, 5, dx, ko)
    err = np.abs(Gd - Go).max() / np.abs(Go).max()
    print('assembly rel err', err, flush=True)
    assert err < 1e-12



@pytest.mark.parametrize('variant', ['fused', 'split'])
@pytest.mark.parametrize('shape,fft', [((8, 6, 4), (0, 0, 0)), ((9, 4, 5), (21, 9, 9)), ((20, 12, 6), (0, 0, 0)),
                                       ((4, 11, 9), (0, 0, 0))])
def test_matvec_variants(vp_ctx, monkeypatch, variant, shape, fft):
    # Both yz-plane strategies (fused residue-class kernel, split y/z passes)
    # against the dense Toeplitz matrix; 'transp' is its conjugate transpose.
    from voxphoton_amd.vie import NOperator
    monkeypatch.setenv('VP_MATVEC_VARIANT', variant)
    G, Mr, Mc, Gram, idx, x = random_problem(*shape, seed=21, with_mr=True)
    A = ovie.dense_AN(G, Mr, Mc, Gram, idx)
    op = NOperator.from_toeplitz(G, Mr, Mc, Gram, idx, fft_shape=fft, ctx=vp_ctx)
    for flag, ref in (('notransp', A @ x), ('transp', A.conj().T @ x)):
        e64 = rel(op.apply(x, flag, 'fp64'), ref)
        e32 = rel(op.apply(x, flag, 'fp32'), ref)
        print(variant, shape, fft, flag, 'rel', e64, e32, flush=True)
        assert e64 < 1e-12 and e32 < 1e-5
