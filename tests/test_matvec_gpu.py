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


@pytest.mark.parametrize('shape,box', [((20, 12, 10), (5, 4, 3)), ((24, 16, 8), (11, 3, 8)), ((9, 30, 6), (9, 7, 2))])
def test_caller_fN_resampled_to_the_box(vp_ctx, monkeypatch, shape, box):
    # a caller's fN at the reference's sizes (2L, 2M, 2N) on a grid whose
    # unknowns fill a smaller box: the tensor is resampled to circulants of
    # the box (the shortest smooth lengths here) and gives the same operator
    from voxphoton_amd.vie import NOperator
    L, M, N = shape
    rng = np.random.default_rng(L * M + N)
    G, Mr, Mc, Gram, _, _ = random_problem(L, M, N, seed=L + M + N, with_mr=True)
    mask = np.zeros(shape, bool)
    o = [rng.integers(0, s - b + 1) for s, b in zip(shape, box)]
    sub = tuple(slice(a, a + b) for a, b in zip(o, box))
    mask[sub] = rng.random(box) < 0.7
    mask[tuple(a for a in o)] = True
    mask[tuple(a + b - 1 for a, b in zip(o, box))] = True
    idx = ovie.idx_from_mask(mask)
    x = rng.standard_normal(idx.size) + 1j * rng.standard_normal(idx.size)
    fN = ovie.fN_from_toeplitz(G)
    monkeypatch.setenv('VP_FFT_SIZES', 'min')
    op = NOperator.from_fN(fN, Mr, Mc, Gram, idx, ctx=vp_ctx)
    info = op.info()
    assert info['box'] == box
    for f in (info['fft_fp64'], info['fft_fp32']):
        assert all(2 * b - 1 <= p <= 2 * s for p, b, s in zip(f, box, shape)), (f, box)
        assert any(p < 2 * s for p, s in zip(f, shape)), f
    monkeypatch.setenv('VP_FN_RESAMPLE', '0')
    op_kept = NOperator.from_fN(fN, Mr, Mc, Gram, idx, ctx=vp_ctx)
    assert op_kept.info()['fft_fp64'] == (2 * L, 2 * M, 2 * N)
    for flag in ('notransp', 'transp'):
        ref = ovie.mv_AN_const(x, fN, Mr, Mc, Gram, flag, idx)
        y = op.apply(x, flag, 'fp64')
        print(shape, box, info['fft_fp64'], flag, 'fp64 rel', rel(y, ref), flush=True)
        assert rel(y, ref) < 1e-12
        assert rel(op.apply(x, flag, 'fp32'), ref) < 1e-5
        assert rel(op_kept.apply(x, flag, 'fp64'), ref) < 1e-12
    op.close()
    op_kept.close()


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


@pytest.mark.parametrize('near', ['1', '0'])
@pytest.mark.parametrize('shape', [(7, 6, 5), (3, 8, 2)])
def test_assembly_matches_oracle(vp_ctx, monkeypatch, shape, near):
    # near cells (every offset below 5) integrated by one workgroup each
    # (VP_ASSEMBLE_NEAR=1) or by one thread each
    from voxphoton_amd.vie import assemble_toeplitz
    monkeypatch.setenv('VP_ASSEMBLE_NEAR', near)
    dx, ko = 20e-9, 2 * np.pi * np.sqrt(2.085) / 1550e-9
    Gd = assemble_toeplitz(shape, dx, ko)
    Go = ogr.assemble_G(*shape, dx, ko)
    err = np.abs(Gd - Go).max() / np.abs(Go).max()
    print(shape, 'near', near, 'assembly rel err', err, flush=True)
    assert err < 1e-12



@pytest.mark.parametrize('variant', ['fused', 'split', 'split_zgen', 'split_zcol'])
@pytest.mark.parametrize('shape,fft', [((8, 6, 4), (0, 0, 0)), ((9, 4, 5), (21, 9, 9)), ((20, 12, 6), (0, 0, 0)),
                                       ((4, 11, 9), (0, 0, 0)), ((6, 5, 22), (0, 0, 44)), ((5, 7, 24), (0, 0, 48)),
                                       ((4, 5, 22), (0, 0, 88)), ((3, 4, 40), (0, 0, 88)), ((5, 10, 6), (0, 0, 64)),
                                       ((2, 100, 3), (0, 200, 0)), ((3, 48, 2), (0, 96, 0)),
                                       ((30, 6, 4), (200, 0, 0)), ((25, 6, 3), (400, 0, 0))])
def test_matvec_variants(vp_ctx, monkeypatch, variant, shape, fft):
    # The pass strategies (fused residue-class yz kernel; split y/z passes
    # with the compile-time x and y passes (pencil-major x spectra where the
    # class length has them) and the two-stage fused z kernel, with the
    # generic LDS x, y and z passes, or with the register z-column kernel)
    # against the dense Toeplitz matrix; 'transp' is its conjugate transpose.
    from voxphoton_amd.vie import NOperator
    monkeypatch.setenv('VP_MATVEC_VARIANT', variant.split('_')[0])
    monkeypatch.setenv('VP_ZCOL', '1' if variant.endswith('zcol') else '0')
    monkeypatch.setenv('VP_ZFUSED', '0' if variant.endswith('zgen') else '1')
    monkeypatch.setenv('VP_YCT', '0' if variant.endswith('zgen') else '1')
    G, Mr, Mc, Gram, idx, x = random_problem(*shape, seed=21, with_mr=True)
    A = ovie.dense_AN(G, Mr, Mc, Gram, idx)
    op = NOperator.from_toeplitz(G, Mr, Mc, Gram, idx, fft_shape=fft, ctx=vp_ctx)
    for flag, ref in (('notransp', A @ x), ('transp', A.conj().T @ x)):
        e64 = rel(op.apply(x, flag, 'fp64'), ref)
        e32 = rel(op.apply(x, flag, 'fp32'), ref)
        print(variant, shape, fft, flag, 'rel', e64, e32, flush=True)
        assert e64 < 1e-12 and e32 < 1e-5


@pytest.mark.parametrize('variant', ['fused', 'split'])
@pytest.mark.parametrize('prune', ['1', '0'])
@pytest.mark.parametrize('form', ['structured', 'generic'])
def test_matvec_unknowns_in_sub_box(vp_ctx, monkeypatch, variant, prune, form):
    # Unknowns confined to an offset sub-box with a hole: the passes run on
    # the bounding box only (prune=1) or on the full grid (prune=0).
    from voxphoton_amd.vie import NOperator
    monkeypatch.setenv('VP_MATVEC_VARIANT', variant)
    monkeypatch.setenv('VP_MATVEC_PRUNE', prune)
    L, M, N = 9, 8, 7
    rng = np.random.default_rng(33)
    G, Mr, Mc, Gram, _, _ = random_problem(L, M, N, seed=33, with_mr=True)
    mask = np.zeros((L, M, N), bool)
    mask[2:7, 3:7, 1:4] = True
    mask[4, 5, 2] = False
    idx = ovie.idx_from_mask(mask)
    if form == 'generic':
        idx = rng.permutation(idx)[: idx.size - 7]
    x = rng.standard_normal(idx.size) + 1j * rng.standard_normal(idx.size)
    A = ovie.dense_AN(G, Mr, Mc, Gram, idx)
    op = NOperator.from_toeplitz(G, Mr, Mc, Gram, idx, ctx=vp_ctx)
    for flag, ref in (('notransp', A @ x), ('transp', A.conj().T @ x)):
        e64 = rel(op.apply(x, flag, 'fp64'), ref)
        e32 = rel(op.apply(x, flag, 'fp32'), ref)
        print(variant, prune, form, flag, 'rel', e64, e32, flush=True)
        assert e64 < 1e-12 and e32 < 1e-5


@pytest.mark.parametrize('variant', ['fused', 'split'])
@pytest.mark.parametrize('sym', ['1', '0'])
def test_matvec_fN_with_and_without_parity_structure(vp_ctx, monkeypatch, variant, sym):
    # A caller's fN whose odd components do not vanish on their zero planes
    # lacks the octant symmetry: it must be stored in full.  A symmetric fN
    # is stored as its octant unless VP_FN_SYMMETRY=0.
    from voxphoton_amd.vie import NOperator
    monkeypatch.setenv('VP_MATVEC_VARIANT', variant)
    monkeypatch.setenv('VP_FN_SYMMETRY', sym)
    rng = np.random.default_rng(44)
    L, M, N = 7, 6, 5
    G, Mr, Mc, Gram, idx, x = random_problem(L, M, N, seed=44)
    Gasym = G + 0.5 * (rng.standard_normal(G.shape) + 1j * rng.standard_normal(G.shape))
    for GG in (G, Gasym):
        fN = ovie.fN_from_toeplitz(GG)
        op = NOperator.from_fN(fN, None, Mc, Gram, idx, ctx=vp_ctx)
        for flag in ('notransp', 'transp'):
            ref = ovie.mv_AN_const(x, fN, Mr, Mc, Gram, flag, idx)
            e64 = rel(op.apply(x, flag, 'fp64'), ref)
            e32 = rel(op.apply(x, flag, 'fp32'), ref)
            print(variant, sym, GG is G, flag, e64, e32, flush=True)
            assert e64 < 1e-12 and e32 < 1e-5


@pytest.mark.parametrize('pair', ['1', '0'])
@pytest.mark.parametrize('shape,fft', [((30, 6, 3), (3600, 0, 0)), ((20, 6, 3), (7200, 0, 0))])
def test_matvec_long_x_classes(vp_ctx, monkeypatch, shape, fft, pair):
    # x lines zero padded to 3600 and 7200 points: compile-time class
    # transforms of 1800 and 3600 points; the FP32 inverse runs both classes
    # of a pencil as one batch (VP_XINV_PAIR=1) or one after the other
    from voxphoton_amd.vie import NOperator
    monkeypatch.setenv('VP_MATVEC_VARIANT', 'split')
    monkeypatch.setenv('VP_XINV_PAIR', pair)
    G, Mr, Mc, Gram, idx, x = random_problem(*shape, seed=77, with_mr=True)
    A = ovie.dense_AN(G, Mr, Mc, Gram, idx)
    op = NOperator.from_toeplitz(G, Mr, Mc, Gram, idx, fft_shape=fft, ctx=vp_ctx)
    for flag, ref in (('notransp', A @ x), ('transp', A.conj().T @ x)):
        e64 = rel(op.apply(x, flag, 'fp64'), ref)
        e32 = rel(op.apply(x, flag, 'fp32'), ref)
        print(shape, fft, 'pair', pair, flag, 'rel', e64, e32, flush=True)
        assert e64 < 1e-12 and e32 < 1e-5
