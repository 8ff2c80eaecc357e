"""GPU parity of the preconditioners, GMRES and mixed-precision solves (C ABI)
against the FP64 oracle."""
import numpy as np
import pytest

from oracle import vie as ovie, precond as opre, krylov as okry, greens as ogr
from tests.helpers import random_problem, random_toeplitz, rel

pytestmark = pytest.mark.gpu


def _blocks(rng, L, M, N):
    n3 = 3 * N
    return (rng.standard_normal((L, M, n3, n3)) + 1j * rng.standard_normal((L, M, n3, n3))) / n3


@pytest.mark.parametrize('shape', [(4, 3, 2), (10, 6, 3), (16, 5, 4), (8, 1, 2)])
def test_chan2_apply_matches_oracle(vp_ctx, shape):
    from voxphoton_amd.vie import Preconditioner
    L, M, N = shape
    rng = np.random.default_rng(L * 100 + M)
    B = _blocks(rng, L, M, N)
    mask = rng.random((L, M, N)) < 0.7
    mask.flat[0] = True
    idx = ovie.idx_from_mask(mask)
    x = rng.standard_normal(idx.size) + 1j * rng.standard_normal(idx.size)
    ref = opre.chan_2_mvp_for_parallel_idx(B, x, L, M, N, idx)
    P = Preconditioner.chan2(B, L, M, N, idx, ctx=vp_ctx)
    y = P.apply(x)
    print(shape, 'chan2 rel', rel(y, ref), flush=True)
    assert rel(y, ref) < 1e-12
    assert rel(P.apply(x, 'fp32'), ref) < 1e-5


def test_chan1_apply_matches_oracle(vp_ctx):
    from voxphoton_amd.vie import Preconditioner
    L, M, N = 6, 3, 2
    rng = np.random.default_rng(5)
    n3 = 3 * M * N
    B = (rng.standard_normal((L, n3, n3)) + 1j * rng.standard_normal((L, n3, n3))) / n3
    idx = ovie.idx_from_mask(rng.random((L, M, N)) < 0.8)
    x = rng.standard_normal(idx.size) + 1j * rng.standard_normal(idx.size)
    ref = opre.chan_mvp_idx(B, x, L, M, N, idx)
    P = Preconditioner.chan1(B, L, M, N, idx, ctx=vp_ctx)
    assert rel(P.apply(x), ref) < 1e-12


@pytest.mark.parametrize('mode', ['fused', 'split', 'quad'])
@pytest.mark.parametrize('sym', ['1', '0'])
@pytest.mark.parametrize('shape', [(6, 4, 3), (10, 5, 2), (7, 5, 3), (9, 8, 4)])
def test_level2_device_setup_matches_oracle(vp_ctx, monkeypatch, shape, sym, mode):
    # sym=1 stores the (kx <= L/2, ky <= M/2) block quadrant, sym=0 all blocks;
    # the plane runs fused, as y-row DFT / block-product / inverse passes
    # (split), or with one workgroup per stored block (quad, sym=1 only)
    from voxphoton_amd.vie import Preconditioner
    monkeypatch.setenv('VP_PC_SYMMETRY', sym)
    monkeypatch.setenv('VP_PC_SPLIT', '1' if mode == 'split' else '0')
    monkeypatch.setenv('VP_PC_QUAD', '1' if mode == 'quad' else '0')
    L, M, N = shape
    dx, ko = 20e-9, 5.85e6
    G = ogr.assemble_G(L, M, N, dx, ko)
    mc, gram = 0.8278, dx ** 3
    idx = ovie.idx_from_mask(np.ones((L, M, N), bool))
    Binv = opre.level_2_parallel_func(G, mc, 1.0, L, M, N, gram)
    rng = np.random.default_rng(0)
    x = rng.standard_normal(idx.size) + 1j * rng.standard_normal(idx.size)
    ref = opre.chan_2_mvp_for_parallel_idx(Binv, x, L, M, N, idx)
    P = Preconditioner.level2(G, mc, 1.0, gram, idx, ctx=vp_ctx)
    y = P.apply(x)
    print(shape, 'level2 rel', rel(y, ref), flush=True)
    assert rel(y, ref) < 1e-10


@pytest.mark.parametrize('shape', [(6, 4, 3), (7, 5, 4), (5, 10, 9)])
def test_level1_setup_matches_oracle(vp_ctx, shape):
    # level_1_parallel_func built and inverted on the device (3MN = 36, 60:
    # Gauss-Jordan kernels; 270: rocSOLVER's batched LU), scalar contrast and
    # contrast field, applied as chan_mvp_idx
    from voxphoton_amd.vie import Preconditioner
    L, M, N = shape
    dx, ko = 20e-9, 5.85e6
    G = ogr.assemble_G(L, M, N, dx, ko)
    rng = np.random.default_rng(L * M * N)
    mask = rng.random(shape) < 0.8
    mask.flat[0] = True
    idx = ovie.idx_from_mask(mask)
    x = rng.standard_normal(idx.size) + 1j * rng.standard_normal(idx.size)
    mc = 0.8278 - 0.02j
    ref = opre.chan_mvp_idx(opre.level_1_parallel_func(G, mc, 1.0, L, M, N, dx ** 3), x, L, M, N, idx)
    P = Preconditioner.level1(G, mc, 1.0, dx ** 3, idx, ctx=vp_ctx)
    e64, e32 = rel(P.apply(x), ref), rel(P.apply(x, 'fp32'), ref)
    Pa = Preconditioner.level1_assembled(shape, dx, ko, mc, 1.0, idx, ctx=vp_ctx)
    ea = rel(Pa.apply(x), ref)
    Mc = np.where(rng.random(shape) < 0.6, 0.83, 0.0) + 0.01 * rng.standard_normal(shape)
    reff = opre.chan_mvp_idx(opre.level_1_parallel_func(G, Mc, 1.0, L, M, N, dx ** 3), x, L, M, N, idx)
    ef = rel(Preconditioner.level1(G, Mc, 1.0, dx ** 3, idx, ctx=vp_ctx).apply(x), reff)
    print(shape, '3MN', 3 * M * N, 'level1', e64, e32, 'assembled', ea, 'field', ef, flush=True)
    assert e64 < 1e-10 and ea < 1e-10 and ef < 1e-10 and e32 < 1e-4


def test_level1_face_matches_oracle(vp_ctx):
    # the face-averaged 1-level form of Bragg_MOR_script.m:61-72
    from voxphoton_amd.vie import Preconditioner
    L, M, N = 8, 5, 4
    dx, ko = 20e-9, 5.85e6
    G = ogr.assemble_G(L, M, N, dx, ko)
    rng = np.random.default_rng(12)
    Mr_mid = 1.0 + 4.8 * (rng.random((L, M, N)) < 0.5) + 0.05j * rng.random((L, M, N))
    hey = np.real(np.sum(Mr_mid, axis=0) / L)
    face = (hey - 1.0) / hey
    idx = ovie.idx_from_mask(np.ones((L, M, N), bool))
    x = rng.standard_normal(idx.size) + 1j * rng.standard_normal(idx.size)
    ref = opre.chan_mvp_idx(opre.level_1_parallel_func_N_face(G, Mr_mid, dx ** 3), x, L, M, N, idx)
    P = Preconditioner.level1_face((L, M, N), dx, ko, face, idx, ctx=vp_ctx)
    e64 = rel(P.apply(x), ref)
    print('level1 face', e64, flush=True)
    assert e64 < 1e-10


@pytest.mark.parametrize('shape', [(6, 4, 3), (9, 8, 4), (12, 10, 2)])
def test_level2_contrast_field_matches_oracle(vp_ctx, shape):
    # level_2_parallel_func with an L x M x N contrast array (the resonator
    # script's disk square): Toeplitz entries scaled by Mc at their offsets
    from voxphoton_amd.vie import Preconditioner
    L, M, N = shape
    dx, ko = 20e-9, 5.85e6
    G = ogr.assemble_G(L, M, N, dx, ko)
    rng = np.random.default_rng(L + M)
    Mc = np.where(rng.random(shape) < 0.6, 0.83 - 0.05j, 0.0) + 0.02 * rng.standard_normal(shape)
    mask = rng.random(shape) < 0.8
    mask.flat[0] = True
    idx = ovie.idx_from_mask(mask)
    Binv = opre.level_2_parallel_func(G, Mc, 1.0, L, M, N, dx ** 3)
    x = rng.standard_normal(idx.size) + 1j * rng.standard_normal(idx.size)
    ref = opre.chan_2_mvp_for_parallel_idx(Binv, x, L, M, N, idx)
    P = Preconditioner.level2(G, Mc, 1.0, dx ** 3, idx, ctx=vp_ctx)
    e64 = rel(P.apply(x), ref)
    Pa = Preconditioner.level2_assembled(shape, dx, ko, Mc, 1.0, idx, ctx=vp_ctx)
    ea = rel(Pa.apply(x), ref)
    e32 = rel(Pa.apply(x, 'fp32'), ref)
    print(shape, 'contrast field: host G', e64, 'assembled', ea, 'fp32', e32, flush=True)
    assert e64 < 1e-10 and ea < 1e-10 and e32 < 2e-3
    # a uniform field is the scalar form
    Pu = Preconditioner.level2(G, np.full(shape, 0.83 - 0.05j), 1.0, dx ** 3, idx, ctx=vp_ctx)
    Ps = Preconditioner.level2(G, 0.83 - 0.05j, 1.0, dx ** 3, idx, ctx=vp_ctx)
    assert rel(Pu.apply(x), Ps.apply(x)) < 1e-13


@pytest.mark.parametrize('half', ['1', '0', 'dot2'])
def test_level2_blocks_wider_than_a_wavefront(vp_ctx, monkeypatch, half):
    # 3N = 66 > 64: block inversion spans several wavefronts.  The FP32 apply
    # of the quadrant form streams binary16 blocks by default (11-bit
    # significands: ~5e-4 per entry) or FP32 blocks with VP_PC_HALF=0; with
    # VP_PC_DOT2=1 the block product also rounds the vector to binary16.
    from voxphoton_amd.vie import Preconditioner
    monkeypatch.setenv('VP_PC_HALF', '0' if half == '0' else '1')
    monkeypatch.setenv('VP_PC_DOT2', '1' if half == 'dot2' else '0')
    L, M, N = 4, 3, 22
    dx, ko = 20e-9, 5.85e6
    G = ogr.assemble_G(L, M, N, dx, ko)
    idx = ovie.idx_from_mask(np.ones((L, M, N), bool))
    Binv = opre.level_2_parallel_func(G, 0.8278, 1.0, L, M, N, dx ** 3)
    rng = np.random.default_rng(6)
    x = rng.standard_normal(idx.size) + 1j * rng.standard_normal(idx.size)
    ref = opre.chan_2_mvp_for_parallel_idx(Binv, x, L, M, N, idx)
    P = Preconditioner.level2(G, 0.8278, 1.0, dx ** 3, idx, ctx=vp_ctx)
    assert rel(P.apply(x), ref) < 1e-10
    e32 = rel(P.apply(x, 'fp32'), ref)
    print({'1': 'fp32 apply, binary16 blocks', '0': 'fp32 apply', 'dot2': 'fp32 apply, dot2'}[half], e32, flush=True)
    assert e32 < (1e-4 if half == '0' else 2e-3)
    if half != '0':
        assert e32 > 1e-6      # the binary16 copy is what ran
    if half == 'dot2':         # the dot2 product is what ran
        y_dot = P.apply(x, 'fp32')
        monkeypatch.setenv('VP_PC_DOT2', '0')
        assert not np.array_equal(y_dot, P.apply(x, 'fp32'))


def test_block_product_window_changes_no_result(vp_ctx, monkeypatch):
    # the binary16 block product keeps 8, 16 or 32 columns of block loads in
    # flight (tuned when the FP32 preconditioner is made, VP_PC_QU overrides);
    # every row still adds its columns in order, so the applies are equal bit
    # for bit, and the tuned default is one of them
    from voxphoton_amd.vie import Preconditioner
    monkeypatch.setenv('VP_PC_HALF', '1')
    monkeypatch.setenv('VP_PC_DOT2', '0')
    for (L, M, N) in [(6, 4, 22), (8, 6, 15), (5, 4, 60)]:
        dx, ko = 20e-9, 5.85e6
        G = ogr.assemble_G(L, M, N, dx, ko)
        idx = ovie.idx_from_mask(np.ones((L, M, N), bool))
        rng = np.random.default_rng(L + N)
        x = rng.standard_normal(idx.size) + 1j * rng.standard_normal(idx.size)
        monkeypatch.delenv('VP_PC_QU', raising=False)
        P = Preconditioner.level2(G, 0.8278, 1.0, dx ** 3, idx, ctx=vp_ctx)
        y_tuned = P.apply(x, 'fp32')
        ys = {}
        for qu in ('8', '16', '32'):
            monkeypatch.setenv('VP_PC_QU', qu)
            ys[qu] = P.apply(x, 'fp32')
        assert np.array_equal(ys['8'], ys['16']) and np.array_equal(ys['8'], ys['32'])
        assert np.array_equal(y_tuned, ys['16'])
        ref = opre.chan_2_mvp_for_parallel_idx(opre.level_2_parallel_func(G, 0.8278, 1.0, L, M, N, dx ** 3), x,
                                               L, M, N, idx)
        assert rel(y_tuned, ref) < 2e-3


@pytest.mark.parametrize('quad', ['1', '0'])
@pytest.mark.parametrize('sym', ['1', '0'])
def test_level2_face_matches_oracle(vp_ctx, monkeypatch, sym, quad):
    from voxphoton_amd.vie import Preconditioner
    monkeypatch.setenv('VP_PC_SYMMETRY', sym)
    monkeypatch.setenv('VP_PC_QUAD', quad)
    L, M, N = 8, 5, 4
    dx, ko = 20e-9, 5.85e6
    G = ogr.assemble_G(L, M, N, dx, ko)
    face = np.array([0.3, 0.8, 0.8, 0.5], complex)
    idx = ovie.idx_from_mask(np.ones((L, M, N), bool))
    Binv = opre.level_2_face_blocks(G, face, dx ** 3)
    rng = np.random.default_rng(4)
    x = rng.standard_normal(idx.size) + 1j * rng.standard_normal(idx.size)
    ref = opre.chan_2_mvp_for_parallel_idx(Binv, x, L, M, N, idx)
    P = Preconditioner.level2_face((L, M, N), dx, ko, face, idx, ctx=vp_ctx)
    assert rel(P.apply(x), ref) < 1e-10
    assert rel(P.apply(x, 'fp32'), ref) < (2e-3 if sym == '1' and quad == '1' else 1e-4)


@pytest.mark.parametrize('ct', ['1', '0'])
@pytest.mark.parametrize('shape', [(100, 10, 3), (200, 11, 2)])
def test_level2_pencil_major_quadrant(vp_ctx, monkeypatch, shape, ct):
    # quadrant form over pencil-major spectra (compile-time x passes of length
    # L and row passes of length M) against the plane-major form and the
    # oracle; the Toeplitz tensor is the device assembly (pinned elsewhere)
    from voxphoton_amd.vie import Preconditioner, assemble_toeplitz
    monkeypatch.setenv('VP_PC_CT', ct)
    L, M, N = shape
    dx, ko = 20e-9, 5.85e6
    G = assemble_toeplitz(shape, dx, ko, ctx=vp_ctx)
    mc, gram = 0.8278, dx ** 3
    mask = np.ones(shape, bool)
    mask[::7, 1, :] = False
    idx = ovie.idx_from_mask(mask)
    Binv = opre.level_2_parallel_func(G, mc, 1.0, L, M, N, gram)
    rng = np.random.default_rng(L + M)
    x = rng.standard_normal(idx.size) + 1j * rng.standard_normal(idx.size)
    ref = opre.chan_2_mvp_for_parallel_idx(Binv, x, L, M, N, idx)
    P = Preconditioner.level2(G, mc, 1.0, gram, idx, ctx=vp_ctx)
    e64, e32 = rel(P.apply(x), ref), rel(P.apply(x, 'fp32'), ref)
    print(shape, 'ct', ct, 'rel', e64, e32, flush=True)
    assert e64 < 1e-10 and e32 < 2e-3


@pytest.mark.parametrize('shape', [(100, 22, 3), (100, 48, 2), (200, 12, 2)])
def test_q_contiguous_block_product_matches_gathers(vp_ctx, monkeypatch, shape):
    # the block product over q-contiguous spectra (row transforms that change
    # the layout: fused for M = 22 and 12 in FP32 and M = 12 in FP64, the
    # pencil-major row pass with separate transposes for M = 22 in FP64, whose
    # tile exceeds 64 KiB, and for M = 48, a length of three radices) equals
    # the same preconditioner gathering its columns from the pencil-major
    # spectra, and the oracle
    from voxphoton_amd.vie import Preconditioner, assemble_toeplitz
    L, M, N = shape
    dx, ko = 20e-9, 5.85e6
    G = assemble_toeplitz(shape, dx, ko, ctx=vp_ctx)
    mc, gram = 0.8278, dx ** 3
    mask = np.ones(shape, bool)
    mask[::7, 1, :] = False
    idx = ovie.idx_from_mask(mask)
    rng = np.random.default_rng(L + 3 * M)
    x = rng.standard_normal(idx.size) + 1j * rng.standard_normal(idx.size)
    monkeypatch.setenv('VP_PC_QLAY', '1')
    Pq = Preconditioner.level2(G, mc, 1.0, gram, idx, ctx=vp_ctx)
    monkeypatch.setenv('VP_PC_QLAY', '0')
    Pg = Preconditioner.level2(G, mc, 1.0, gram, idx, ctx=vp_ctx)
    Binv = opre.level_2_parallel_func(G, mc, 1.0, L, M, N, gram)
    ref = opre.chan_2_mvp_for_parallel_idx(Binv, x, L, M, N, idx)
    q64, g64 = Pq.apply(x), Pg.apply(x)
    q32, g32 = Pq.apply(x, 'fp32'), Pg.apply(x, 'fp32')
    print(shape, 'q vs gathers', rel(q64, g64), rel(q32, g32), 'oracle', rel(q64, ref), rel(q32, ref), flush=True)
    assert rel(q64, g64) < 1e-13 and rel(q32, g32) < 1e-5
    assert rel(q64, ref) < 1e-10 and rel(q32, ref) < 2e-3


def test_level2_plane_larger_than_lds(vp_ctx):
    # 3N x M plane of 90 x 60 complex128 exceeds LDS: the split passes run
    from voxphoton_amd.vie import Preconditioner
    L, M, N = 4, 60, 30
    rng = np.random.default_rng(8)
    G = random_toeplitz(L, M, N, rng) * 1e-3
    idx = ovie.idx_from_mask(np.ones((L, M, N), bool))
    Binv = opre.level_2_parallel_func(G, 0.5, 1.0, L, M, N, 1.0)
    x = rng.standard_normal(idx.size) + 1j * rng.standard_normal(idx.size)
    ref = opre.chan_2_mvp_for_parallel_idx(Binv, x, L, M, N, idx)
    P = Preconditioner.level2(G, 0.5, 1.0, 1.0, idx, ctx=vp_ctx)
    assert rel(P.apply(x), ref) < 1e-10
    assert rel(P.apply(x, 'fp32'), ref) < 2e-3      # binary16 blocks of the quadrant form


@pytest.mark.parametrize('reg', ['1', '0'])
@pytest.mark.parametrize('N', [6, 12, 20, 26, 30, 40, 50, 60, 64])
def test_level2_block_inversion_every_shape(vp_ctx, monkeypatch, N, reg):
    # 3N x 3N blocks of 18..192 rows: the register Gauss-Jordan kernel at every
    # elements-per-thread shape (1..6 x 1..6 per thread, 256 or 512 threads),
    # the panel kernel from 133 to 192 rows, the global-memory kernel with
    # VP_PC_INVERT_REG=0
    from voxphoton_amd.vie import Preconditioner
    monkeypatch.setenv('VP_PC_INVERT_REG', reg)
    L, M = 3, 2
    rng = np.random.default_rng(N)
    # unit-scale tensor: blocks far from the identity (condition numbers up to
    # ~5e2), so partial pivoting moves most rows
    G = random_toeplitz(L, M, N, rng)
    idx = ovie.idx_from_mask(np.ones((L, M, N), bool))
    Binv = opre.level_2_parallel_func(G, 0.5, 1.0, L, M, N, 1.0)
    x = rng.standard_normal(idx.size) + 1j * rng.standard_normal(idx.size)
    ref = opre.chan_2_mvp_for_parallel_idx(Binv, x, L, M, N, idx)
    P = Preconditioner.level2(G, 0.5, 1.0, 1.0, idx, precisions=('fp64',), ctx=vp_ctx)
    e = rel(P.apply(x), ref)
    print('3N', 3 * N, 'register kernel', reg, 'rel', e, flush=True)
    assert e < 1e-9


def test_level2_without_parity_zeros_keeps_all_blocks(vp_ctx):
    # Odd components that do not vanish on their zero planes break the
    # mirror symmetry of the Chan blocks: every block must be stored.
    from voxphoton_amd.vie import Preconditioner
    L, M, N = 6, 5, 3
    rng = np.random.default_rng(12)
    G = random_toeplitz(L, M, N, rng) * 1e-3
    G[0, :, :, 1] += 2e-4 * (1 + 1j)
    idx = ovie.idx_from_mask(np.ones((L, M, N), bool))
    Binv = opre.level_2_parallel_func(G, 0.5, 1.0, L, M, N, 1.0)
    x = rng.standard_normal(idx.size) + 1j * rng.standard_normal(idx.size)
    ref = opre.chan_2_mvp_for_parallel_idx(Binv, x, L, M, N, idx)
    P = Preconditioner.level2(G, 0.5, 1.0, 1.0, idx, ctx=vp_ctx)
    assert rel(P.apply(x), ref) < 1e-10


def test_embedded_preconditioner_block_jacobi(vp_ctx):
    # box preconditioner acting on part of a larger unknown vector (prec_fun_disk_idx.m semantics)
    from voxphoton_amd.vie import Preconditioner
    rng = np.random.default_rng(9)
    L, M, N = 6, 4, 2
    B = _blocks(rng, L, M, N)
    idx = ovie.idx_from_mask(np.ones((L, M, N), bool))
    n_outer = idx.size + 17
    sub = rng.permutation(n_outer)[: idx.size] + 1
    P = Preconditioner.chan2(B, L, M, N, idx, ctx=vp_ctx).set_embedding(n_outer, sub)
    x = rng.standard_normal(n_outer) + 1j * rng.standard_normal(n_outer)
    ref = x.copy()
    ref[sub - 1] = opre.chan_2_mvp_for_parallel_idx(B, x[sub - 1], L, M, N, idx)
    assert rel(P.apply(x), ref) < 1e-12


def _physical(L, M, N, seed=0, frac=1.0):
    dx = 20e-9
    ko = 2 * np.pi * np.sqrt(2.085) / 1550e-9
    G = ogr.assemble_G(L, M, N, dx, ko)
    rng = np.random.default_rng(seed)
    mask = rng.random((L, M, N)) < frac if frac < 1 else np.ones((L, M, N), bool)
    mask.flat[0] = True
    idx = ovie.idx_from_mask(mask)
    er = 12.11 / 2.085
    Mc = np.full((L, M, N), (er - 1) / er + 0j)
    Gram = dx ** 3
    b = (rng.standard_normal(idx.size) + 1j * rng.standard_normal(idx.size)) * Gram
    return G, Mc, Gram, idx, b, dx, ko


@pytest.mark.parametrize('orth', ['cgs', 'dgks', 'cgs2'])
def test_gmres_fp64_matches_oracle(vp_ctx, monkeypatch, orth):
    # one Gram-Schmidt pass per step (default, as pgmres_iter.m), the
    # DGKS-selected second pass and the second pass at every step: the
    # residual histories of the oracle's single and double pass GMRES
    from voxphoton_amd.vie import NOperator, pgmres
    monkeypatch.setenv('VP_DGKS_ETA', '0.7071' if orth == 'dgks' else '0')
    monkeypatch.setenv('VP_GMRES_CGS2', '1' if orth == 'cgs2' else '0')
    G, Mc, Gram, idx, b, dx, ko = _physical(6, 5, 4, seed=1, frac=0.7)
    op = NOperator.from_toeplitz(G, None, Mc, Gram, idx, ctx=vp_ctx)
    fN = ovie.fN_from_toeplitz(G)
    Mr = np.ones(Mc.shape, complex)
    A = lambda v: ovie.mv_AN_const(v, fN, Mr, Mc, Gram, 'notransp', idx)   # noqa: E731
    xo, flo, rro, ito, rvo = okry.pgmres(A, b, 10, 1e-10, 20, orth='cgs' if orth == 'cgs' else 'cgs2')
    x, flag, relres, it, rv = pgmres(op, b, 10, 1e-10, 20)
    print('gmres iters', len(rv), len(rvo), 'relres', relres, rro, flush=True)
    assert flag == 0 and flo == 0
    k = min(len(rv), len(rvo), 8)
    assert np.allclose(rv[:k], rvo[:k], rtol=1e-6, atol=1e-12)
    Ad = ovie.dense_AN(G, Mr, Mc, Gram, idx)
    assert rel(x, np.linalg.solve(Ad, b)) < 1e-8
    assert np.linalg.norm(b - Ad @ x) / np.linalg.norm(b) <= 1e-10


@pytest.mark.parametrize('side', ['left', 'right'])
def test_gmres_preconditioned(vp_ctx, side):
    from voxphoton_amd.vie import NOperator, Preconditioner, pgmres
    L, M, N = 8, 4, 3
    G, Mc, Gram, idx, b, dx, ko = _physical(L, M, N, seed=2)
    op = NOperator.from_toeplitz(G, None, Mc, Gram, idx, ctx=vp_ctx)
    P = Preconditioner.level2(G, Mc.flat[0], 1.0, Gram, idx, ctx=vp_ctx)
    kw = {'L1': P} if side == 'left' else {'R1': P}
    x, flag, relres, it, rv = pgmres(op, b, 20, 1e-9, 10, **kw)
    x0, flag0, relres0, it0, rv0 = pgmres(op, b, 20, 1e-9, 10)
    print(side, 'prec iters', len(rv), 'unprec iters', len(rv0), flush=True)
    Ad = ovie.dense_AN(G, np.ones(Mc.shape, complex), Mc, Gram, idx)
    assert flag == 0
    assert np.linalg.norm(b - Ad @ x) / np.linalg.norm(b) <= 1e-9 * 1.01
    assert len(rv) < len(rv0)


def test_gmres_step_chunks_give_identical_results(vp_ctx, monkeypatch):
    # the Arnoldi steps run in chunks gated on the device; the chunk length
    # (host round trips) changes nothing: same iterates, histories, flags,
    # for single and batched solves in which problems stop at different steps
    from voxphoton_amd.vie import NOperator, Preconditioner, pgmres, solve_ir_batched
    G, Mc, Gram, idx, b, dx, ko = _physical(8, 5, 4, seed=6)
    op = NOperator.from_toeplitz(G, None, Mc, Gram, idx, ctx=vp_ctx)
    P = Preconditioner.level2(G, Mc.flat[0], 1.0, Gram, idx, ctx=vp_ctx)
    rng = np.random.default_rng(6)
    B = np.array([b, rng.standard_normal(b.size) + 1j * rng.standard_normal(b.size), 0.5 * b])
    out = {}
    for chunk in ('1', '3', '64'):
        monkeypatch.setenv('VP_GMRES_CHUNK', chunk)
        single = pgmres(op, b, 17, 1e-10, 6, R1=P)
        X, infos = solve_ir_batched([op] * 3, [P] * 3, B, tol=1e-9, restart=13, inner_tol=1e-4, inner_maxit=3,
                                    max_outer=10, side='right')
        out[chunk] = (single, X, infos)
    s1, X1, i1 = out['1']
    for chunk in ('3', '64'):
        s, X, infos = out[chunk]
        assert np.array_equal(s[0], s1[0]) and s[1] == s1[1] and s[2] == s1[2]
        assert np.array_equal(s[4], s1[4])
        assert np.array_equal(X, X1)
        for a, c in zip(infos, i1):
            assert a['inner_iters'] == c['inner_iters'] and a['relres'] == c['relres']
            assert np.array_equal(a['history'], c['history'])
    print('chunked GMRES: single', len(s1[4]) - 1, 'its; batched inner its', [c['inner_iters'] for c in i1],
          flush=True)


def test_gmres_graph_replays_equal_stream_launches(vp_ctx, monkeypatch):
    # VP_GMRES_GRAPH=1 captures each chunk of Arnoldi steps into a hipGraph on
    # first use and replays it in later restart cycles and later inner solves
    # of a refinement run.  Every replay must give what the stream launches
    # give: same iterates, histories and flags (bit for bit), for right, left
    # and no preconditioning, single and batched, with eager work and
    # synchronisations between the solves that replay the graphs.
    from voxphoton_amd.vie import NOperator, Preconditioner, pgmres, solve_ir_batched
    G, Mc, Gram, idx, b, dx, ko = _physical(8, 5, 4, seed=6)
    op = NOperator.from_toeplitz(G, None, Mc, Gram, idx, ctx=vp_ctx)
    P = Preconditioner.level2(G, Mc.flat[0], 1.0, Gram, idx, ctx=vp_ctx)
    rng = np.random.default_rng(7)
    B = np.array([b, rng.standard_normal(b.size) + 1j * rng.standard_normal(b.size), 0.5 * b])

    def run():
        out = [pgmres(op, b, 9, 1e-11, 8, R1=P), pgmres(op, b, 9, 1e-11, 8, L1=P),
               pgmres(op, b, 9, 1e-11, 8, prec='fp32')]
        vp_ctx.synchronize()
        X, infos = solve_ir_batched([op] * 3, [P] * 3, B, tol=1e-10, restart=11, inner_tol=1e-4, inner_maxit=3,
                                    max_outer=12, side='right')
        return out, X, infos

    monkeypatch.setenv('VP_GMRES_CHUNK', '3')
    monkeypatch.setenv('VP_GMRES_GRAPH', '0')
    ref = run()
    monkeypatch.setenv('VP_GMRES_GRAPH', '1')
    for rep in range(3):
        got = run()
        # eager work between the graph runs
        y = op.apply(b)
        z = P.apply(y, prec='fp32')
        assert np.isfinite(z).all()
        vp_ctx.synchronize()
        for (s, s0) in zip(got[0], ref[0]):
            assert np.array_equal(s[0], s0[0]) and s[1] == s0[1] and s[2] == s0[2]
            assert np.array_equal(s[4], s0[4])
        assert np.array_equal(got[1], ref[1])
        for a, c in zip(got[2], ref[2]):
            assert a['inner_iters'] == c['inner_iters'] and a['relres'] == c['relres']
            assert np.array_equal(a['history'], c['history'])
    cycles = [len(s0[4]) - 1 for s0 in ref[0]]
    print('graph GMRES equals stream launches: iterations (right, left, fp32 none)', cycles,
          'batched inner its', [c['inner_iters'] for c in ref[2]], flush=True)
    assert max(cycles) > 18     # several restart cycles: the chunk graphs were replayed


def test_gmres_maxit_flag_and_breakdown(vp_ctx):
    from voxphoton_amd.vie import NOperator, pgmres
    G, Mc, Gram, idx, b, dx, ko = _physical(5, 4, 3, seed=4)
    op = NOperator.from_toeplitz(G, None, Mc, Gram, idx, ctx=vp_ctx)
    x, flag, relres, it, rv = pgmres(op, b, 2, 1e-14, 1)
    assert flag == 1 and relres > 1e-14            # non-convergence is a flag, not an error
    op0 = NOperator.from_toeplitz(G, None, 0 * Mc, Gram, idx, ctx=vp_ctx)   # A = Gram I
    x, flag, relres, it, rv = pgmres(op0, b, 10, 1e-12, 3)
    assert flag == 0 and len(rv) <= 3 and rel(x, b / Gram) < 1e-12


def test_mixed_precision_ir_reaches_fp64_tolerance(vp_ctx):
    from voxphoton_amd.vie import NOperator, Preconditioner, solve_ir
    L, M, N = 12, 6, 4
    G, Mc, Gram, idx, b, dx, ko = _physical(L, M, N, seed=3)
    op = NOperator.from_toeplitz(G, None, Mc, Gram, idx, ctx=vp_ctx)
    P = Preconditioner.level2(G, Mc.flat[0], 1.0, Gram, idx, ctx=vp_ctx)
    x, info = solve_ir(op, b, tol=1e-10, restart=30, inner_tol=1e-4, inner_maxit=3, max_outer=20, M=P)
    fN = ovie.fN_from_toeplitz(G)
    r = b - ovie.mv_AN_const(x, fN, np.ones(Mc.shape, complex), Mc, Gram, 'notransp', idx)
    true_rel = np.linalg.norm(r) / np.linalg.norm(b)
    print('IR', info, 'true relres', true_rel, flush=True)
    assert info['flag'] == 0 and true_rel <= 1e-10 * 1.01
    assert abs(true_rel - info['relres']) <= 1e-3 * true_rel + 1e-15


def test_recycle_width_of_the_context(vp_ctx, monkeypatch):
    # vp_ctx_set_recycle_width: width 0 keeps the whole last cycle (equal bit
    # for bit to the default without VP_IR_RECYCLE_K), width k the k harmonic
    # Ritz directions; with a narrow space the solve still reaches the
    # tolerance (the oracle recomputes the residual) and batched = solo
    from voxphoton_amd.vie import NOperator, Preconditioner, solve_ir, solve_ir_batched
    monkeypatch.delenv('VP_IR_RECYCLE_K', raising=False)
    L, M, N = 12, 6, 4
    G, Mc, Gram, idx, b, dx, ko = _physical(L, M, N, seed=3)
    op = NOperator.from_toeplitz(G, None, Mc, Gram, idx, ctx=vp_ctx)
    P = Preconditioner.level2(G, Mc.flat[0], 1.0, Gram, idx, ctx=vp_ctx)
    fN = ovie.fN_from_toeplitz(G)
    ones = np.ones(Mc.shape, complex)
    # short inner solves that run their whole cycle, so the cycle is wider
    # than the narrow space
    prm = dict(tol=1e-10, restart=12, inner_tol=1e-4, inner_maxit=1, max_outer=60, M=P, side='right')
    try:
        x_def, i_def = solve_ir(op, b, **prm)
        vp_ctx.set_recycle_width(0)
        x0, i0 = solve_ir(op, b, **prm)
        assert np.array_equal(x_def, x0) and i_def['inner_iters'] == i0['inner_iters']
        vp_ctx.set_recycle_width(4)
        x4, i4 = solve_ir(op, b, **prm)
        r = b - ovie.mv_AN_const(x4, fN, ones, Mc, Gram, 'notransp', idx)
        true_rel = np.linalg.norm(r) / np.linalg.norm(b)
        print(f'recycled width 4: {i4["ir_steps"]} steps {i4["inner_iters"]} inner its relres {i4["relres"]:.3e} '
              f'(oracle {true_rel:.3e}); whole cycle {i0["ir_steps"]} steps {i0["inner_iters"]} its', flush=True)
        assert i4['flag'] == 0 and true_rel <= 1e-10 * 1.01
        assert abs(true_rel - i4['relres']) <= 1e-3 * true_rel + 1e-15
        assert not np.array_equal(x4, x0)
        rng = np.random.default_rng(6)
        B = np.array([b, (rng.standard_normal(b.size) + 1j * rng.standard_normal(b.size)) * np.abs(b).max()])
        bprm = dict(prm)
        bprm.pop('M')
        X, infos = solve_ir_batched([op, op], [P, P], B, **bprm)
        for i in range(2):
            xs, inf_s = solve_ir(op, B[i], **prm)
            assert np.array_equal(X[i], xs) and infos[i]['inner_iters'] == inf_s['inner_iters']
        from voxphoton_amd._lib import VoxPhotonError
        with pytest.raises(VoxPhotonError):
            vp_ctx.set_recycle_width(-2)
    finally:
        vp_ctx.set_recycle_width(-1)


@pytest.mark.parametrize('outer_m', [30, 3])
def test_flexible_refinement_reaches_fp64_tolerance(vp_ctx, outer_m):
    # vp_solve_fgmres_ir: FGMRES in FP64 over the inner FP32 solves.  The true
    # residual (recomputed by the oracle) meets the tolerance and equals the
    # reported one; the history falls monotonically inside a cycle (the outer
    # step minimises the residual over all corrections of the cycle); cycles
    # of 3 outer steps restart from the true residual; with weak inner solves
    # it needs no more steps than plain refinement; batched = solo bit for bit.
    from voxphoton_amd.vie import NOperator, Preconditioner, solve_ir, solve_ir_batched
    L, M, N = 12, 6, 4
    G, Mc, Gram, idx, b, dx, ko = _physical(L, M, N, seed=3)
    op = NOperator.from_toeplitz(G, None, Mc, Gram, idx, ctx=vp_ctx)
    P = Preconditioner.level2(G, Mc.flat[0], 1.0, Gram, idx, ctx=vp_ctx)
    fN = ovie.fN_from_toeplitz(G)
    ones = np.ones(Mc.shape, complex)
    # weak inner solves (restart 6, one cycle) so that many outer steps are needed
    prm = dict(tol=1e-10, restart=6, inner_tol=1e-2, inner_maxit=1, max_outer=60, M=P, side='right')
    x, info = solve_ir(op, b, outer_restart=outer_m, **prm)
    r = b - ovie.mv_AN_const(x, fN, ones, Mc, Gram, 'notransp', idx)
    true_rel = np.linalg.norm(r) / np.linalg.norm(b)
    x_ir, info_ir = solve_ir(op, b, **prm)
    print(f'flexible refinement (cycles of {outer_m}): {info["ir_steps"]} steps, {info["inner_iters"]} inner its, '
          f'relres {info["relres"]:.3e}, oracle {true_rel:.3e}; plain refinement {info_ir["ir_steps"]} steps '
          f'({info_ir["relres"]:.3e})', flush=True)
    assert info['flag'] == 0 and true_rel <= 1e-10 * 1.01
    assert abs(true_rel - info['relres']) <= 1e-3 * true_rel
    h = np.asarray(info['history'])
    assert h[0] == 1.0 and len(h) == info['ir_steps'] + 1
    if outer_m == 30:
        # estimates fall monotonically; the last entry is the recomputed residual
        assert np.all(np.diff(h[:-1]) <= 1e-12 * h[:-2])
        assert info['ir_steps'] <= info_ir['ir_steps']
    rng = np.random.default_rng(4)
    B = np.array([b, (rng.standard_normal(b.size) + 1j * rng.standard_normal(b.size)) * np.abs(b).max()])
    bprm = dict(prm)
    bprm.pop('M')
    X, infos = solve_ir_batched([op, op], [P, P], B, outer_restart=outer_m, **bprm)
    for i in range(2):
        xs, inf_s = solve_ir(op, B[i], outer_restart=outer_m, **prm)
        assert infos[i]['flag'] == 0 and inf_s['flag'] == 0
        assert np.array_equal(X[i], xs) and infos[i]['ir_steps'] == inf_s['ir_steps']


@pytest.mark.parametrize('variant', ['fused', 'split'])
def test_batched_ir_alone_equals_in_batch(vp_ctx, monkeypatch, variant):
    from voxphoton_amd.vie import NOperator, Preconditioner, solve_ir, solve_ir_batched
    monkeypatch.setenv('VP_MATVEC_VARIANT', variant)
    L, M, N = 10, 5, 3
    dx = 20e-9
    ops, precs, rhs = [], [], []
    rng = np.random.default_rng(12)
    idx = ovie.idx_from_mask(np.ones((L, M, N), bool))
    er = 12.11 / 2.085
    Mc = np.full((L, M, N), (er - 1) / er + 0j)
    for lam in (1500e-9, 1550e-9, 1600e-9):
        ko = 2 * np.pi * np.sqrt(2.085) / lam
        ops.append(NOperator.assembled((L, M, N), dx, ko, None, Mc, idx, ctx=vp_ctx))
        precs.append(Preconditioner.level2_assembled((L, M, N), dx, ko, Mc.flat[0], 1.0, idx, ctx=vp_ctx))
    for i in range(4):   # two RHS share an operator
        rhs.append((rng.standard_normal(idx.size) + 1j * rng.standard_normal(idx.size)) * dx ** 3)
    bops = [ops[0], ops[0], ops[1], ops[2]]
    bprecs = [precs[0], precs[0], precs[1], precs[2]]
    X, infos = solve_ir_batched(bops, bprecs, np.array(rhs), tol=1e-8, restart=20, inner_tol=1e-3,
                                inner_maxit=3, max_outer=15)
    for i in range(4):
        xa, ia = solve_ir(bops[i], rhs[i], tol=1e-8, restart=20, inner_tol=1e-3, inner_maxit=3, max_outer=15,
                          M=bprecs[i])
        print(i, infos[i]['relres'], ia['relres'], rel(X[i], xa), flush=True)
        assert infos[i]['flag'] == 0 and ia['flag'] == 0
        assert rel(X[i], xa) <= 1e-12      # same result alone and in the mix


@pytest.mark.parametrize('side', ['none', 'right', 'left'])
def test_pgmresDR_matches_oracle(vp_ctx, side):
    # deflated restarts: same residual history and solution as the oracle
    from voxphoton_amd.vie import NOperator, Preconditioner
    from voxphoton_amd.krylov_dr import pgmresDR
    L, M, N = 10, 6, 4
    G, Mc, Gram, idx, b, dx, ko = _physical(L, M, N, seed=6, frac=0.8)
    op = NOperator.from_toeplitz(G, None, Mc, Gram, idx, ctx=vp_ctx)
    P = Preconditioner.level2(G, Mc.flat[0], 1.0, Gram, idx, ctx=vp_ctx)
    fN = ovie.fN_from_toeplitz(G)
    ones = np.ones(Mc.shape, complex)
    Aref = lambda v: ovie.mv_AN_const(v, fN, ones, Mc, Gram, 'notransp', idx)    # noqa: E731
    Binv = opre.level_2_parallel_func(G, Mc.flat[0], 1.0, L, M, N, Gram)
    Pref = lambda v: opre.chan_2_mvp_for_parallel_idx(Binv, v, L, M, N, idx)      # noqa: E731
    kw = dict(R1=P) if side == 'right' else dict(L1=P) if side == 'left' else {}
    kwo = dict(R1=Pref) if side == 'right' else dict(L1=Pref) if side == 'left' else {}
    restart, tol, maxit, ritz = 8, 1e-9, 40, 4
    x, flag, relres, it, rv = pgmresDR(op, b, restart, tol, maxit, ritz, **kw)
    xo, flago, relreso, ito, rvo = okry.pgmresDR(Aref, b, restart, tol, maxit, ritz, **kwo)
    print(side, 'device', len(rv), it, relres, 'oracle', len(rvo), ito, relreso, flush=True)
    assert flag == flago == 0 and it == ito and len(rv) == len(rvo)
    assert np.allclose(rv, rvo, rtol=1e-6, atol=1e-13)
    assert rel(x, xo) < 1e-8


@pytest.mark.parametrize('chany', ['fft', 'direct'])
@pytest.mark.parametrize('shape', [(6, 10, 2), (5, 12, 3), (4, 49, 2)])
def test_level2_chan_y_by_fft_matches_oracle(vp_ctx, monkeypatch, shape, chany):
    # the y-Chan transform of the block builder as one FFT per line (long y
    # axes) or as direct sums: the same blocks as the oracle's
    from voxphoton_amd.vie import Preconditioner
    monkeypatch.setenv('VP_PC_CHANY', chany)
    L, M, N = shape
    G = ogr.assemble_G(L, M, N, 20e-9, 5.85e6) if M < 40 else random_toeplitz(L, M, N, np.random.default_rng(3),
                                                                                 scale=1e-24)
    mc, gram = 0.8278 + 0.01j, 8e-24
    idx = ovie.idx_from_mask(np.ones((L, M, N), bool))
    Binv = opre.level_2_parallel_func(G, mc, 1.0, L, M, N, gram)
    rng = np.random.default_rng(1)
    x = rng.standard_normal(idx.size) + 1j * rng.standard_normal(idx.size)
    ref = opre.chan_2_mvp_for_parallel_idx(Binv, x, L, M, N, idx)
    P = Preconditioner.level2(G, mc, 1.0, gram, idx, ctx=vp_ctx)
    e = rel(P.apply(x), ref)
    print(shape, chany, 'level2 rel', e, flush=True)
    assert e < 1e-10
