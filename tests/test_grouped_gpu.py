"""GPU tests of grouped products (qfx_cgemm_grouped_c64): products of uneven
shapes in one persistent launch give the complex128 result to the
precision's tolerance, and bit for bit what each product gives alone on the
256 x 128 tile; products outside the large tiles run one by one."""
import time

import numpy as np
import pytest

from qflex_amd import capi

pytestmark = pytest.mark.gpu

TOL = {capi.SPLIT_FP16: 5e-6, capi.SPLIT_BF16: 3e-4, capi.SPLIT_BF16X3: 5e-6, capi.FP32: 2e-6}

SHAPES = [(16384, 256, 256), (33000, 130, 74), (300, 130, 74), (64, 64, 32), (1000, 20, 16),
          (512, 512, 48), (7, 300, 128), (2048, 18, 40)]


def rand_c(rng, shape):
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)


def make(ctx, rng, shapes):
    probs, host = [], []
    for (m, n, k) in shapes:
        a, b = rand_c(rng, (m, k)), rand_c(rng, (k, n))
        host.append((a, b))
        probs.append((m, n, k, ctx.upload(a), ctx.upload(b), ctx.alloc(m * n * 8)))
    return probs, host


def free(probs):
    for p in probs:
        for buf in p[3:]:
            buf.free()


@pytest.mark.parametrize("prec", [capi.SPLIT_FP16, capi.SPLIT_BF16, capi.SPLIT_BF16X3, capi.FP32])
def test_grouped_uneven_shapes(ctx, prec):
    rng = np.random.default_rng(prec + 40)
    probs, host = make(ctx, rng, SHAPES)
    ctx.cgemm_grouped(prec, probs)
    got = [ctx.download(p[5], np.complex64, (p[0], p[1])) for p in probs]
    for (m, n, k), (a, b), g in zip(SHAPES, host, got):
        ref = a.astype(np.complex128) @ b.astype(np.complex128)
        err = np.linalg.norm(g - ref) / np.linalg.norm(ref)
        assert err <= TOL[prec], ((m, n, k), prec, err)
    # products large enough for the 256 x 128 tile alone give the same bits
    for i in (0, 1):
        m, n, k, A, B, C = probs[i]
        single = ctx.alloc(m * n * 8)
        ctx.cgemm(prec, m, n, k, A, B, single)
        assert np.array_equal(ctx.download(single, np.complex64, (m, n)), got[i]), (i, prec)
        single.free()
    free(probs)


def test_grouped_many_small_slices(ctx):
    """256 products of one slice-like shape: one launch, against one call
    each (same kernel class for products below the large-tile block count)."""
    rng = np.random.default_rng(7)
    shapes = [(256, 128, 64)] * 256
    probs, host = make(ctx, rng, shapes)
    ctx.cgemm_grouped(capi.SPLIT_FP16, probs)
    ctx.sync()
    t0 = time.perf_counter()
    ctx.cgemm_grouped(capi.SPLIT_FP16, probs)
    ctx.sync()
    t_grouped = time.perf_counter() - t0
    got = [ctx.download(p[5], np.complex64, (256, 128)) for p in probs]
    t0 = time.perf_counter()
    for (m, n, k, A, B, C) in probs:
        ctx.cgemm(capi.SPLIT_FP16, m, n, k, A, B, C)
    ctx.sync()
    t_single = time.perf_counter() - t0
    print("256 x (256x128x64): grouped %.3f ms, one by one %.3f ms" % (1e3 * t_grouped,
                                                                       1e3 * t_single))
    for (a, b), g in zip(host[:8], got[:8]):
        ref = a.astype(np.complex128) @ b.astype(np.complex128)
        assert np.linalg.norm(g - ref) / np.linalg.norm(ref) <= TOL[capi.SPLIT_FP16]
    free(probs)


def test_grouped_empty_and_refused_while_recording(ctx):
    ctx.cgemm_grouped(capi.SPLIT_FP16, [])
    rng = np.random.default_rng(9)
    probs, _ = make(ctx, rng, [(512, 256, 64)])
    ctx.graph_begin()
    try:
        with pytest.raises(RuntimeError, match="recorded"):
            ctx.cgemm_grouped(capi.SPLIT_FP16, probs)
    finally:
        ctx.graph_end().destroy()
    free(probs)


ROW_SHAPES = [(65536, 8, 2), (1, 2, 2), (300, 16, 16), (4097, 2, 2), (777, 7, 11), (256, 16, 1),
              (20000, 8, 2)]


@pytest.mark.parametrize("prec", [capi.SPLIT_FP16, capi.FP32])
def test_grouped_row_kernel_products_same_bits_as_alone(ctx, prec):
    """k, n <= 16 (the time-line products of a network build): one grouped
    launch of the exact-FP32 row kernel, each result bit-identical to the
    product run alone; a large-tile product in the same call is unaffected."""
    rng = np.random.default_rng(prec + 60)
    shapes = ROW_SHAPES + [(16384, 256, 256)]
    probs, host = make(ctx, rng, shapes)
    ctx.profile(True)
    ctx.profile_reset()
    ctx.cgemm_grouped(prec, probs)
    prof = ctx.profile_read()
    ctx.profile(False)
    assert prof["cgemm_small"]["launches"] == 1
    for (m, n, k, A, B, C), (a, b) in zip(probs, host):
        got = ctx.download(C, np.complex64, (m, n))
        single = ctx.alloc(m * n * 8)
        ctx.cgemm(prec, m, n, k, A, B, single)
        assert np.array_equal(ctx.download(single, np.complex64, (m, n)), got), ((m, n, k), prec)
        single.free()
        ref = a.astype(np.complex128) @ b.astype(np.complex128)
        assert np.linalg.norm(got - ref) / np.linalg.norm(ref) <= TOL[prec], (m, n, k)
    free(probs)


# Products with at least two 256 x 128 (or 128 x 256, 512 x 32) tiles per
# compute unit and an even number of 8-deep k-stages run as chains of tiles
# (each block computes several tiles, loading the next tile's first stages
# during the current tile's last ones); the grouped launch runs one tile at a
# time.  Same arithmetic per element: the same bits.
CHAIN_SHAPES = [(65536, 256, 256), (131072, 512, 256), (70000, 300, 144), (262144, 32, 256),
                (65536, 256, 64)]


@pytest.mark.parametrize("m,n,k", CHAIN_SHAPES)
def test_chained_tiles_bit_identical_to_single_tiles(ctx, m, n, k):
    rng = np.random.default_rng(m + n + k)
    a, b = rand_c(rng, (m, k)), rand_c(rng, (k, n))
    A, B = ctx.upload(a), ctx.upload(b)
    C, G = ctx.alloc(m * n * 8), ctx.alloc(m * n * 8)
    ctx.cgemm(capi.SPLIT_FP16, m, n, k, A, B, C)
    ctx.cgemm_grouped(capi.SPLIT_FP16, [(m, n, k, A, B, G)])
    got = ctx.download(C, np.complex64, (m, n))
    want = ctx.download(G, np.complex64, (m, n))
    assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), (m, n, k)
    rows = np.concatenate([rng.integers(0, m, 500), np.arange(m - 4, m)])
    ref = a[rows].astype(np.complex128) @ b.astype(np.complex128)
    assert np.linalg.norm(got[rows] - ref) / np.linalg.norm(ref) <= 5e-6
    for buf in (A, B, C, G):
        buf.free()


@pytest.mark.parametrize("a_inner", ["left", "common"])
@pytest.mark.parametrize("left,right", [([16] * 4, [16, 16]),        # 256 x 128 tiles
                                        ([2] + [16] * 4, [2, 16, 16])])  # 128 x 256 tiles
def test_chained_mapped_tiles_bit_identical_to_single_tiles(ctx, a_inner, left, right):
    """Operands read in their stored index orders (A along m or along k) by
    chained tiles: the same bits as the grouped launch on the reordered
    matrices."""
    from test_mapped_gpu import make_case
    a_dims, a_l, a_c, b_dims, b_c, b_r = make_case(left, [16, 16], right, a_inner, "right", seed=7)
    rng = np.random.default_rng(7)
    a, b = rand_c(rng, a_dims), rand_c(rng, b_dims)
    m, k, n = int(np.prod(left)), 256, int(np.prod(right))
    am = np.ascontiguousarray(a.transpose(a_l + a_c).reshape(m, k))
    bm = np.ascontiguousarray(b.transpose(b_c + b_r).reshape(k, n))
    maps = (capi.index_map(a_dims, a_l), capi.index_map(a_dims, a_c),
            capi.index_map(b_dims, b_c), capi.index_map(b_dims, b_r))
    A, B, C = ctx.upload(a), ctx.upload(b), ctx.alloc(m * n * 8)
    Am, Bm, G = ctx.upload(am), ctx.upload(bm), ctx.alloc(m * n * 8)
    assert ctx.cgemm_mapped(capi.SPLIT_FP16, A, maps[0], maps[1], B, maps[2], maps[3], C)
    ctx.cgemm_grouped(capi.SPLIT_FP16, [(m, n, k, Am, Bm, G)])
    got = ctx.download(C, np.complex64, (m, n))
    want = ctx.download(G, np.complex64, (m, n))
    assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), a_inner
    for buf in (A, B, C, Am, Bm, G):
        buf.free()
