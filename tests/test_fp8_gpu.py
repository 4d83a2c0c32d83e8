"""GPU parity of the FP8 low-fidelity precision (QFX_PREC_FP8).

The FP8 form quantizes A and the real 2x2 expansion of B per 128 x 128 tile
(power-of-two scale, E4M3 round to nearest even) and sums the products in
FP32.  oracle/fp8.py does the same quantization and sums in float64, so
the GPU result must match it to FP32 summation error, far below the FP8
quantization error itself (a few percent), which is also checked.

Products with m n <= 2^16 (and k <= 4096) run in exact FP32 on the vector
ALUs in every precision, FP8 included, so the FP8 shapes here are larger."""
import numpy as np
import pytest

from oracle import fp8
from qflex_amd import capi

from test_mapped_gpu import make_case

pytestmark = pytest.mark.gpu

# The FP8 matrix instruction does not round each sum to FP32: its outputs on
# this hardware come out on a grid about 2^-16 of the largest partial sum
# (measured relative difference to the float64 sums of the same quantized
# values: 1.5e-5).  The quantization itself costs a few 1e-2, and a wrong
# scale, tile or sign costs more than that.
TOL_VS_ORACLE = 1e-4


def rand_c(rng, shape, scale=1.0):
    return ((rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) * scale).astype(
        np.complex64)


def gpu_cgemm(ctx, a, b, prec=capi.FP8):
    m, k = a.shape
    n = b.shape[1]
    A, B, C = ctx.upload(a), ctx.upload(b), ctx.alloc(m * n * 8)
    ctx.cgemm(prec, m, n, k, A, B, C)
    got = ctx.download(C, np.complex64, (m, n))
    for buf in (A, B, C):
        buf.free()
    return got


def relerr(got, ref):
    return float(np.linalg.norm(got - ref) / np.linalg.norm(ref))


SHAPES = [(512, 256, 128), (2048, 64, 64), (1000, 70, 100), (1024, 512, 256), (520, 200, 1000),
          (4096, 256, 256)]


@pytest.mark.parametrize("shape", SHAPES)
def test_fp8_cgemm_matches_oracle(ctx, shape):
    m, n, k = shape
    print("fp8 cgemm", shape, flush=True)
    rng = np.random.default_rng(m + n + k)
    a, b = rand_c(rng, (m, k)), rand_c(rng, (k, n))
    got = gpu_cgemm(ctx, a, b)
    want = fp8.cgemm_fp8(a, b)
    assert relerr(got, want) <= TOL_VS_ORACLE, (shape, relerr(got, want))
    exact = a.astype(np.complex128) @ b.astype(np.complex128)
    err = relerr(got, exact)
    assert 1e-3 < err < 0.1, (shape, err)  # the FP8 rounding really happened, and only it


def test_fp8_cgemm_small_integers_exact(ctx):
    """Small integers are exact in E4M3 at any tile scale, and their FP32 sums
    are exact: the result is the complex product bit for bit."""
    rng = np.random.default_rng(11)
    m, n, k = 384, 192, 320
    a = (rng.integers(-8, 9, (m, k)) + 1j * rng.integers(-8, 9, (m, k))).astype(np.complex64)
    b = (rng.integers(-8, 9, (k, n)) + 1j * rng.integers(-8, 9, (k, n))).astype(np.complex64)
    a[:128, :64] *= 2.0 ** 6
    b[64:, :] *= 2.0 ** -5
    got = gpu_cgemm(ctx, a, b)
    want = (a.astype(np.complex128) @ b.astype(np.complex128)).astype(np.complex64)
    assert np.array_equal(got, want)


def test_fp8_cgemm_tiles_of_very_different_scales(ctx):
    """Per-tile scales: a tile 2^40 larger than its neighbours does not wipe
    out the small tiles (one scale per matrix would)."""
    rng = np.random.default_rng(12)
    m, n, k = 1024, 128, 256
    a, b = rand_c(rng, (m, k)), rand_c(rng, (k, n))
    a[128:256, 64:] *= 2.0 ** 40
    b[:64, :] *= 2.0 ** -20
    got = gpu_cgemm(ctx, a, b)
    want = fp8.cgemm_fp8(a, b)
    # rows 128..255 meet the large tile; compare the row tiles separately
    for r0 in range(0, m, 128):
        assert relerr(got[r0:r0 + 128], want[r0:r0 + 128]) <= TOL_VS_ORACLE, r0


def test_fp8_zero_operand_tile(ctx):
    rng = np.random.default_rng(13)
    a, b = rand_c(rng, (1024, 256)), rand_c(rng, (256, 128))
    a[:, :64] = 0
    got = gpu_cgemm(ctx, a, b)
    assert relerr(got, fp8.cgemm_fp8(a, b)) <= TOL_VS_ORACLE


def test_fp8_small_shapes_run_as_fp32(ctx):
    """Below the FP8 tiles (here n < 64) the product is computed in FP32."""
    rng = np.random.default_rng(14)
    a, b = rand_c(rng, (512, 256)), rand_c(rng, (256, 32))
    got = gpu_cgemm(ctx, a, b)
    exact = a.astype(np.complex128) @ b.astype(np.complex128)
    assert relerr(got, exact) < 2e-6


# (left, common, right, a_inner, b_inner) as in test_mapped_gpu, with k >= 64.
FP8_MAPPED = [
    ([16, 16, 16], [16, 8], [8, 16], "left", "right"),
    ([16, 16, 16], [8, 16], [16, 8], "common", "right"),
    ([16, 4, 16], [16, 16], [2, 16, 4], "common", "right"),
    ([4, 16, 16], [16, 16], [16, 8], "left", "common"),
    ([4, 16, 16], [4, 16, 2], [16, 8], "common", "common"),
]


@pytest.mark.parametrize("case", range(len(FP8_MAPPED)))
def test_fp8_mapped_matches_oracle(ctx, case):
    """Operands read in their stored index orders (every form is accepted:
    the quantize kernels gather through the digit maps)."""
    a_dims, a_left, a_common, b_dims, b_common, b_right = make_case(*FP8_MAPPED[case], seed=case)
    print("fp8 mapped case", case, flush=True)
    rng = np.random.default_rng(100 + case)
    a, b = rand_c(rng, a_dims), rand_c(rng, b_dims)
    m = int(np.prod([a_dims[x] for x in a_left]))
    k = int(np.prod([a_dims[x] for x in a_common]))
    n = int(np.prod([b_dims[x] for x in b_right]))
    am = a.transpose(a_left + a_common).reshape(m, k)
    bm = b.transpose(b_common + b_right).reshape(k, n)
    maps = (capi.index_map(a_dims, a_left), capi.index_map(a_dims, a_common),
            capi.index_map(b_dims, b_common), capi.index_map(b_dims, b_right))
    A, B, C = ctx.upload(a), ctx.upload(b), ctx.alloc(m * n * 8)
    assert ctx.cgemm_mapped(capi.FP8, A, maps[0], maps[1], B, maps[2], maps[3], C)
    got = ctx.download(C, np.complex64, (m, n))
    for buf in (A, B, C):
        buf.free()
    assert relerr(got, fp8.cgemm_fp8(am, bm)) <= TOL_VS_ORACLE
    exact = am.astype(np.complex128) @ bm.astype(np.complex128)
    assert 1e-3 < relerr(got, exact) < 0.1


def test_fp8_mapped_declines_small_k(ctx):
    """k below the FP8 tiles: the mapped entry declines (the engine then
    reorders and runs the row-major entry, which uses FP32 MFMA)."""
    a_dims, a_left, a_common, b_dims, b_common, b_right = make_case(
        [16, 16, 16], [16, 2], [8, 16], "left", "right", seed=0)
    maps = (capi.index_map(a_dims, a_left), capi.index_map(a_dims, a_common),
            capi.index_map(b_dims, b_common), capi.index_map(b_dims, b_right))
    A = ctx.alloc(int(np.prod(a_dims)) * 8)
    B = ctx.alloc(int(np.prod(b_dims)) * 8)
    C = ctx.alloc(4096 * 128 * 8)
    assert not ctx.cgemm_mapped(capi.FP8, A, maps[0], maps[1], B, maps[2], maps[3], C)
    for buf in (A, B, C):
        buf.free()
