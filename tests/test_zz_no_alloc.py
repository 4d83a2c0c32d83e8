"""qfx_cgemm_noalloc_c64 (no device allocation; scratch from the library's
fixed per-GPU array) against qfx_cgemm_c64 on the same inputs: the same
kernels run, so the results must be bit-identical.  Each case calls the
no-allocation entry twice in a row, the second time with a smaller product,
and covers a path that uses scratch: split-K partial sums (tiny outputs and
small outputs over long k), the transposed GEMV, and the split-FP16 operand
maxima of the large tiles.  The arena must not grow, and a product whose
scratch exceeds the fixed array is refused without a launch."""
import numpy as np
import pytest

from qflex_amd import capi

pytestmark = pytest.mark.gpu

# (precision, first shape, second smaller shape)
CASES = [
    ("split_fp16", (4096, 256, 256), (1024, 128, 64)),     # large tiles + maxima partials
    ("split_fp16", (8, 8, 1 << 17), (4, 4, 4096)),          # tiny m, n: chunked partial sums
    ("fp32", (32, 64, 8192), (16, 32, 2048)),               # small outputs, long k: split-K
    ("split_bf16x3", (1, 512, 4096), (1, 128, 1024)),       # transposed GEMV partials
    ("fp8", (1024, 512, 256), (256, 128, 128)),              # FP8 codes and tile scales
    ("fp32_exact", (300, 20, 40), (10, 3, 17)),             # no scratch at all
]


def rand_c(rng, shape):
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)


@pytest.mark.parametrize("prec,first,second", CASES)
def test_noalloc_matches_cgemm_two_calls(prec, first, second):
    print("noalloc", prec, first, second, flush=True)
    ctx = capi.Context(0)
    try:
        p = capi.PRECISIONS[prec]
        rng = np.random.default_rng(sum(first) + sum(second))
        bufs = []
        for (m, n, k) in (first, second):
            a, b = rand_c(rng, (m, k)), rand_c(rng, (k, n))
            A, B = ctx.upload(a), ctx.upload(b)
            C1, C2 = ctx.alloc(m * n * 8), ctx.alloc(m * n * 8)
            bufs.append((m, n, k, A, B, C1, C2))
        ctx.sync()
        before = ctx.arena_stats()
        for (m, n, k, A, B, C1, C2) in bufs:
            ctx.cgemm_noalloc(p, m, n, k, A, B, C2)
        ctx.sync()
        assert ctx.arena_stats() == before  # nothing allocated
        for (m, n, k, A, B, C1, C2) in bufs:
            ctx.cgemm(p, m, n, k, A, B, C1)
            want = ctx.download(C1, np.complex64, (m, n))
            got = ctx.download(C2, np.complex64, (m, n))
            assert np.array_equal(got, want), (prec, m, n, k)
    finally:
        ctx.close()


def test_noalloc_refuses_products_beyond_the_fixed_scratch():
    """FP8 codes of a 4096^3 product need 96 MiB of scratch: refused (no
    launch), and the entry works again right after."""
    ctx = capi.Context(0)
    try:
        m = n = k = 4096
        A, B, C = ctx.alloc(m * k * 8), ctx.alloc(k * n * 8), ctx.alloc(m * n * 8)
        ctx.zero(A)
        ctx.zero(B)
        with pytest.raises(capi.QfxError, match="fixed"):
            ctx.cgemm_noalloc(capi.FP8, m, n, k, A, B, C)
        rng = np.random.default_rng(1)
        a, b = rand_c(rng, (64, 32)), rand_c(rng, (32, 16))
        A2, B2, C2, C3 = ctx.upload(a), ctx.upload(b), ctx.alloc(64 * 16 * 8), ctx.alloc(64 * 16 * 8)
        ctx.cgemm_noalloc(capi.FP32, 64, 16, 32, A2, B2, C2)
        ctx.cgemm(capi.FP32, 64, 16, 32, A2, B2, C3)
        assert np.array_equal(ctx.download(C2, np.complex64, (64, 16)),
                              ctx.download(C3, np.complex64, (64, 16)))
    finally:
        ctx.close()
