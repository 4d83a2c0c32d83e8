"""GPU parity of the C-ABI primitives against the oracle's numpy semantics.

Reorders are pure data movement and must be bit-exact.  Complex GEMMs are
compared norm-wise against complex128 (numpy) with the tolerance written per
precision; the defined-order FP32_EXACT path is also compared bit-exactly on
integer-valued inputs, where every partial sum is exact.
"""
import numpy as np
import pytest

from qflex_amd import capi

pytestmark = pytest.mark.gpu


def rand_c64(rng, shape):
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)


def relerr(x, ref):
    return float(np.linalg.norm((x - ref).ravel()) / np.linalg.norm(ref.ravel()))


REORDER_CASES = [
    # (dims, perm)
    ([2, 2, 2], [1, 2, 0]),                       # reference known answer
    ([2] * 10, [9, 4, 6, 2, 5, 1, 7, 8, 3, 0]),   # reference RightTen (tensor_test.cpp:271)
    ([2] * 12, [11, 10, 9, 8, 7, 6, 5, 4, 3, 2, 1, 0]),
    ([16, 16, 16, 16], [3, 0, 2, 1]),
    ([16, 16, 16, 16, 16], [4, 3, 2, 1, 0]),
    ([16, 16, 16, 16, 16, 2], [5, 0, 1, 2, 3, 4]),
    ([2, 16, 16, 16, 16, 16, 16], [1, 2, 0, 3, 4, 5, 6]),
    ([16] * 6, [0, 1, 2, 3, 5, 4]),               # inner swap, outer kept
    ([16] * 6, [1, 0, 2, 3, 4, 5]),               # outer swap, inner kept
    ([256, 4096], [1, 0]),                        # plain transpose
    ([4, 1, 8, 1], [2, 3, 0, 1]),                 # unit axes
    ([3, 5, 6], [2, 0, 1]),                       # non powers of two
    ([7], [0]),
    ([2, 3, 2, 5], [3, 1, 0, 2]),
]


@pytest.mark.parametrize("dims,perm", REORDER_CASES)
def test_reorder_bit_exact(ctx, dims, perm):
    rng = np.random.default_rng(len(dims) * 31 + sum(dims))
    x = rand_c64(rng, dims)
    ref = np.ascontiguousarray(x.transpose(perm))
    src = ctx.upload(x)
    dst = ctx.alloc(x.nbytes)
    ctx.reorder(src, dst, dims, perm)
    got = ctx.download(dst, np.complex64, ref.shape)
    assert np.array_equal(got.view(np.uint64), ref.view(np.uint64))
    src.free()
    dst.free()


def test_reorder_random_binary_permutations(ctx):
    rng = np.random.default_rng(7)
    for nbits in (11, 16, 20, 22):
        x = rand_c64(rng, [2] * nbits)
        src = ctx.upload(x)
        dst = ctx.alloc(x.nbytes)
        for _ in range(3):
            perm = [int(p) for p in rng.permutation(nbits)]
            ctx.reorder(src, dst, [2] * nbits, perm)
            got = ctx.download(dst, np.complex64, x.shape)
            assert np.array_equal(got, np.ascontiguousarray(x.transpose(perm))), (nbits, perm)
        src.free()
        dst.free()


def test_reorder_large_rank28_like(ctx):
    # 2^24 complex64 (128 MiB): a 6-axis bond tensor permutation.
    rng = np.random.default_rng(3)
    dims = [16, 16, 16, 16, 16, 16]
    x = rand_c64(rng, dims)
    perm = [2, 5, 0, 4, 1, 3]
    src = ctx.upload(x)
    dst = ctx.alloc(x.nbytes)
    ctx.reorder(src, dst, dims, perm)
    got = ctx.download(dst, np.complex64, [dims[p] for p in perm])
    assert np.array_equal(got, np.ascontiguousarray(x.transpose(perm)))
    src.free()
    dst.free()


def test_reorder_2gib_rank28(ctx):
    """2^28 complex64 (2 GiB, the largest tensors of the 7x7 workloads): a
    rank-28 binary permutation, checked at sampled output positions."""
    rng = np.random.default_rng(4)
    n = 1 << 28
    x = np.empty(n, np.complex64)
    x.real = np.arange(n, dtype=np.float32)  # exact below 2^24, distinct enough above
    x.imag = rng.random(n, dtype=np.float32)
    dims = [2] * 28
    perm = list(rng.permutation(28))
    src = ctx.upload(x)
    dst = ctx.alloc(x.nbytes)
    ctx.reorder(src, dst, dims, perm)
    got = ctx.download(dst, np.complex64, (n,))
    src.free()
    dst.free()
    # output position o has output bits o_j (bit 27 - j is output axis j);
    # its source position has bit 27 - perm[j] set to o_j
    outs = rng.integers(0, n, 4096, dtype=np.int64)
    srcs = np.zeros_like(outs)
    for j, p in enumerate(perm):
        srcs |= ((outs >> (27 - j)) & 1) << (27 - p)
    assert np.array_equal(got[outs], x[srcs])


@pytest.mark.parametrize("prec", [capi.FP32, capi.SPLIT_FP16])
def test_cgemm_operand_above_2gib(ctx, prec):
    """A of 2.2 GB, past the 32-bit byte offsets of the large-tile kernels:
    the product runs on the 64-bit indexed kernels; rows sampled against
    complex128."""
    m, k, n = (1 << 20) + 5, 264, 20
    rng = np.random.default_rng(5)
    a = np.empty((m, k), np.complex64)
    a.real = rng.random((m, k), dtype=np.float32) - 0.5
    a.imag = rng.random((m, k), dtype=np.float32) - 0.5
    b = rand_c64(rng, (k, n))
    A, B, C = ctx.upload(a), ctx.upload(b), ctx.alloc(m * n * 8)
    ctx.cgemm(prec, m, n, k, A, B, C)
    got = ctx.download(C, np.complex64, (m, n))
    for buf in (A, B, C):
        buf.free()
    rows = np.concatenate([rng.integers(0, m, 2000), np.arange(m - 8, m)])
    ref = a[rows].astype(np.complex128) @ b.astype(np.complex128)
    err = np.linalg.norm(got[rows] - ref) / np.linalg.norm(ref)
    assert err < (2e-6 if prec == capi.FP32 else 5e-6), err


GEMM_SHAPES = [
    (2, 2, 2), (4, 2, 2), (16, 16, 16), (64, 16, 8), (100, 50, 37), (128, 64, 16),
    (256, 256, 256), (4096, 256, 256), (1024, 512, 256), (513, 129, 65), (64, 4096, 16),
    (1, 64, 4096), (1, 3, 5), (300, 1, 77), (5, 1, 3), (300, 130, 40), (1000, 200, 24),
    (4096, 4096, 64), (8, 8, 1 << 20), (3, 5, 5000), (64, 64, 1024), (1024, 1024, 40),
    # small common dimension (vector-ALU tiles of 256x16 and 64x64)
    (4096, 512, 16), (1000, 20, 16), (300, 7, 3), (70000, 16, 16), (257, 65, 9),
    # one thread per row, more than 2048 tiles of 256 rows: each block walks several tiles
    (524365, 16, 15),
    # at most 8 x 8 outputs with long k (split-K with register accumulators)
    (1, 8, 4096), (7, 3, 1500), (8, 1, 2048),
    # small outputs, k a multiple of 16: split-K over the vector-ALU tiles
    (100, 50, 48), (9, 1000, 64),
    # large tiles 256x128 and 128x256 with ragged edges and odd stage counts
    (32768, 256, 24), (33000, 130, 74), (20000, 600, 36),
]
TOL = {capi.FP32: 2e-6, capi.FP32_EXACT: 2e-6, capi.SPLIT_BF16X3: 5e-6, capi.SPLIT_BF16: 3e-4,
       capi.SPLIT_FP16: 5e-6}


@pytest.mark.parametrize("m,n,k", GEMM_SHAPES)
@pytest.mark.parametrize("prec", [capi.FP32, capi.FP32_EXACT, capi.SPLIT_BF16, capi.SPLIT_BF16X3,
                                  capi.SPLIT_FP16])
def test_cgemm_vs_complex128(ctx, m, n, k, prec):
    rng = np.random.default_rng(m * 7 + n * 3 + k)
    a = rand_c64(rng, (m, k))
    b = rand_c64(rng, (k, n))
    ref = a.astype(np.complex128) @ b.astype(np.complex128)
    A, B, C = ctx.upload(a), ctx.upload(b), ctx.alloc(m * n * 8)
    ctx.cgemm(prec, m, n, k, A, B, C)
    got = ctx.download(C, np.complex64, (m, n))
    err = relerr(got.astype(np.complex128), ref)
    tol = TOL[prec]
    if prec == capi.FP32_EXACT:
        # one sequential fmaf chain per output: rounding error grows ~ sqrt(k)
        tol = max(tol, 1e-7 * np.sqrt(k))
    assert err <= tol, (m, n, k, prec, err)
    for buf in (A, B, C):
        buf.free()


def test_cgemm_reference_known_answer(ctx):
    # tests/src/tensor_test.cpp:496-530: A(a,b,c)=0..7, B(b,c,d)=8..1 -> {20,14,100,78}
    a = np.arange(8, dtype=np.complex64).reshape(2, 4)
    b = np.arange(8, 0, -1).astype(np.complex64).reshape(4, 2)
    for prec in (capi.FP32, capi.FP32_EXACT, capi.SPLIT_BF16):
        A, B, C = ctx.upload(a), ctx.upload(b), ctx.alloc(32)
        ctx.cgemm(prec, 2, 2, 4, A, B, C)
        got = ctx.download(C, np.complex64, (2, 2))
        assert np.array_equal(got.ravel(), np.array([20, 14, 100, 78], dtype=np.complex64))


@pytest.mark.parametrize("m,n,k", [(70000, 16, 16), (3000, 7, 11), (1000, 100, 16), (5000, 9, 6),
                                   (524365, 16, 15), (530000, 3, 7)])
def test_small_k_kernels_bit_identical_to_defined_order(ctx, m, n, k):
    """k <= 16 runs on the vector-ALU kernels (one thread per row for n <= 16,
    tiles otherwise); appending a zero column to A and a zero row to B makes
    the FP32_EXACT product go to the scalar kernel for k = 16 (one fmaf chain
    per output in increasing k), and for smaller k to the same kernels with
    one more k (for odd k and more than 2048 tiles of 256 rows: the
    tile-walking row kernel against the one reading whole rows of A): the
    extra fmaf adds exact zeros, so the results must agree bit for bit."""
    rng = np.random.default_rng(m + n + k)
    a = rand_c64(rng, (m, k))
    b = rand_c64(rng, (k, n))
    a1 = np.concatenate([a, np.zeros((m, 1), np.complex64)], axis=1)
    b1 = np.concatenate([b, np.zeros((1, n), np.complex64)], axis=0)
    A, B, C = ctx.upload(a), ctx.upload(b), ctx.alloc(m * n * 8)
    A1, B1, C1 = ctx.upload(a1), ctx.upload(b1), ctx.alloc(m * n * 8)
    ctx.cgemm(capi.FP32_EXACT, m, n, k, A, B, C)
    ctx.cgemm(capi.FP32_EXACT, m, n, k + 1, A1, B1, C1)
    got = ctx.download(C, np.complex64, (m, n))
    want = ctx.download(C1, np.complex64, (m, n))
    assert np.array_equal(got, want)
    for buf in (A, B, C, A1, B1, C1):
        buf.free()


def test_cgemm_exact_integer_inputs_bit_exact(ctx):
    rng = np.random.default_rng(11)
    for (m, n, k) in [(64, 16, 8), (200, 96, 40), (130, 70, 300), (300, 130, 40)]:
        a = (rng.integers(-8, 8, (m, k)) + 1j * rng.integers(-8, 8, (m, k))).astype(np.complex64)
        b = (rng.integers(-8, 8, (k, n)) + 1j * rng.integers(-8, 8, (k, n))).astype(np.complex64)
        ref = (a.astype(np.complex128) @ b.astype(np.complex128)).astype(np.complex64)
        for prec in (capi.FP32, capi.FP32_EXACT, capi.SPLIT_BF16, capi.SPLIT_BF16X3,
                     capi.SPLIT_FP16):
            A, B, C = ctx.upload(a), ctx.upload(b), ctx.alloc(m * n * 8)
            ctx.cgemm(prec, m, n, k, A, B, C)
            got = ctx.download(C, np.complex64, (m, n))
            assert np.array_equal(got, ref), (m, n, k, prec)


@pytest.mark.parametrize("m,k", [(1, 1), (7, 300), (4096, 256), (3, 65536)])
def test_cgemv(ctx, m, k):
    rng = np.random.default_rng(m + k)
    a = rand_c64(rng, (m, k))
    x = rand_c64(rng, (k,))
    xt = rand_c64(rng, (m,))
    A, X, Y = ctx.upload(a), ctx.upload(x), ctx.alloc(m * 8)
    ctx.cgemv(0, m, k, A, X, Y)
    got = ctx.download(Y, np.complex64, (m,))
    assert relerr(got.astype(np.complex128), a.astype(np.complex128) @ x) < 2e-6
    XT, YT = ctx.upload(xt), ctx.alloc(k * 8)
    ctx.cgemv(1, m, k, A, XT, YT)
    got = ctx.download(YT, np.complex64, (k,))
    assert relerr(got.astype(np.complex128), xt.astype(np.complex128) @ a) < 2e-6


@pytest.mark.parametrize("n", [1, 5, 256, 65537, 1 << 22])
def test_cdotu_unconjugated(ctx, n):
    rng = np.random.default_rng(n)
    x, y = rand_c64(rng, (n,)), rand_c64(rng, (n,))
    X, Y, O = ctx.upload(x), ctx.upload(y), ctx.alloc(8)
    ctx.cdotu(n, X, Y, O)
    got = complex(ctx.download(O, np.complex64, (1,))[0])
    ref = complex(np.sum(x.astype(np.complex128) * y.astype(np.complex128)))
    assert abs(got - ref) <= 2e-6 * np.sqrt(n) * max(1.0, abs(ref))


def test_accumulate_fp64(ctx):
    rng = np.random.default_rng(5)
    amps = np.zeros(80, dtype=np.complex128)
    amps[:] = rng.standard_normal(80)
    D = ctx.upload(amps)
    total = amps.copy()
    for s in range(3):
        v = rand_c64(rng, (64,))
        V = ctx.upload(v)
        ctx.accumulate(V, 64, D, offset=16)
        total[16:] += v.astype(np.complex128)
    got = ctx.download(D, np.complex128, (80,))
    assert np.array_equal(got, total)


def test_profile_counts_launches(ctx):
    rng = np.random.default_rng(1)
    a = rand_c64(rng, (1024, 512))
    A, B, C = ctx.upload(a), ctx.upload(a), ctx.alloc(a.nbytes * 2)
    ctx.profile(True)
    ctx.profile_reset()
    ctx.cgemm(capi.FP32, 1024, 512, 64, A, B, C)
    ctx.reorder(A, C, [1024, 512], [1, 0])
    prof = ctx.profile_read()
    ctx.profile(False)
    assert prof["cgemm"]["launches"] == 1 and prof["permute"]["launches"] == 1
    assert prof["cgemm"]["flops"] == 8.0 * 1024 * 512 * 64
    assert prof["cgemm"]["ms"] > 0


@pytest.mark.parametrize("pre,dim,post,picks", [(1, 16, 4096, [3]), (16, 16, 256, [0]),
                                                  (4, 8, 33, [7, 0, 5]), (3, 130, 5, list(range(130)))])
def test_project_gathers_index_values(ctx, pre, dim, post, picks):
    """Tensor::project onto index values (src/tensor.cpp:280-316) in one launch."""
    rng = np.random.default_rng(pre + dim + post)
    x = rand_c64(rng, (pre, dim, post))
    X, Y = ctx.upload(x), ctx.alloc(pre * len(picks) * post * 8)
    ctx.project(X, Y, pre, dim, post, picks)
    got = ctx.download(Y, np.complex64, (pre, len(picks), post))
    assert np.array_equal(got, x[:, picks, :])
    with pytest.raises(RuntimeError):
        ctx.project(X, Y, pre, dim, post, [dim])
    X.free(); Y.free()


@pytest.mark.parametrize("n", [1, 1000, 16384 * 5 + 3, 1 << 23])
def test_amax_partials(ctx, n):
    rng = np.random.default_rng(n)
    x = rand_c64(rng, (n,))
    x[n // 2] = 7.5 - 30j
    X, out = ctx.upload(x), ctx.alloc(256 * 4)
    cnt = ctx.amax_partials(X, n, out)
    parts = ctx.download(out, np.float32, (256,))[:cnt]
    assert 1 <= cnt <= 256
    assert parts.max() == np.float32(30.0)
    assert parts.max() == max(np.abs(x.real).max(), np.abs(x.imag).max())
    X.free(); out.free()


def test_staged_uploads_keep_the_values_at_call_time(ctx):
    """qfx_copy_h2d_async: the host array is reused right after each call;
    20 uploads of 1 MiB also wrap the 16 MiB staging area once."""
    n = (1 << 20) // 8
    host = np.zeros(n, np.complex64)
    bufs = [ctx.alloc(n * 8) for _ in range(20)]
    for i, b in enumerate(bufs):
        host[:] = i + 1j * (i + 1)
        ctx.upload_async(b, host)
    host[:] = -1
    for i, b in enumerate(bufs):
        got = ctx.download(b, np.complex64, (n,))
        assert np.all(got == np.complex64(i + 1j * (i + 1))), i
        b.free()


def _upload_at(ctx, buf, offset, arr):
    arr = np.ascontiguousarray(arr)
    capi.check(capi.load().qfx_copy_h2d_async(ctx.handle, buf.ptr + offset, arr.ctypes.data,
                                              arr.nbytes))


def test_staged_uploads_later_copy_into_pending_bytes_wins(ctx):
    """Staged copies wait in a batch until the next call that puts work on
    the stream; a copy into bytes the batch still has to write starts a new
    batch, so the values of the later call win."""
    b = ctx.alloc(4096)
    _upload_at(ctx, b, 0, np.full(512, 1, np.uint64))
    _upload_at(ctx, b, 1024, np.full(16, 2, np.uint64))
    _upload_at(ctx, b, 1000, np.full(4, 3, np.uint64))  # overlaps both
    want = np.full(512, 1, np.uint64)
    want[128:144] = 2
    want_b = want.view(np.uint8).copy()
    want_b[1000:1032] = np.full(4, 3, np.uint64).view(np.uint8)
    assert np.array_equal(ctx.download(b, np.uint8, (4096,)), want_b)
    b.free()


def test_staged_uploads_every_alignment_in_one_batch(ctx):
    """16-byte, 8-byte and byte-aligned copies of one batch, each whole."""
    b16, b8, b1 = ctx.alloc(4096), ctx.alloc(256), ctx.alloc(256)
    x16 = np.arange(512, dtype=np.uint64) * 3
    x8 = np.array([7, 8, 9], np.uint64)
    x1 = np.arange(1, 14, dtype=np.uint8)
    _upload_at(ctx, b16, 0, x16)
    _upload_at(ctx, b8, 8, x8)
    _upload_at(ctx, b1, 3, x1)
    ctx.zero(b8, 8)  # a launch: the batch goes first
    assert np.array_equal(ctx.download(b16, np.uint64, (512,)), x16)
    got8 = ctx.download(b8, np.uint64, (4,))
    assert got8[0] == 0 and np.array_equal(got8[1:], x8)
    assert np.array_equal(ctx.download(b1, np.uint8, (16,))[3:16], x1)
    for b in (b16, b8, b1):
        b.free()


def test_staged_uploads_one_launch_per_batch(ctx):
    bufs = [ctx.alloc(128) for _ in range(200)]
    ctx.profile(True)
    ctx.profile_reset()
    for i, b in enumerate(bufs):
        ctx.upload_async(b, np.full(16, i, np.uint64))
    ctx.sync()
    prof = ctx.profile_read()
    ctx.profile(False)
    assert prof["copy"]["launches"] == 1 and prof["copy"]["bytes"] == 2.0 * 200 * 128
    for i, b in enumerate(bufs):
        assert np.all(ctx.download(b, np.uint64, (16,)) == i), i
        b.free()


def test_staged_upload_into_freed_block_then_reused(ctx):
    """A block freed with its copy still batched and handed out again: the
    new owner's values are the ones that stay."""
    b = ctx.alloc(2048)
    ptr = b.ptr
    _upload_at(ctx, b, 0, np.full(256, 5, np.uint64))
    b.free()
    c = ctx.alloc(2048)
    assert c.ptr == ptr  # the arena hands the block out again
    _upload_at(ctx, c, 0, np.full(256, 6, np.uint64))
    assert np.all(ctx.download(c, np.uint64, (256,)) == 6)
    ctx.zero(c)
    assert np.all(ctx.download(c, np.uint64, (256,)) == 0)
    c.free()
