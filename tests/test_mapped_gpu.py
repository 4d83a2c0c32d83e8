"""GPU parity of qfx_cgemm_mapped_c64: contractions read straight from the
stored index orders of both tensors (no reorder), against numpy complex128
of the reordered matrices (the reference's reorder-then-cgemm,
src/tensor.cpp:1100-1128)."""
import numpy as np
import pytest

from qflex_amd import capi

pytestmark = pytest.mark.gpu

TOL = {capi.FP32: 2e-6, capi.FP32_EXACT: 2e-6, capi.SPLIT_BF16X3: 5e-6, capi.SPLIT_BF16: 3e-4,
       capi.SPLIT_FP16: 5e-6}


def make_case(left, common, right, a_inner, b_inner, seed):
    """Random storage orders of A (left + common axes) and B (common + right).

    left/common/right are lists of dimensions; a_inner / b_inner name the
    group ("left", "common", "right") whose last axis is stored innermost.
    Returns dims and the axis lists of the matrix views (most significant
    axis first; the common axes listed in the same order for A and B).
    """
    rng = np.random.default_rng(seed)
    names_l = [("l", i) for i in range(len(left))]
    names_c = [("c", i) for i in range(len(common))]
    names_r = [("r", i) for i in range(len(right))]
    dim = {**{n: d for n, d in zip(names_l, left)}, **{n: d for n, d in zip(names_c, common)},
           **{n: d for n, d in zip(names_r, right)}}
    groups = {"left": names_l, "common": names_c, "right": names_r}

    def order(names, inner):
        last = groups[inner][-1]
        rest = [n for n in names if n != last]
        return [rest[i] for i in rng.permutation(len(rest))] + [last]

    a_names = order(names_l + names_c, a_inner)
    b_names = order(names_c + names_r, b_inner)
    a_dims = [dim[n] for n in a_names]
    b_dims = [dim[n] for n in b_names]
    return (a_dims, [a_names.index(n) for n in names_l], [a_names.index(n) for n in names_c],
            b_dims, [b_names.index(n) for n in names_c], [b_names.index(n) for n in names_r])


def run(ctx, precs, case, integer=False, seed=0):
    a_dims, a_left, a_common, b_dims, b_common, b_right = case
    rng = np.random.default_rng(seed)
    if integer:
        a = (rng.integers(-8, 8, a_dims) + 1j * rng.integers(-8, 8, a_dims)).astype(np.complex64)
        b = (rng.integers(-8, 8, b_dims) + 1j * rng.integers(-8, 8, b_dims)).astype(np.complex64)
    else:
        a = (rng.standard_normal(a_dims) + 1j * rng.standard_normal(a_dims)).astype(np.complex64)
        b = (rng.standard_normal(b_dims) + 1j * rng.standard_normal(b_dims)).astype(np.complex64)
    m = int(np.prod([a_dims[x] for x in a_left]))
    k = int(np.prod([a_dims[x] for x in a_common]))
    n = int(np.prod([b_dims[x] for x in b_right]))
    am = a.transpose(a_left + a_common).reshape(m, k).astype(np.complex128)
    bm = b.transpose(b_common + b_right).reshape(k, n).astype(np.complex128)
    ref = am @ bm
    maps = (capi.index_map(a_dims, a_left), capi.index_map(a_dims, a_common),
            capi.index_map(b_dims, b_common), capi.index_map(b_dims, b_right))
    A, B, C = ctx.upload(a), ctx.upload(b), ctx.alloc(m * n * 8)
    out = {}
    for prec in precs:
        if not ctx.cgemm_mapped(prec, A, maps[0], maps[1], B, maps[2], maps[3], C):
            out[prec] = None
            continue
        got = ctx.download(C, np.complex64, (m, n))
        if integer:
            out[prec] = bool(np.array_equal(got, ref.astype(np.complex64)))
        else:
            out[prec] = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
    for buf in (A, B, C):
        buf.free()
    return out


SPLITS = [capi.SPLIT_FP16, capi.SPLIT_BF16, capi.SPLIT_BF16X3]

# (left, common, right, a_inner, b_inner): contraction patterns of grid
# patches (bond dimension 16, terminal outputs 2), at test sizes with at
# least 256 tiles.
BIG_CASES = [
    ([16, 16, 16, 16], [16, 2], [8, 16], "left", "right"),          # 256x128, A along m
    ([16, 16, 16, 16], [2, 16], [16, 8], "common", "right"),        # 256x128, A along k
    ([16, 16, 16, 4], [2, 16], [16, 2, 16], "common", "right"),     # 128x256, A along k
    ([16, 16, 4, 16], [16, 2], [2, 16, 16], "left", "right"),       # 128x256, A along m
    ([16, 16, 2, 16, 2, 16], [16, 2], [2, 16], "left", "right"),    # 512x32
]


@pytest.mark.parametrize("case", range(len(BIG_CASES)))
def test_mapped_large_tiles(ctx, case):
    res = run(ctx, SPLITS, make_case(*BIG_CASES[case], seed=case), seed=case)
    for prec, err in res.items():
        if err is None:
            # BF16x3 has no 512x32 tile (its three planes exceed the LDS)
            assert prec == capi.SPLIT_BF16X3 and case == 4, (case, prec)
            continue
        assert err <= TOL[prec], (case, prec, err)


def test_mapped_large_tiles_integer_inputs_bit_exact(ctx):
    for case in (0, 2, 4):
        res = run(ctx, SPLITS, make_case(*BIG_CASES[case], seed=case), integer=True, seed=case)
        for prec, ok in res.items():
            assert ok is not False, (case, prec)


SMALL_CASES = [
    ([16, 16, 16, 16], [16], [16], "left", "right"),         # k = 16, n = 16, A along m
    ([16, 16, 16], [16], [4, 16], "common", "common"),       # n = 64, A and B along k
    ([16, 32], [2, 4], [8], "left", "right"),                # k = 8, ragged tiles
    ([2, 2, 2], [16, 16, 16, 16, 16], [2, 2, 2], "common", "common"),  # 8 x 8 x 2^20
    ([16, 16], [16, 4], [16, 16], "common", "right"),        # 256 x 256 x 64, split-K
    ([16, 16], [4, 16], [16, 16], "left", "common"),         # same, A along m, B along k
]


@pytest.mark.parametrize("case", range(len(SMALL_CASES)))
def test_mapped_vector_kernels(ctx, case):
    precs = [capi.FP32, capi.FP32_EXACT, capi.SPLIT_FP16]
    res = run(ctx, precs, make_case(*SMALL_CASES[case], seed=20 + case), seed=case)
    for prec, err in res.items():
        if case >= 3 and prec == capi.FP32_EXACT:
            assert err is None  # the split-K forms have no defined order
            continue
        assert err is not None and err <= 2e-6, (case, prec, err)


def test_mapped_unsupported_forms(ctx):
    # FP32 precision at MFMA sizes: the caller reorders instead
    res = run(ctx, [capi.FP32], make_case(*BIG_CASES[0], seed=0))
    assert res[capi.FP32] is None
    # contiguous run of 2 elements along the read axis of A
    res = run(ctx, [capi.SPLIT_FP16], make_case([16, 16, 16, 16, 2], [16, 2], [8, 16], "left",
                                                "right", seed=5))
    assert res[capi.SPLIT_FP16] is None
    # at most 8 x 8 outputs with B stored along n (k's lowest digit at stride
    # 2^18): the loads along k would be scattered, so the caller reorders B
    res = run(ctx, [capi.SPLIT_FP16], make_case([2, 2, 2], [16, 16, 16, 16, 16], [2, 2, 2],
                                                "common", "right", seed=6))
    assert res[capi.SPLIT_FP16] is None



def test_scale_chain_matches_recomputed_scales(ctx):
    # GEMM 1 writes per-block maxima of |C1|; GEMM 2 reads C1 as A with
    # those maxima instead of a pass over C1: the same scale, so the same bits
    rng = np.random.default_rng(12)
    m, k1, k2, n = 65536, 32, 256, 128
    a = (rng.standard_normal((m, k1)) + 1j * rng.standard_normal((m, k1))).astype(np.complex64)
    b1 = (rng.standard_normal((k1, k2)) + 1j * rng.standard_normal((k1, k2))).astype(np.complex64)
    b2 = (rng.standard_normal((k2, n)) + 1j * rng.standard_normal((k2, n))).astype(np.complex64)
    A, B1, B2 = ctx.upload(a), ctx.upload(b1), ctx.upload(b2)
    C1, C2, C2ref = ctx.alloc(m * k2 * 8), ctx.alloc(m * n * 8), ctx.alloc(m * n * 8)
    cap = 4096
    CM = ctx.alloc(4 * cap)
    nat = lambda r, c: (capi.index_map([r, c], [0]), capi.index_map([r, c], [1]))
    io1 = capi.ScaleIO(None, 0, None, 0, CM.ptr, cap, 0)
    assert ctx.cgemm_mapped_ex(capi.SPLIT_FP16, A, *nat(m, k1), B1, *nat(k1, k2), C1, io1)
    assert 0 < io1.c_max_n <= cap
    cmax = ctx.download(CM, np.float32, (io1.c_max_n,))
    c1 = ctx.download(C1, np.complex64, (m, k2))
    assert cmax.max() == max(np.abs(c1.real).max(), np.abs(c1.imag).max())
    io2 = capi.ScaleIO(CM.ptr, io1.c_max_n, None, 0, None, 0, 0)
    assert ctx.cgemm_mapped_ex(capi.SPLIT_FP16, C1, *nat(m, k2), B2, *nat(k2, n), C2, io2)
    assert ctx.cgemm_mapped(capi.SPLIT_FP16, C1, *nat(m, k2), B2, *nat(k2, n), C2ref)
    got = ctx.download(C2, np.complex64, (m, n))
    ref = ctx.download(C2ref, np.complex64, (m, n))
    assert np.array_equal(got.view(np.uint64), ref.view(np.uint64))
    for buf in (A, B1, B2, C1, C2, C2ref, CM):
        buf.free()


@pytest.mark.parametrize("k1,n1,a_along_m", [(16, 16, False), (16, 256, False), (16, 16, True)])
def test_scale_chain_from_small_k_products(ctx, k1, n1, a_along_m):
    # The vector-ALU products (k <= 16: one thread per row for n <= 16, tiles
    # otherwise) also write per-block maxima of |C| for a split-FP16
    # consumer; the next product reads C with them: same scale, same bits.
    # a_along_m: A stored k-major, read one value per lane and k by the
    # row kernel that walks several 256-row tiles per block (2048 here).
    rng = np.random.default_rng(13 + n1 + a_along_m)
    m, n = 524288 if n1 == 16 else 32768, 128  # the second product needs >= 128 large tiles
    a = (rng.standard_normal((m, k1)) + 1j * rng.standard_normal((m, k1))).astype(np.complex64)
    b1 = (rng.standard_normal((k1, n1)) + 1j * rng.standard_normal((k1, n1))).astype(np.complex64)
    b2 = (rng.standard_normal((n1, n)) + 1j * rng.standard_normal((n1, n))).astype(np.complex64)
    A = ctx.upload(np.ascontiguousarray(a.T) if a_along_m else a)
    B1, B2 = ctx.upload(b1), ctx.upload(b2)
    C1, C2, C2ref = ctx.alloc(m * n1 * 8), ctx.alloc(m * n * 8), ctx.alloc(m * n * 8)
    cap = 8192
    CM = ctx.alloc(4 * cap)
    nat = lambda r, c: (capi.index_map([r, c], [0]), capi.index_map([r, c], [1]))
    a_maps = ((capi.index_map([k1, m], [1]), capi.index_map([k1, m], [0])) if a_along_m
              else nat(m, k1))
    io1 = capi.ScaleIO(None, 0, None, 0, CM.ptr, cap, 0)
    assert ctx.cgemm_mapped_ex(capi.SPLIT_FP16, A, *a_maps, B1, *nat(k1, n1), C1, io1)
    assert 0 < io1.c_max_n <= cap
    cmax = ctx.download(CM, np.float32, (io1.c_max_n,))
    c1 = ctx.download(C1, np.complex64, (m, n1))
    assert cmax.max() == max(np.abs(c1.real).max(), np.abs(c1.imag).max())
    if n1 == 16:  # one maximum per 256 rows of C
        per = np.maximum(np.abs(c1.real), np.abs(c1.imag)).reshape(-1, 256 * n1).max(axis=1)
        assert io1.c_max_n == m // 256 and np.array_equal(cmax, per)
    # C1 = A B1 exactly as the defined-order kernels compute it
    ref1 = a.astype(np.complex128) @ b1.astype(np.complex128)
    assert np.linalg.norm(c1 - ref1) / np.linalg.norm(ref1) < 2e-6
    io2 = capi.ScaleIO(CM.ptr, io1.c_max_n, None, 0, None, 0, 0)
    if n1 == 16:
        # m x 16 read as (m / 16) x 256: the next product's k covers whole rows of C1
        mm, kk = m // 16, 256
    else:
        mm, kk = m, n1
    b2k = np.resize(b2, (kk, n)).astype(np.complex64) if kk != n1 else b2
    B2k = ctx.upload(b2k)
    assert ctx.cgemm_mapped_ex(capi.SPLIT_FP16, C1, *nat(mm, kk), B2k, *nat(kk, n), C2, io2)
    assert ctx.cgemm_mapped(capi.SPLIT_FP16, C1, *nat(mm, kk), B2k, *nat(kk, n), C2ref)
    got = ctx.download(C2, np.complex64, (mm, n))
    ref = ctx.download(C2ref, np.complex64, (mm, n))
    assert np.array_equal(got.view(np.uint64), ref.view(np.uint64))
    for buf in (A, B1, B2, B2k, C1, C2, C2ref, CM):
        buf.free()


# k <= 16, n <= 16: one thread per row of C.  With A's k digits stored
# innermost and in order, the rows of A are read whole into LDS.
ROWS_CASES = [
    ([16, 16, 16, 16], [16], [16], "common", "right"),   # 2^16 x 16 x 16, A along k
    ([16, 16, 8], [2, 8], [8], "common", "common"),      # k in two digits, B along k
    ([16, 16, 16, 4], [16], [16], "left", "right"),      # A along m
    # 2^20 x 16 x 16, A along m: 4096 tiles of 256 rows, each block walks four
    ([16, 16, 16, 16, 16], [16], [16], "left", "right"),
]


@pytest.mark.parametrize("case", range(len(ROWS_CASES)))
def test_mapped_rows_bit_identical_to_defined_order(ctx, case):
    """Same fmaf chain per output as the scalar kernel: the reordered
    matrices with a zero column / row appended (k + 1 goes to the scalar
    kernel in FP32_EXACT) give the same bits."""
    a_dims, a_left, a_common, b_dims, b_common, b_right = make_case(*ROWS_CASES[case],
                                                                    seed=40 + case)
    rng = np.random.default_rng(40 + case)
    a = (rng.standard_normal(a_dims) + 1j * rng.standard_normal(a_dims)).astype(np.complex64)
    b = (rng.standard_normal(b_dims) + 1j * rng.standard_normal(b_dims)).astype(np.complex64)
    m = int(np.prod([a_dims[x] for x in a_left]))
    k = int(np.prod([a_dims[x] for x in a_common]))
    n = int(np.prod([b_dims[x] for x in b_right]))
    am = a.transpose(a_left + a_common).reshape(m, k)
    bm = b.transpose(b_common + b_right).reshape(k, n)
    maps = (capi.index_map(a_dims, a_left), capi.index_map(a_dims, a_common),
            capi.index_map(b_dims, b_common), capi.index_map(b_dims, b_right))
    A, B, C = ctx.upload(a), ctx.upload(b), ctx.alloc(m * n * 8)
    am1 = np.ascontiguousarray(np.concatenate([am, np.zeros((m, 1), np.complex64)], axis=1))
    bm1 = np.ascontiguousarray(np.concatenate([bm, np.zeros((1, n), np.complex64)], axis=0))
    A1, B1, C1 = ctx.upload(am1), ctx.upload(bm1), ctx.alloc(m * n * 8)
    for prec in (capi.FP32_EXACT, capi.SPLIT_FP16):
        assert ctx.cgemm_mapped(prec, A, maps[0], maps[1], B, maps[2], maps[3], C)
        ctx.cgemm(capi.FP32_EXACT, m, n, k + 1, A1, B1, C1)
        got = ctx.download(C, np.complex64, (m, n))
        want = ctx.download(C1, np.complex64, (m, n))
        assert np.array_equal(got, want), (case, prec)
    for buf in (A, B, C, A1, B1, C1):
        buf.free()
