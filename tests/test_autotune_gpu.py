"""GPU tests of the tile autotuner (qfx_autotune_*): every tile gives the same
bits, the tuned table holds a valid tile per product, and a whole
contraction with tuning on reproduces the untuned amplitudes exactly."""
import numpy as np
import pytest

from qflex_amd import capi, circuits, qflex

from test_mapped_gpu import make_case

pytestmark = pytest.mark.gpu

SPLITS = [capi.SPLIT_FP16, capi.SPLIT_BF16]


def rand_c(rng, shape):
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)


@pytest.mark.parametrize("prec", SPLITS)
def test_every_tile_same_bits_row_major(ctx, prec):
    m, n, k = 16384, 256, 256  # all three tiles can run it (>= 128 blocks of 256 x 128)
    rng = np.random.default_rng(1)
    a, b = rand_c(rng, (m, k)), rand_c(rng, (k, n))
    A, B, C = ctx.upload(a), ctx.upload(b), ctx.alloc(m * n * 8)
    outs = {}
    try:
        for tile in (capi.TILE_256x128, capi.TILE_128x256, capi.TILE_512x32):
            ctx.autotune_set(prec, capi.FORM_ROW_MAJOR, m, n, k, tile)
            ctx.autotune(True)
            ctx.cgemm(prec, m, n, k, A, B, C)
            outs[tile] = ctx.download(C, np.complex64, (m, n))
    finally:
        ctx.autotune(False)
        ctx.autotune_set(prec, capi.FORM_ROW_MAJOR, m, n, k, 0)
    ref = a.astype(np.complex128) @ b.astype(np.complex128)
    base = outs[capi.TILE_256x128]
    assert np.linalg.norm(base - ref) / np.linalg.norm(ref) < 5e-4
    for tile, got in outs.items():
        assert np.array_equal(got, base), tile
    for buf in (A, B, C):
        buf.free()


@pytest.mark.parametrize("inner", ["left", "common"])
def test_tuning_mapped_product_same_bits_and_table(ctx, inner):
    a_dims, a_left, a_common, b_dims, b_common, b_right = make_case(
        [16, 16, 16, 16], [16, 16], [16, 16], inner, "right", seed=3)
    rng = np.random.default_rng(3)
    a, b = rand_c(rng, a_dims), rand_c(rng, b_dims)
    m, k, n = 65536, 256, 256
    maps = (capi.index_map(a_dims, a_left), capi.index_map(a_dims, a_common),
            capi.index_map(b_dims, b_common), capi.index_map(b_dims, b_right))
    A, B, C = ctx.upload(a), ctx.upload(b), ctx.alloc(m * n * 8)
    form = capi.FORM_A_ALONG_K if inner == "common" else capi.FORM_A_ALONG_M
    try:
        assert ctx.cgemm_mapped(capi.SPLIT_FP16, A, *maps[:2], B, *maps[2:], C)
        plain = ctx.download(C, np.complex64, (m, n))
        ctx.autotune(True)
        assert ctx.cgemm_mapped(capi.SPLIT_FP16, A, *maps[:2], B, *maps[2:], C)
        tuned = ctx.download(C, np.complex64, (m, n))
        entries = [e for e in ctx.autotune_entries() if e[:5] == (capi.SPLIT_FP16, form, m, n, k)]
        assert len(entries) == 1 and entries[0][5] in (1, 2, 3), ctx.autotune_entries()
        # a second call uses the table (no new entry, same bits)
        assert ctx.cgemm_mapped(capi.SPLIT_FP16, A, *maps[:2], B, *maps[2:], C)
        again = ctx.download(C, np.complex64, (m, n))
    finally:
        ctx.autotune(False)
        ctx.autotune_set(capi.SPLIT_FP16, form, m, n, k, 0)
    assert np.array_equal(plain, tuned) and np.array_equal(plain, again)
    for buf in (A, B, C):
        buf.free()


def test_autotune_set_rejects_tiles_that_cannot_run():
    c = capi.Context(0)
    try:
        with pytest.raises(RuntimeError):
            c.autotune_set(capi.SPLIT_BF16X3, capi.FORM_ROW_MAJOR, 4096, 256, 256,
                           capi.TILE_128x256)  # BF16x3 has the 256 x 128 tile only
        with pytest.raises(RuntimeError):
            c.autotune_set(capi.SPLIT_FP16, capi.FORM_ROW_MAJOR, 4096, 256, 256, 7)
        c.autotune_set(capi.SPLIT_FP16, capi.FORM_ROW_MAJOR, 4096, 256, 256, capi.TILE_128x256)
        assert c.autotune_entries() == [(capi.SPLIT_FP16, 0, 4096, 256, 256, 2)]
    finally:
        c.close()


def test_engine_with_autotune_reproduces_untuned_amplitudes():
    """One slice of the 7x7, depth 1+32+1 workload (its products are large
    enough for the large tiles)."""
    circ = circuits.random_circuit(7, 7, 32, 0)
    order = circuits.corner_inward_7x7(([0], [0]))
    opts = {"circuit": circ.splitlines(), "ordering": order.splitlines(),
            "grid": circuits.grid_text(7, 7).splitlines(), "initial_state": "0" * 49,
            "final_state": "", "memory_limit": 1 << 36, "precision": "split_fp16"}
    plain = qflex.evaluate(dict(opts, autotune=False))
    try:
        tuned = qflex.evaluate(dict(opts, autotune=True))
        table = qflex.autotune_table()
        again = qflex.evaluate(dict(opts, autotune=True))
    finally:
        qflex.set_autotune(False)
    assert table, "no product was tuned"
    assert all(t[5] in (1, 2, 3) for t in table)
    assert [s for s, _ in tuned] == [s for s, _ in plain]
    assert all(x == y for (_, x), (_, y) in zip(tuned, plain))
    assert all(x == y for (_, x), (_, y) in zip(again, plain))
