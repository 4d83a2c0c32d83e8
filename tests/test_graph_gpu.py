"""GPU tests of recorded sequences (qfx_graph_*, HIP stream capture): a
sequence of GEMMs, a reorder, a copy, a zero fill and an FP64 accumulation
is recorded once and replayed several times, with other work and
synchronizations between the replays; every replay must give the eager
results bit for bit."""
import numpy as np
import pytest

from qflex_amd import capi

pytestmark = pytest.mark.gpu


def rand_c(rng, shape):
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)


def test_recorded_sequence_replays_match_eager(ctx):
    rng = np.random.default_rng(31)
    m, n, k = 16384, 256, 256          # large-tile split-FP16 product
    a, b = rand_c(rng, (m, k)), rand_c(rng, (k, n))
    s = rand_c(rng, (16, 16))          # k = 16 product on the vector ALUs
    A, B, S = ctx.upload(a), ctx.upload(b), ctx.upload(s)
    C1, C2, C3, C4 = (ctx.alloc(m * n * 8) for _ in range(4))
    amps = ctx.alloc(64 * 16)
    other_a, other_c = ctx.upload(rand_c(rng, (512, 512))), ctx.alloc(512 * 512 * 8)

    def sequence():
        ctx.cgemm(capi.SPLIT_FP16, m, n, k, A, B, C1)          # C1 = A B
        ctx.reorder(C1, C2, [m // 16, 16, n], [1, 0, 2])       # C2 = C1 with rows permuted
        ctx.cgemm(capi.FP32, m * n // 16, 16, 16, C2, S, C3)   # C3 = C2 (as rows of 16) . S
        ctx.copy(C4, C3, m * n * 8)
        ctx.zero(amps)
        ctx.accumulate(C4, 64, amps)                            # amps = first 64 entries
        ctx.accumulate(C4, 64, amps)                            # twice: 2 x C4[:64]

    def read():
        return (ctx.download(C1, np.complex64, (m, n)), ctx.download(C4, np.complex64, (m, n)),
                ctx.download(amps, np.complex128, (64,)))

    sequence()
    ctx.sync()
    eager = read()
    ref = a.astype(np.complex128) @ b.astype(np.complex128)
    assert np.linalg.norm(eager[0] - ref) / np.linalg.norm(ref) < 1e-5
    assert np.array_equal(eager[2], 2 * eager[1].reshape(-1)[:64].astype(np.complex128))

    ctx.graph_begin(workspace_bytes=64 << 20)
    try:
        sequence()
    finally:
        graph = ctx.graph_end()
    try:
        for rep in range(3):
            for buf in (C1, C2, C3, C4, amps):
                ctx.zero(buf)
            ctx.cgemm(capi.SPLIT_FP16, 512, 512, 512, other_a, other_a, other_c)  # other work
            ctx.sync()
            graph.launch()
            ctx.sync()
            got = read()
            for g, e in zip(got, eager):
                assert np.array_equal(g, e), rep
    finally:
        graph.destroy()
    for buf in (A, B, S, C1, C2, C3, C4, amps, other_a, other_c):
        buf.free()


def test_synchronizing_calls_refused_while_recording(ctx):
    buf = ctx.alloc(1 << 12)
    ctx.graph_begin()
    try:
        with pytest.raises(RuntimeError, match="recorded"):
            ctx.sync()
        with pytest.raises(RuntimeError, match="recorded"):
            ctx.download(buf, np.uint8, (16,))
        ctx.zero(buf)  # a fill kernel is recorded
    finally:
        graph = ctx.graph_end()
    graph.launch()
    ctx.sync()
    assert not ctx.download(buf, np.uint8, (1 << 12,)).any()
    graph.destroy()
    buf.free()


def test_workspace_must_be_reserved_before_recording():
    """A product whose workspace exceeds the reservation fails while
    recording (instead of synchronizing to grow it)."""
    c = capi.Context(0)
    try:
        rng = np.random.default_rng(5)
        m, n, k = 256, 256, 4096      # split-K partials: 16 x 256 x 256 x 8 bytes = 8 MiB
        A, B, C = c.upload(rand_c(rng, (m, k))), c.upload(rand_c(rng, (k, n))), c.alloc(m * n * 8)
        c.graph_begin(workspace_bytes=0)
        try:
            with pytest.raises(RuntimeError, match="workspace"):
                c.cgemm(capi.SPLIT_FP16, m, n, k, A, B, C)
        finally:
            c.graph_end().destroy()
        c.cgemm(capi.SPLIT_FP16, m, n, k, A, B, C)  # eager: the workspace grows
        c.sync()
    finally:
        c.close()
