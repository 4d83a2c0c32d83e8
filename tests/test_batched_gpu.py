"""Bond cuts contracted a batch of values at a time (simulate's "slice_batch").

The reference contracts one cut value at a time and sums the slices' amplitudes
in FP64 (ContractionData::ContractGrid, src/contraction_utils.cpp:283-297 and
:352).  A batch keeps the cut index, restricted to the batch's values, as a
batch index: operands that carry it on one side become one taller (or wider)
product, operands that carry it on both sides one product per batch value, and
work that does not depend on the cut runs once per batch instead of once per
value.  The slices' amplitudes are still summed in FP64, one slice at a time in
value order (qfx_accumulate_batched_c128).

The batched products have other shapes than the per-slice ones (the batch
index makes m or n larger, and can turn a matrix-vector product into a
matrix product), so they may run on other kernels with other summation
orders: the amplitudes agree with the one-value-at-a-time run to FP32
rounding and meet the same error bars against the oracle."""
import json
import os

import numpy as np
import pytest

from oracle.contraction import evaluate_circuit as oracle_evaluate
from qflex_amd import circuits, qflex

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def run(circ, order, grid, init, final, precision, slice_batch=None):
    opts = {"circuit": circ.splitlines(), "ordering": order.splitlines(),
            "grid": grid.splitlines(), "initial_state": init, "final_state": final,
            "precision": precision, "memory_limit": "200GB"}
    if slice_batch is not None:
        opts["slice_batch"] = slice_batch
    return qflex.evaluate(opts)


def relerr(got, ref):
    got, ref = np.array(got), np.array(ref)
    return float(np.linalg.norm(got - ref) / np.linalg.norm(ref))


SUBSET_ORDER = "\n".join(["cut (0,1) 1 4", "cut () 3 4", "expand a 0", "expand a 1", "expand a 3",
                          "cut (1,0) 8", "cut () 7", "expand b 8", "expand b 7", "expand b 5",
                          "expand b 2", "expand b 6", "expand b 4", "merge a b"])


@pytest.mark.parametrize("slice_batch", [2, 4, [2, 1], [1, 4], [2, 3]])
@pytest.mark.parametrize("precision", ["fp32", "fp32_exact"])
def test_subset_cuts_batched(slice_batch, precision):
    """Value subsets, a terminal cut between the bond cuts, uneven batches."""
    circ = circuits.random_circuit(3, 3, 16, 6)
    grid = circuits.grid_text(3, 3)
    seq = run(circ, SUBSET_ORDER, grid, "0" * 9, "000000000", precision)
    seq_stats = qflex.last_stats()
    got = run(circ, SUBSET_ORDER, grid, "0" * 9, "000000000", precision, slice_batch)
    stats = qflex.last_stats()
    assert [s for s, _ in got] == [s for s, _ in seq]
    assert relerr([a for _, a in got], [a for _, a in seq]) < 2e-6, (slice_batch, got, seq)
    assert stats["slices"] == seq_stats["slices"]
    assert stats["multiplies"] < seq_stats["multiplies"]
    o = oracle_evaluate(circ, SUBSET_ORDER, grid, "0" * 9, "000000000")
    assert relerr([a for _, a in got], [a for _, a in o]) < 1e-5


@pytest.mark.parametrize("slice_batch", [4, [2, 4], [4, 1], [1, 4], [3, 2]])
@pytest.mark.parametrize("precision", ["fp32_exact", "fp32", "split_fp16", "split_bf16x3"])
def test_corner_inward_7x7_d16_batched(slice_batch, precision):
    """The headline ordering at bond 4 (two bond cuts of 4 values, six
    terminal cuts): batched runs against the sequential run and the oracle."""
    circ = circuits.random_circuit(7, 7, 16, 11)
    order = circuits.corner_inward_7x7()
    grid = circuits.grid_text(7, 7)
    seq = run(circ, order, grid, "0" * 49, "", precision)
    seq_stats = qflex.last_stats()
    got = run(circ, order, grid, "0" * 49, "", precision, slice_batch)
    stats = qflex.last_stats()
    print(f"7x7 d16 {precision} batch {slice_batch}: multiplies {stats['multiplies']} "
          f"(sequential {seq_stats['multiplies']}), flops {stats['flops']:.3e} "
          f"({seq_stats['flops']:.3e})")
    assert [s for s, _ in got] == [s for s, _ in seq]
    assert stats["slices"] == seq_stats["slices"] == 16
    assert stats["flops"] <= seq_stats["flops"] * (1 + 1e-12)
    o = oracle_evaluate(circ, order, grid, "0" * 49, "")
    ref = [a for _, a in o]
    assert relerr([a for _, a in got], ref) < 1e-5
    assert relerr([a for _, a in got], [a for _, a in seq]) < 1e-5


def headline(slices):
    ref = json.load(open(os.path.join(GOLDEN, f"headline_7x7_d32_seed0_slices{slices}.json")))
    cut_values = tuple(ref["cut_values"]) if ref.get("cut_values") else (None, None)
    want = np.array([complex(re, im) for _, re, im in ref["amplitudes"]])
    return circuits.random_circuit(7, 7, 32, 0), circuits.corner_inward_7x7(cut_values), ref, want


@pytest.mark.parametrize("precision", ["fp32", "split_fp16"])
def test_headline_16_slices_one_batch(precision):
    """The 7x7 depth-32 workload (bond 16) with its inner cut's 16 values as
    one batch: the stacked products exceed 2 GiB per operand (the largest is
    2^29 complex64 elements), against the complex128 fixture."""
    circ, order, ref, want = headline(16)
    got = run(circ, order, circuits.grid_text(7, 7), "0" * 49, "", precision, [1, 16])
    stats = qflex.last_stats()
    d = dict(got)
    err = relerr([d[s] for s, _, _ in ref["amplitudes"]], want)
    print(f"7x7 d32 16 slices in one batch, {precision}: relerr {err:.3e}, "
          f"{stats['multiplies']} multiplies")
    assert stats["slices"] == 16
    assert err <= 1e-5


def test_headline_full_job_batched_split_fp16():
    """All 256 slices with the inner cut batched (16 batches of 16 values):
    the bench's batched configuration, against the complex128 fixture."""
    circ, order, ref, want = headline(256)
    got = run(circ, order, circuits.grid_text(7, 7), "0" * 49, "", "split_fp16", [1, 16])
    stats = qflex.last_stats()
    d = dict(got)
    err = relerr([d[s] for s, _, _ in ref["amplitudes"]], want)
    print(f"7x7 d32 256 slices, batches of 16: relerr {err:.3e}, {stats['multiplies']} "
          f"multiplies, {stats['seconds']:.3f} s")
    assert stats["slices"] == 256
    assert err <= 1e-5


def test_accumulate_batched_matches_sequential_accumulates(ctx):
    """qfx_accumulate_batched_c128 == nbatch calls of qfx_accumulate_c128 (same
    FP64 additions in the same order), bit for bit, at an offset."""
    from qflex_amd import capi
    rng = np.random.default_rng(3)
    n, nb, off = 1000, 7, 24
    src = (rng.standard_normal((nb, n)) + 1j * rng.standard_normal((nb, n))).astype(np.complex64)
    base = rng.standard_normal(n + off) + 1j * rng.standard_normal(n + off)
    S = ctx.upload(src)
    A1, A2 = ctx.upload(base.astype(np.complex128)), ctx.upload(base.astype(np.complex128))
    ctx.accumulate_batched(S, n, nb, A1, off)
    for b in range(nb):
        part = ctx.upload(src[b])
        ctx.accumulate(part, n, A2, off)
        part.free()
    got = ctx.download(A1, np.complex128, (n + off,))
    want = ctx.download(A2, np.complex128, (n + off,))
    assert np.array_equal(got, want)
    assert np.array_equal(got[:off], base[:off])
    for buf in (S, A1, A2):
        buf.free()
    with pytest.raises(capi.QfxError):
        ctx.accumulate_batched(S, n, nb, A1, off)  # null (freed) buffers are refused
