"""Recorded cut passes (simulate's "record_slices"): the innermost bond cut's
first pass runs eagerly, the second is recorded as one HIP graph of all its
launches, and every later pass copies the cut tensors' projections into the
same fixed buffers and replays that graph.  The amplitudes must equal the
eager run's (the recorded passes run the same kernels on the same shapes; the
eager run reads the projections in place, so its kernels may see other
operand alignments: FP32 rounding), and the work counters must count the
replayed passes."""
import json
import os

import numpy as np
import pytest

from qflex_amd import circuits, qflex

pytestmark = pytest.mark.gpu


def run(circ, order, precision, record, slice_batch=None):
    opts = {"circuit": circ.splitlines(), "ordering": order.splitlines(),
            "grid": circuits.grid_text(7, 7).splitlines(), "initial_state": "0" * 49,
            "final_state": "", "precision": precision, "memory_limit": "200GB",
            "record_slices": record}
    if slice_batch is not None:
        opts["slice_batch"] = slice_batch
    return qflex.evaluate(opts), qflex.last_stats()


def relerr(got, ref):
    got, ref = np.array(got), np.array(ref)
    return float(np.linalg.norm(got - ref) / np.linalg.norm(ref))


@pytest.mark.parametrize("precision", ["fp32", "split_fp16"])
@pytest.mark.parametrize("slice_batch", [None, [2, 1]])
def test_recorded_passes_match_eager(precision, slice_batch):
    """7x7 depth 1+16+1 (two bond cuts of 4 values): the inner cut's 4 passes
    (per outer pass) are eager, recorded, replayed, replayed."""
    circ = circuits.random_circuit(7, 7, 16, 11)
    order = circuits.corner_inward_7x7()
    eager, st_e = run(circ, order, precision, False, slice_batch)
    rec, st_r = run(circ, order, precision, True, slice_batch)
    outer = 4 if slice_batch is None else 2
    assert st_e["replayed_passes"] == 0
    assert st_r["replayed_passes"] == outer * 2
    for key in ("slices", "multiplies", "mapped_multiplies"):
        assert st_r[key] == st_e[key], key
    assert abs(st_r["flops"] - st_e["flops"]) <= 1e-12 * st_e["flops"]
    assert [s for s, _ in rec] == [s for s, _ in eager]
    assert relerr([a for _, a in rec], [a for _, a in eager]) < 1e-6


def test_recorded_passes_headline_16_slices():
    """The 7x7 depth 1+32+1 workload restricted to one value of the outer
    cut: 16 inner passes (1 eager, 1 recorded, 14 replayed) against the
    complex128 fixture."""
    ref = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                      "headline_7x7_d32_seed0_slices16.json")))
    circ = circuits.random_circuit(7, 7, 32, 0)
    order = circuits.corner_inward_7x7(tuple(ref["cut_values"]))
    got, st = run(circ, order, "split_fp16", True)
    assert st["replayed_passes"] == 14 and st["slices"] == 16
    d = dict(got)
    want = np.array([complex(re, im) for _, re, im in ref["amplitudes"]])
    err = relerr([d[s] for s, _, _ in ref["amplitudes"]], want)
    print(f"7x7 d32, 16 recorded inner passes: relerr {err:.3e}")
    assert err <= 1e-5


def test_no_recording_while_profiling():
    circ = circuits.random_circuit(7, 7, 16, 11)
    order = circuits.corner_inward_7x7()
    qflex.profile_enable(True)
    try:
        _, st = run(circ, order, "fp32", True)
    finally:
        qflex.profile_enable(False)
    assert st["replayed_passes"] == 0
