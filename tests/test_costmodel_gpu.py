"""The cost model on the GPU: its features come from the engine's symbolic
plan (checked against the flops the engine then executes), every candidate
contracts to the same amplitudes, and a small trained model's choices are
no slower than the plain ordering on held-out circuits."""
import numpy as np
import pytest

from qflex_amd import capi, circuits, costmodel, qflex

pytestmark = pytest.mark.gpu


def test_plan_matches_executed_contraction():
    rng = np.random.default_rng(4)
    circ = circuits.random_circuit(5, 5, 16, 7)
    for order in costmodel.candidate_set(5, 5, rng, 6):
        opts = costmodel.options(circ, order, 5, 5, "split_fp16")
        plan = qflex.plan(opts)
        qflex.evaluate(opts)
        stats = qflex.last_stats()
        assert stats["multiplies"] == sum(s["multiplicity"] for s in plan["steps"])
        assert abs(stats["flops"] - costmodel.plan_flops(plan)) <= 1e-9 * stats["flops"]
        assert stats["seconds"] > 0


def test_candidates_give_the_same_amplitude():
    # the model only changes speed: every ordering contracts to one value
    rng = np.random.default_rng(5)
    circ = circuits.random_circuit(4, 5, 16, 9)
    vals = []
    for order in costmodel.candidate_set(4, 5, rng, 8):
        r = qflex.evaluate(costmodel.options(circ, order, 4, 5, "fp32"))
        vals.append(r[0][1])
    vals = np.array(vals)
    assert np.max(np.abs(vals - vals[0])) <= 1e-5 * np.abs(vals[0])


def test_trained_model_not_slower_than_plain():
    class Args:
        train_circuits, candidates, heldout, heldout_candidates = 4, 16, 4, 10
        steps, precision, seed = 300, "split_fp16", 0
        shapes, max_log2_elems, max_flops = "7x7x24", 26, 3e13
    rep = costmodel.run(Args())
    assert rep["train_samples"] >= 20
    assert rep["fit_loss_first_last"][1] < rep["fit_loss_first_last"][0]
    # the chosen orderings, timed, against the plain (row snake) ordering
    assert rep["total_model_s"] <= 1.05 * rep["total_plain_s"], rep


@pytest.mark.parametrize("I,J,depth,seed", [(4, 5, 16, 9), (5, 5, 24, 3)])
def test_greedy_orderings_contract_exactly_and_as_planned(I, J, depth, seed):
    """The greedy family (no cut, greedy cuts by slice or by total cost):
    every ordering gives the plain ordering's amplitude, and the engine's
    planned and executed flops are 8 x the greedy's own step times summed
    over the slices."""
    from qflex_amd import greedy
    circ = circuits.random_circuit(I, J, depth, seed)
    ref = qflex.evaluate(costmodel.options(circ, costmodel.plain_ordering(I, J), I, J, "fp32"))
    for max_cuts, per_slice in ((0, True), (2, True), (2, False)):
        order, info = greedy.greedy_ordering(circ, max_cuts=max_cuts, per_slice=per_slice)
        opts = costmodel.options(circ, order, I, J, "fp32")
        plan = qflex.plan(opts)
        r = qflex.evaluate(opts)
        stats = qflex.last_stats()
        want = 8.0 * info["slice_time"] * info["slices"]
        assert costmodel.plan_flops(plan) == want
        assert abs(stats["flops"] - want) <= 1e-9 * want
        assert abs(r[0][1] - ref[0][1]) <= 1e-5 * abs(ref[0][1])


def test_low_fidelity_cuts_sum_the_chosen_slices():
    """A fidelity-narrowed cut evaluates a subset of its values: the result
    is the sum of the single-value slices."""
    from qflex_amd import greedy
    I, J = 4, 5
    circ = circuits.random_circuit(I, J, 16, 9)
    order, info = greedy.greedy_ordering(circ, max_cuts=1, fidelity=0.5)
    lines = order.splitlines()
    cut = next(l for l in lines if l.startswith("cut"))
    vals = [int(v) for v in cut.split()[1].strip("()").split(",")]
    assert info["fidelity"] == 0.5 and len(vals) == info["slices"]
    low = qflex.evaluate(costmodel.options(circ, order, I, J, "fp32"))[0][1]
    parts = []
    for v in vals:
        one = [f"cut ({v}) " + " ".join(cut.split()[2:]) if l == cut else l for l in lines]
        parts.append(qflex.evaluate(costmodel.options(circ, "\n".join(one), I, J, "fp32"))[0][1])
    assert abs(low - sum(parts)) <= 1e-5 * max(abs(low), 1e-12)
