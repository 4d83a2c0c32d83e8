"""A queue of uneven grid circuits (6x6, 7x7 corner-inward, 8x7 with greedy
orderings and cuts), each with several base bitstrings and a six-qubit open
region, run through simulate() with "final_states" and batched cuts, against
the complex128 oracle state by state (at depth 1+16+1, bond 4, where the
oracle takes seconds)."""
import numpy as np
import pytest

from oracle.contraction import evaluate_circuit as oracle_evaluate
from qflex_amd import jobqueue, qflex

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["fp32", "split_fp16"])
def test_small_queue_vs_oracle(precision):
    for job in jobqueue.queue(states=3, depths=[16, 16, 16]):
        opts = dict(job["options"], precision=precision)
        got = qflex.evaluate(opts)
        assert len(got) == 3 * 64, job["name"]
        o = opts
        want = []
        for final in o["final_states"]:
            want += oracle_evaluate("\n".join(o["circuit"]) + "\n", "\n".join(o["ordering"]) + "\n",
                                    "\n".join(o["grid"]) + "\n", o["initial_state"], final)
        assert [s for s, _ in got] == [s for s, _ in want], job["name"]
        g = np.array([a for _, a in got])
        w = np.array([a for _, a in want])
        err = float(np.linalg.norm(g - w) / np.linalg.norm(w))
        print(f"{job['name']} at depth {job['depth']}, {precision}: relerr {err:.2e}")
        assert err < 1e-5, (job["name"], err)
