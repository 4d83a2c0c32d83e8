"""Several final states of one circuit (simulate's "final_states" list).

The reference builds and flattens the network once per EvaluateCircuit call,
i.e. once per final state (src/evaluate_circuit.cpp:160-217), and its
simulate() accepts one final state (src/pybind_main.cpp:112-113).
EvaluateCircuitStates builds and flattens the network once, with an open
output index on every qubit whose bit differs between the states, and
contracts each state from zero-copy projections of those sites' tensors: the
amplitudes must equal the per-state runs (to FP32 rounding: the flattened site
tensors are formed in another order) and the FP64 state vector."""
import numpy as np
import pytest

from oracle import statevector
from oracle.formats import Circuit
from qflex_amd import circuits, qflex

pytestmark = pytest.mark.gpu


def run(circ, order, grid, init, finals, precision="fp32", slice_batch=None):
    opts = {"circuit": circ.splitlines(), "ordering": order.splitlines(),
            "grid": grid.splitlines(), "initial_state": init, "final_states": finals,
            "precision": precision, "memory_limit": "200GB"}
    if slice_batch is not None:
        opts["slice_batch"] = slice_batch
    return qflex.evaluate(opts)


def relerr(got, ref):
    got, ref = np.array(got), np.array(ref)
    return float(np.linalg.norm(got - ref) / np.linalg.norm(ref))


@pytest.mark.parametrize("precision", ["fp32", "split_fp16"])
def test_terminal_region_many_states_vs_state_vector(precision):
    I = J = 4
    circ = circuits.random_circuit(I, J, 16, 3)
    region = [(3, 1), (3, 2), (3, 3)]
    order = circuits.snake_ordering(I, J, final_region=region)
    grid = circuits.grid_text(I, J)
    rng = np.random.default_rng(7)
    states = []
    for _ in range(6):
        b = list("".join(str(x) for x in rng.integers(0, 2, I * J)))
        for (i, j) in region:
            b[i * J + j] = "x"
        states.append("".join(b))
    many = run(circ, order, grid, "0" * 16, states, precision)
    stats = qflex.last_stats()
    assert len(many) == 6 * 8 and stats["slices"] == 6
    one = []
    for s in states:
        one += run(circ, order, grid, "0" * 16, [s], precision)
    assert [s for s, _ in many] == [s for s, _ in one]
    assert relerr([a for _, a in many], [a for _, a in one]) < 1e-6
    psi = statevector.final_state_vector(Circuit(circ), I, J, [], "0" * 16)
    ref = [psi[tuple(int(c) for c in s)] for s, _ in many]
    assert relerr([a for _, a in many], ref) < 1e-5


@pytest.mark.parametrize("slice_batch", [None, [1, 4]])
def test_corner_inward_7x7_many_states(slice_batch):
    """Two bond cuts, six terminal cuts and three base bitstrings (192
    amplitudes from one network build), batched cuts or not."""
    circ = circuits.random_circuit(7, 7, 16, 5)
    order = circuits.corner_inward_7x7()
    grid = circuits.grid_text(7, 7)
    rng = np.random.default_rng(11)
    states = ["".join(str(x) for x in rng.integers(0, 2, 49)) for _ in range(3)]
    many = run(circ, order, grid, "0" * 49, states, "split_fp16", slice_batch)
    assert len(many) == 3 * 64
    one = []
    for s in states:
        one += run(circ, order, grid, "0" * 49, [s], "split_fp16", slice_batch)
    assert [s for s, _ in many] == [s for s, _ in one]
    assert relerr([a for _, a in many], [a for _, a in one]) < 1e-5
