"""A queue of grid circuits of uneven shapes, each with many base bitstrings.

Each job is one circuit with an open output region of six qubits (a 2 x 3 or
3 x 2 patch: 64 amplitudes per base bitstring) and a list of base bitstrings
for the other qubits.  A job is one simulate() call with "final_states": the
engine builds and flattens the network once and contracts every base
bitstring from projections of the flattened site tensors, with the bond cuts
of the job's ordering contracted a batch of values at a time ("slice_batch").

The queue mixes 6x6, 7x7 and 8x7 grids of the reference's circuit family
(circuits.random_circuit) at depths 1+24+1 (bond dimension 8 on every edge)
and 1+32+1 (bond 16).  The 7x7 job uses the corner-inward ordering of the
headline workload; the others use the greedy ordering and cut search
(greedy.greedy_ordering) with the region's terminal cuts first.
"""
import numpy as np

from qflex_amd import circuits, greedy

# name, grid, depth, ordering source, bond cuts of the greedy search, slice batches
JOBS = [
    ("6x6 depth 1+24+1", 6, 6, 24, "greedy", 1, [8]),
    ("7x7 depth 1+32+1", 7, 7, 32, "corner", 2, [1, 16]),
    ("8x7 depth 1+24+1", 8, 7, 24, "greedy", 3, [1, 1, 8]),
]


def region(I, J, source):
    """The open output region: the corner-inward ordering's own 3 x 2 patch,
    otherwise the 2 x 3 patch in the top-right corner."""
    if source == "corner":
        return list(circuits.CORNER_7X7_FINAL)
    return [(i, j) for i in range(2) for j in range(J - 3, J)]


def job(name, I, J, depth, source, cuts, slice_batch, states, seed=0, depth_override=None):
    """One queue entry: the simulate() options and the open region."""
    d = depth_override or depth
    circ = circuits.random_circuit(I, J, d, seed)
    reg = region(I, J, source)
    if source == "corner":
        order = circuits.corner_inward_7x7()
    else:
        order, _ = greedy.greedy_ordering(circ, max_cuts=cuts)
        order = "".join(f"cut () {i * J + j}\n" for (i, j) in reg) + order
    rng = np.random.default_rng(seed + 1000 * I + J)
    finals = []
    for _ in range(states):
        bits = list("".join(str(int(b)) for b in rng.integers(0, 2, I * J)))
        for (i, j) in reg:
            bits[i * J + j] = "x"
        finals.append("".join(bits))
    opts = {"circuit": circ.splitlines(), "ordering": order.splitlines(),
            "grid": circuits.grid_text(I, J).splitlines(), "initial_state": "0" * (I * J),
            "final_states": finals, "memory_limit": "200GB", "slice_batch": slice_batch}
    return {"name": name, "I": I, "J": J, "depth": d, "options": opts, "region": reg,
            "states": states}


def queue(states=16, seed=0, depths=None):
    """The jobs of JOBS with `states` base bitstrings each (depths may
    override each job's depth, e.g. for a small test)."""
    out = []
    for n, spec in enumerate(JOBS):
        out.append(job(*spec, states=states, seed=seed,
                       depth_override=depths[n] if depths else None))
    return out
