"""Writes complex128 reference amplitudes of the 7x7, depth 1+32+1 workload.

The numbers come from the oracle (numpy complex128 restatement of the
reference's EvaluateCircuit) and are committed so that GPU runs can state
their norm-wise relative error without re-running the CPU oracle:

    python tests/golden/make_headline_reference.py --slices 4    # ~1 min
    python tests/golden/make_headline_reference.py --slices 256  # full job

--slices S < 256 restricts the bond cuts to a subset of their values: the
amplitudes are then partial sums over S slices.  The subsets are stored
("cut_values"); amplitudes do not depend on the ordering beyond which
slices are summed, so any ordering with the same two cuts and the same
value subsets must reproduce them.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle.contraction import evaluate_circuit  # noqa: E402
from qflex_amd import circuits  # noqa: E402


def cut_subsets(n_slices):
    """Value subsets of the two bond cuts for an n_slices sub-job."""
    if n_slices >= 256:
        return None, None
    if n_slices < 16:
        return [0], list(range(n_slices))
    return list(range(n_slices // 16)), None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--slices", type=int, default=4)
    ap.add_argument("--seed", type=int, default=0)
    args = ap.parse_args()
    circ = circuits.random_circuit(7, 7, 32, args.seed)
    cut_values = cut_subsets(args.slices)
    order = circuits.corner_inward_7x7(cut_values)
    grid = circuits.grid_text(7, 7)
    t = time.time()
    r = evaluate_circuit(circ, order, grid, "0" * 49, "", dtype=np.complex128)
    out = {
        "_what": "complex128 oracle amplitudes of the 7x7 1+32+1 corner-inward workload",
        "_made_by": "tests/golden/make_headline_reference.py",
        "circuit": {"I": 7, "J": 7, "depth": 32, "seed": args.seed, "layout": "8pattern",
                    "gates": "xyt"},
        "slices": args.slices,
        "cut_values": list(cut_values),
        "ordering": order.splitlines(),
        "cpu_seconds": time.time() - t,
        "amplitudes": [[s, a.real, a.imag] for s, a in r],
    }
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)),
                        f"headline_7x7_d32_seed{args.seed}_slices{args.slices}.json")
    json.dump(out, open(path, "w"), indent=1)
    print(path, out["cpu_seconds"])


if __name__ == "__main__":
    main()
