"""Writes fixtures from the reference's own input files (this container only:
it reads the reference tree, which GPU runs do not have).

* reference_inputs.json -- input data of two reference workloads and the
  complex128 oracle's amplitudes for them:
    - the CI smoke of the reference (.travis.yml:16): circuit
      config/circuits/bristlecone_48_1-24-1_0.txt, ordering
      config/ordering/bristlecone_48.txt (bond cuts narrowed to some values,
      two open terminal cuts: 4 amplitudes), grid config/grid/bristlecone_48.txt,
      default initial and final states;
    - config/circuits/rectangular_4x4_1-16-1_0.txt with
      config/grid/rectangular_4x4.txt, for which the reference ships no
      ordering: a row-snake ordering, eight final bitstrings sampled from the
      FP64 state vector of the same circuit (which also pins the oracle).
  The circuit, ordering and grid are the reference's data files, kept as
  lists of lines so that a GPU test can feed them to simulate() as files.
* reference_circuit_structure.json -- digests of the gate positions of the
  reference's rectangular circuit files (per cycle: CZ pairs, and the qubits
  of H, of x_1_2/y_1_2 and of T gates), which
  qflex_amd.circuits.random_circuit must reproduce (only its seeded choice
  between x_1_2 and y_1_2 may differ), with gate counts.

    python tests/golden/make_reference_inputs.py
"""
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
REF = "/root/reference/config"

from oracle import statevector  # noqa: E402
from oracle.contraction import evaluate_circuit  # noqa: E402
from oracle.formats import Circuit  # noqa: E402
from qflex_amd import circuits  # noqa: E402


def read(*parts):
    with open(os.path.join(REF, *parts)) as f:
        return f.read().splitlines()


def structure(lines):
    """{cycle: {"cz": [[a, b]...], "h": [...], "xy": [...], "t": [...]}}"""
    out = {}
    for line in lines[1:]:
        w = line.split("#")[0].split()
        if not w:
            continue
        c, name, qs = int(w[0]), w[1], sorted(int(x) for x in w[2:])
        kind = {"cz": "cz", "h": "h", "t": "t", "x_1_2": "xy", "y_1_2": "xy"}[name]
        d = out.setdefault(str(c), {"cz": [], "h": [], "xy": [], "t": []})
        d[kind].append(qs if kind == "cz" else qs[0])
    for d in out.values():
        for v in d.values():
            v.sort()
    return out


def structure_digest(st):
    """sha256 of the structure as canonical JSON (sorted keys, no spaces)."""
    import hashlib
    text = json.dumps(st, sort_keys=True, separators=(",", ":"))
    return hashlib.sha256(text.encode()).hexdigest()


def main():
    t0 = time.time()
    fixtures = {}
    circ = read("circuits", "bristlecone_48_1-24-1_0.txt")
    order = read("ordering", "bristlecone_48.txt")
    grid = read("grid", "bristlecone_48.txt")
    amps = evaluate_circuit("\n".join(circ), "\n".join(order), "\n".join(grid), "", "",
                            dtype=np.complex128)
    fixtures["bristlecone_48_1-24-1"] = {
        "source": ["config/circuits/bristlecone_48_1-24-1_0.txt",
                   "config/ordering/bristlecone_48.txt", "config/grid/bristlecone_48.txt"],
        "circuit": circ, "ordering": order, "grid": grid,
        "final_state": "",
        "amplitudes": [[s, a.real, a.imag] for s, a in amps]}
    print("bristlecone_48 1-24-1:", len(amps), "amplitudes", time.time() - t0, "s")

    circ = read("circuits", "rectangular_4x4_1-16-1_0.txt")
    grid = read("grid", "rectangular_4x4.txt")
    order = circuits.snake_ordering(4, 4).splitlines()
    psi = statevector.final_state_vector(Circuit("\n".join(circ)), 4, 4, [], "0" * 16)
    bitstrings = circuits.sample_bitstrings(psi, 8, 0)
    rows = []
    for b in bitstrings:
        (s, a), = evaluate_circuit("\n".join(circ), "\n".join(order), "\n".join(grid), "", b,
                                   dtype=np.complex128)
        sv = psi[tuple(int(x) for x in b)]
        assert abs(a - sv) <= 1e-12 * np.abs(psi).max(), (b, a, sv)
        rows.append([s, a.real, a.imag])
    fixtures["rectangular_4x4_1-16-1"] = {
        "source": ["config/circuits/rectangular_4x4_1-16-1_0.txt", "(row-snake ordering)",
                   "config/grid/rectangular_4x4.txt"],
        "circuit": circ, "ordering": order, "grid": grid,
        "amplitudes": rows}
    print("rectangular_4x4 1-16-1: oracle = state vector on", len(rows), "bitstrings")
    json.dump({"_what": "reference input files and complex128 oracle amplitudes",
               "_made_by": "tests/golden/make_reference_inputs.py", **fixtures},
              open(os.path.join(os.path.dirname(__file__), "reference_inputs.json"), "w"),
              indent=0)

    struct = {}
    for name in sorted(os.listdir(os.path.join(REF, "circuits"))):
        if not name.startswith("rectangular_") or name.endswith("2x2_1-2-1_0.txt"):
            continue
        shape, depth = name[len("rectangular_"):].split("_")[:2]
        I, J = (int(x) for x in shape.split("x"))
        d = int(depth.split("-")[1])
        st = structure(read("circuits", name))
        struct[name] = {"I": I, "J": J, "depth": d, "sha256": structure_digest(st),
                        "counts": {k: sum(len(c[k]) for c in st.values())
                                   for k in ("cz", "h", "xy", "t")}}
    json.dump({"_what": "gate positions of the reference's rectangular circuit files "
                        "(sha256 of structure_digest's canonical JSON)",
               "_made_by": "tests/golden/make_reference_inputs.py", "files": struct},
              open(os.path.join(os.path.dirname(__file__),
                                "reference_circuit_structure.json"), "w"), indent=1)
    print("structure of", len(struct), "rectangular circuits")


if __name__ == "__main__":
    main()
