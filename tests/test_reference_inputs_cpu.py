"""The reference's own inputs on the CPU side: the seeded generator places
every gate where the reference's rectangular circuit files
(config/circuits/rectangular_*_0.txt) do, and the FP64 state vector of the
reference's 4x4 circuit reproduces the committed oracle amplitudes."""
import importlib.util
import json
import os

import numpy as np
import pytest

from oracle import statevector
from oracle.formats import Circuit
from qflex_amd import circuits

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
_spec = importlib.util.spec_from_file_location(
    "make_reference_inputs", os.path.join(GOLDEN, "make_reference_inputs.py"))
refgen = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(refgen)

STRUCT = json.load(open(os.path.join(GOLDEN, "reference_circuit_structure.json")))["files"]
INPUTS = json.load(open(os.path.join(GOLDEN, "reference_inputs.json")))


@pytest.mark.parametrize("name", sorted(STRUCT))
def test_generator_places_gates_like_the_reference_files(name):
    f = STRUCT[name]
    st = refgen.structure(circuits.random_circuit(f["I"], f["J"], f["depth"], 0).splitlines())
    counts = {k: sum(len(c[k]) for c in st.values()) for k in ("cz", "h", "xy", "t")}
    assert counts == f["counts"]
    assert refgen.structure_digest(st) == f["sha256"]


def test_state_vector_of_the_reference_4x4_circuit():
    fx = INPUTS["rectangular_4x4_1-16-1"]
    psi = statevector.final_state_vector(Circuit("\n".join(fx["circuit"])), 4, 4, [], "0" * 16)
    for s, re, im in fx["amplitudes"]:
        assert abs(psi[tuple(int(x) for x in s)] - complex(re, im)) <= 1e-12


def test_reference_inputs_parse():
    from qflex_amd import qflex
    for key in ("bristlecone_48_1-24-1", "rectangular_4x4_1-16-1"):
        fx = INPUTS[key]
        n, depth, gates = qflex.load_circuit(fx["circuit"])
        I, J, off = qflex.load_grid(fx["grid"])
        assert n == I * J - len(off)
        ops = qflex.parse_ordering(fx["ordering"], fx["grid"])
        qflex.validate_ordering(ops)
    assert len(INPUTS["bristlecone_48_1-24-1"]["amplitudes"]) == 4


def test_module_installs_into_a_qflexcirq_package(tmp_path):
    """The reference's simulator imports the binding as
    ``from qflexcirq import qflex``; installed into a package of that name,
    this build's module imports and exposes simulate()."""
    import subprocess
    import sys
    from qflex_amd import build
    pkg = tmp_path / "qflexcirq"
    pkg.mkdir()
    (pkg / "__init__.py").write_text("")
    build.install_into_package(str(pkg))
    code = ("from qflexcirq import qflex; import sys; "
            "sys.exit(0 if callable(qflex.simulate) else 1)")
    r = subprocess.run([sys.executable, "-c", code], cwd=str(tmp_path), capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
