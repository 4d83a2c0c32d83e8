"""The reference's own input files through the drop-in entry point on the
GPU: simulate() with the dict the reference's Cirq simulator builds
(qflexcirq/interface/qflex_simulator.py:46-53: circuit, ordering and grid
file names plus one final bitstring), against the complex128 oracle's
amplitudes of the same files (tests/golden/make_reference_inputs.py).

* bristlecone_48, depth 1+24+1, with the reference's ordering: the CI smoke
  of the reference (.travis.yml:16).  Its bond cuts are narrowed to some of
  their values and two terminal cuts stay open: 4 amplitudes, partial sums
  over the chosen slices.
* rectangular_4x4, depth 1+16+1 (a row-snake ordering; the oracle
  amplitudes there equal the FP64 state vector)."""
import json
import os

import numpy as np
import pytest

from qflex_amd import qflex

pytestmark = pytest.mark.gpu

INPUTS = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                     "reference_inputs.json")))


def write_inputs(tmp_path, fx):
    names = {}
    for part in ("circuit", "ordering", "grid"):
        p = tmp_path / f"{part}.txt"
        p.write_text("\n".join(fx[part]) + "\n")
        names[part] = str(p)
    return names


def relerr(got, want):
    got, want = np.asarray(got), np.asarray(want)
    return float(np.linalg.norm(got - want) / np.linalg.norm(want))


def test_bristlecone_smoke_through_simulate(tmp_path):
    fx = INPUTS["bristlecone_48_1-24-1"]
    f = write_inputs(tmp_path, fx)
    options = {"circuit_filename": f["circuit"], "ordering_filename": f["ordering"],
               "grid_filename": f["grid"], "final_state": "0" * 48}
    got = qflex.simulate(options)
    assert [s for s, _ in got] == [s for s, _, _ in fx["amplitudes"]]
    want = [complex(re, im) for _, re, im in fx["amplitudes"]]
    assert relerr([a for _, a in got], want) <= 1e-5


@pytest.mark.parametrize("precision", ["split_fp16", "split_bf16x3", "fp32_exact"])
def test_bristlecone_smoke_other_precisions(tmp_path, precision):
    fx = INPUTS["bristlecone_48_1-24-1"]
    f = write_inputs(tmp_path, fx)
    got = qflex.evaluate({"circuit_filename": f["circuit"], "ordering_filename": f["ordering"],
                          "grid_filename": f["grid"], "precision": precision})
    want = [complex(re, im) for _, re, im in fx["amplitudes"]]
    assert relerr([a for _, a in got], want) <= 1e-5


def test_reference_4x4_circuit_through_simulate(tmp_path):
    fx = INPUTS["rectangular_4x4_1-16-1"]
    f = write_inputs(tmp_path, fx)
    got, want = [], []
    for s, re, im in fx["amplitudes"]:
        r = qflex.simulate({"circuit_filename": f["circuit"], "ordering_filename": f["ordering"],
                            "grid_filename": f["grid"], "final_state": s})
        assert len(r) == 1 and r[0][0] == s
        got.append(r[0][1])
        want.append(complex(re, im))
    assert relerr(got, want) <= 1e-5
