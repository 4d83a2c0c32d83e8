"""End-to-end parity of the C++ host + HIP path (module qflex) against the
oracle (complex128 restatement) and the FP64 state vector, restating the
reference's own Tensor / ContractGrid / EvaluateCircuit tests."""
import json
import os

import numpy as np
import pytest

from oracle import statevector
from oracle.contraction import evaluate_circuit as oracle_evaluate
from oracle.formats import Circuit
from oracle.tensor_ops import brute_force_reorder
from qflex_amd import circuits, qflex

pytestmark = pytest.mark.gpu

KATS = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_kats.json")))


def relerr(x, ref):
    x, ref = np.asarray(x), np.asarray(ref)
    return float(np.linalg.norm(x - ref) / np.linalg.norm(ref))


# ---------------------------------------------------------------- Tensor API

def test_tensor_zero_init_and_rename():
    t = qflex.Tensor(["a", "b"], [2, 4])
    assert t.size() == 8 and t.num_zeros() == 8
    t.rename_index("a", "c")
    assert t.get_indices() == ["c", "b"]


def test_reorder_known_answer():
    k = KATS["simple_index_reordering"]
    t = qflex.Tensor(k["indices"], k["dims"], np.array(k["data_real"], dtype=np.complex64))
    t.reorder(k["new_indices"])
    assert t.get_indices() == k["new_indices"]
    assert t.data().real.tolist() == k["expected_real"]


@pytest.mark.parametrize("case", [c for c in KATS["reorder_cases"] if not c.startswith("_")])
def test_reorder_reference_cases_bit_exact(case):
    c = KATS["reorder_cases"][case]
    old, new = list(c["indices"]), list(c["new"])
    n = len(old)
    vals = np.arange(2 ** n).astype(np.complex64)
    t = qflex.Tensor(old, [2] * n, vals)
    t.reorder(new)
    assert t.data().real.astype(int).tolist() == brute_force_reorder(list(range(2 ** n)), n, old, new)


def test_multiply_known_answer():
    k = KATS["multiply"]
    A = qflex.Tensor(k["a_indices"], [2, 2, 2], np.array(k["a_data_real"], dtype=np.complex64))
    B = qflex.Tensor(k["b_indices"], [2, 2, 2], np.array(k["b_data_real"], dtype=np.complex64))
    C = qflex.Tensor(["x"], [16])
    qflex.multiply(A, B, C)
    assert C.get_indices() == k["c_indices"] and C.get_dimensions() == k["c_dims"]
    assert C.data().real.tolist() == k["expected_real"]


def test_multiply_random_vs_einsum():
    rng = np.random.default_rng(0)
    a = (rng.standard_normal((4, 8, 2, 16)) + 1j * rng.standard_normal((4, 8, 2, 16))).astype(np.complex64)
    b = (rng.standard_normal((16, 8, 32)) + 1j * rng.standard_normal((16, 8, 32))).astype(np.complex64)
    A = qflex.Tensor(list("pqrs"), list(a.shape), a)
    B = qflex.Tensor(list("sqt"), list(b.shape), b)
    C = qflex.Tensor(["c"], [4 * 2 * 32])
    qflex.multiply(A, B, C)
    assert C.get_indices() == ["p", "r", "t"]
    ref = np.einsum("pqrs,sqt->prt", a.astype(np.complex128), b)
    assert relerr(C.data().reshape(ref.shape), ref) < 1e-6
    # A and B were reordered in place, as the reference does
    assert A.get_indices() == ["p", "r", "q", "s"] and B.get_indices() == ["q", "s", "t"]


def test_project_bundle_capacity():
    t = qflex.Tensor(["a", "b", "c"], [2, 2, 2], np.arange(8).astype(np.complex64))
    p = qflex.Tensor(["x"], [4])
    t.project("a", 1, p)
    assert p.get_indices() == ["b", "c"] and p.data().real.tolist() == [4, 5, 6, 7]
    t.bundle(["a", "b"], "ab")
    assert t.get_indices() == ["ab", "c"] and t.get_dimensions() == [4, 2]
    big = qflex.Tensor([""], [64])
    big.set_indices_and_dimensions(["a", "b"], [4, 4])
    big.set_indices_and_dimensions(["k", "m", "n"], [2, 2, 2])
    small = qflex.Tensor(["f", "g"], [16, 16])
    with pytest.raises(RuntimeError, match="The total allocated space: 256, is insufficient for "
                                           "the requested tensor dimensions: 4096."):
        small.assign(qflex.Tensor(["f", "g", "h"], [16, 16, 16]))


def test_tensor_error_texts():
    with pytest.raises(RuntimeError, match="The number of indices: 3, and number of dimensions: 2, "
                                           "should be equal."):
        qflex.Tensor(["a", "b", "c"], [2, 2])
    with pytest.raises(RuntimeError, match="The vector data size: 8, has to match the size of the "
                                           "Tensor: 4"):
        qflex.Tensor(["a", "b"], [2, 2], np.zeros(8, dtype=np.complex64))
    abc = qflex.Tensor(["a", "b", "c"], [2, 2, 2])
    ac = qflex.Tensor(["a", "c"], [2, 2])
    checks = [
        (lambda: abc.project("b", 0, ac), "Index: 'b' has to be equal to indices[0]: 'a'."),
        (lambda: abc.project("a", 2, ac), "index_value: 2 must be contained in [0, 2)."),
        (lambda: abc.project("a", 0, qflex.Tensor(["a", "c"], [2, 1])),
         "The total allocated space: 2, is insufficient for the requested tensor dimensions: 4."),
        (lambda: abc.rename_index("x", "y"), "old_name: x, has to be a valid index."),
        (lambda: abc.rename_index("a", "b"), "new_name: b, cannot be an existing index."),
        (lambda: abc.bundle(["a", "x"], "ax"),
         "indices_to_bundle: {a, x} has to be contained in indices: {a, b, c}."),
        (lambda: abc.bundle(["b", "a"], "ba"),
         "indices_to_bundle: {b, a} must be in its original order: {a, b}."),
        (lambda: abc.reorder(["b", "a"]),
         "new_ordering: {b, a} must be a reordering of current indices: {a, b, c}."),
        (lambda: abc.reorder(["b", "y", "x"]),
         "new_ordering: {b, y, x} must be a reordering of current indices: {a, b, c}."),
        (lambda: abc.reorder(["a", "b"], null_scratch=True), "Scratch copy must be non-null."),
    ]
    cd = qflex.Tensor(["c", "d"], [2, 2])
    abd = qflex.Tensor(["a", "b", "d"], [2, 2, 2])
    checks += [
        (lambda: qflex.multiply(abc, cd, abc),
         "A and C cannot be the same tensor: Tensor of rank 3: a -> 2, b -> 2, c -> 2"),
        (lambda: qflex.multiply(abc, cd, cd),
         "B and C cannot be the same tensor: Tensor of rank 2: c -> 2, d -> 2"),
        (lambda: qflex.multiply(abc, qflex.Tensor(["c", "d"], [4, 4]), abd),
         "Common indices must have matching dimensions, but at index: 0, the dimensions are A: 2, B: 4."),
        (lambda: qflex.multiply(abc, cd, qflex.Tensor(["x"], [2])),
         "C: 2 doesn't have enough space for the product of A*B: 8."),
        (lambda: qflex.multiply(abc, cd, qflex.Tensor(["x"], [2]), null_scratch=True),
         "Scratch copy must be non-null."),
    ]
    for fn, msg in checks:
        with pytest.raises(RuntimeError) as e:
            fn()
        assert msg in str(e.value), (msg, str(e.value))


# ---------------------------------------------------------------- contraction

def test_simple_initialize_data_known_answer():
    g = [[None, None] for _ in range(3)]
    I4, I8, I2x4 = (np.eye(4), np.eye(8), np.eye(2, 4))
    g[0][0] = qflex.Tensor(["(0,0),(0,1)", "(0,0),(1,0)"], [4, 4], I4.astype(np.complex64))
    g[0][1] = qflex.Tensor(["(0,1),(1,1)", "(0,0),(0,1)"], [4, 4], I4.astype(np.complex64))
    g[1][0] = qflex.Tensor(["(0,0),(1,0)", "(1,0),(1,1)"], [4, 4], I4.astype(np.complex64))
    g[1][1] = qflex.Tensor(["(0,1),(1,1)", "(1,1),(2,1)", "(1,0),(1,1)"], [4, 4, 4],
                           I8.astype(np.complex64))
    g[2][1] = qflex.Tensor(["(2,1),(o)", "(1,1),(2,1)"], [2, 4], I2x4.astype(np.complex64))
    ops = [("cut", [[0, 1], [1, 1]], []), ("expand", "a", [0, 1]), ("expand", "a", [0, 0]),
           ("expand", "a", [1, 0]), ("cut", [[2, 1]], []), ("expand", "b", [2, 1]),
           ("expand", "b", [1, 1]), ("merge", "a", "b")]
    amps = qflex.contract_grid(ops, g, 2)
    assert np.allclose(amps, [1.0, 0.0], atol=1e-6)


def test_simple_circuit_known_answer():
    k = KATS["simple_circuit"]
    for prec in ("fp32", "fp32_exact", "split_bf16x3", "split_fp16"):
        r = qflex.evaluate({"circuit": k["circuit"], "ordering": k["ordering"], "grid": k["grid"],
                            "initial_state": k["initial_state"], "final_state": k["final_state"],
                            "precision": prec})
        assert [s for s, _ in r] == [e[0] for e in k["expected"]]
        for (_, amp), (_, re, im) in zip(r, k["expected"]):
            assert abs(amp.real - re) < k["tolerance"] and abs(amp.imag - im) < k["tolerance"]


def test_condense_to_grid_index_orders():
    c = KATS["condense_to_grid"]
    ops = [("cut", o[1], []) if o[0] == "cut" else tuple(o) for o in c["ordering"]]
    got = qflex.flatten_indices(c["circuit"], c["I"], c["J"], c["initial"], c["final"],
                                c["final_region"], c["off"], ops)
    assert got == c["expected"]


def run_engine(circ, order, grid, initial, final, precision="fp32"):
    return qflex.evaluate({"circuit": circ.splitlines(), "ordering": order.splitlines(),
                           "grid": grid.splitlines(), "initial_state": initial,
                           "final_state": final, "precision": precision,
                           "memory_limit": 1 << 36})


TOL = {"fp32": 1e-5, "fp32_exact": 1e-5, "split_bf16x3": 1e-5, "split_bf16": 3e-3,
       "split_fp16": 1e-5}


@pytest.mark.parametrize("I,J,depth,seed", [(3, 3, 8, 1), (4, 4, 16, 2), (5, 4, 16, 3)])
@pytest.mark.parametrize("precision", ["fp32", "fp32_exact", "split_bf16x3", "split_bf16",
                                       "split_fp16"])
def test_snake_vs_state_vector(I, J, depth, seed, precision):
    circ = circuits.random_circuit(I, J, depth, seed)
    psi = statevector.final_state_vector(Circuit(circ), I, J, [], "0" * (I * J))
    bs = circuits.sample_bitstrings(psi, 4, seed)
    got, ref = [], []
    for b in bs:
        r = run_engine(circ, circuits.snake_ordering(I, J), circuits.grid_text(I, J),
                       "0" * (I * J), b, precision)
        got.append(r[0][1])
        ref.append(psi[tuple(int(x) for x in b)])
    assert relerr(got, ref) <= TOL[precision], relerr(got, ref)


def test_config1_abcd_layout_vs_state_vector():
    # 5x4 grid, ABCDCDAB with a one-qubit layer on every qubit, depth 16:
    # bonds of dimension 16, 16 bitstrings sampled from |psi|^2.
    I, J = 5, 4
    circ = circuits.random_circuit(I, J, 16, 0, layout="abcd")
    psi = statevector.final_state_vector(Circuit(circ), I, J, [], "0" * 20)
    bs = circuits.sample_bitstrings(psi, 16, 0)
    got = [run_engine(circ, circuits.snake_ordering(I, J), circuits.grid_text(I, J), "0" * 20, b)[0][1]
           for b in bs]
    ref = [psi[tuple(int(x) for x in b)] for b in bs]
    assert relerr(got, ref) <= 1e-5


def test_terminal_region_all_outputs_vs_oracle():
    I, J = 3, 3
    circ = circuits.random_circuit(I, J, 16, 4)
    region = [(2, 1), (2, 2), (1, 2)]
    order = circuits.snake_ordering(I, J, final_region=region)
    r = run_engine(circ, order, circuits.grid_text(I, J), "", "")
    o = oracle_evaluate(circ, order, circuits.grid_text(I, J))
    assert [s for s, _ in r] == [s for s, _ in o]
    assert relerr([a for _, a in r], [a for _, a in o]) < 1e-5


def test_subset_terminal_values_and_cuts_vs_oracle():
    I, J = 3, 3
    circ = circuits.random_circuit(I, J, 16, 6)
    order = "\n".join(["cut (0,1) 1 4", "cut () 3 4", "expand a 0", "expand a 1", "expand a 3",
                       "cut (1,0) 8", "cut () 7", "expand b 8", "expand b 7", "expand b 5",
                       "expand b 2", "expand b 6", "expand b 4", "merge a b"])
    r = run_engine(circ, order, circuits.grid_text(I, J), "0" * 9, "000000000")
    o = oracle_evaluate(circ, order, circuits.grid_text(I, J), "0" * 9, "000000000")
    assert [s for s, _ in r] == [s for s, _ in o]
    assert relerr([a for _, a in r], [a for _, a in o]) < 1e-5


def test_off_qubits_and_merges_vs_state_vector():
    I, J = 3, 4
    off = [(0, 3), (2, 0)]
    circ = circuits.random_circuit(I, J, 16, 8, off=off)
    grid = circuits.grid_text(I, J, off)
    order = "\n".join(["expand a 0", "expand a 1", "expand a 4", "expand a 5",
                       "expand b 11", "expand b 10", "expand b 7", "expand b 6",
                       "expand c 2", "expand c 9", "merge b c", "merge a c"])
    psi = statevector.final_state_vector(Circuit(circ), I, J, off, "0" * 10)
    for b in circuits.sample_bitstrings(psi, 3, 1):
        r = run_engine(circ, order, grid, "0" * 10, b)
        assert abs(r[0][1] - psi[tuple(int(x) for x in b)]) < 1e-5 * max(abs(psi).max(), 1e-3)


@pytest.mark.parametrize("depth", [8, 16])
def test_corner_inward_7x7_ordering_vs_oracle(depth):
    # The headline ordering (2 cuts, 6 terminal cuts) at small bond dims.
    circ = circuits.random_circuit(7, 7, depth, 11)
    order = circuits.corner_inward_7x7()
    grid = circuits.grid_text(7, 7)
    r = run_engine(circ, order, grid, "0" * 49, "")
    o = oracle_evaluate(circ, order, grid, "0" * 49, "")
    assert len(r) == 64
    assert [s for s, _ in r] == [s for s, _ in o]
    assert relerr([a for _, a in r], [a for _, a in o]) < 1e-5


@pytest.mark.skip(reason="engine-level FP8 is disabled: its first run on this workload ended "
                         "in a GPU memory fault whose cause is not found yet")
def test_fp8_low_fidelity_mode_vs_complex128_reference():
    """The FP8 mode on one slice of the 7x7, depth 1+32+1 workload (bonds of
    dimension 256 after bundling, so the large products run in FP8), against
    the committed complex128 oracle amplitudes of that slice: amplitudes
    within a few percent, fidelity |<ref|psi>|^2 / (|ref|^2 |psi|^2) close
    to 1, and the quantization visibly at work (not an FP32 result)."""
    ref = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                      "headline_7x7_d32_seed0_slices1.json")))
    c = ref["circuit"]
    circ = circuits.random_circuit(c["I"], c["J"], c["depth"], c["seed"])
    order = circuits.corner_inward_7x7(tuple(ref["cut_values"]))
    r = dict(run_engine(circ, order, circuits.grid_text(7, 7), "0" * 49, "", "fp8"))
    got = np.array([r[s] for s, _, _ in ref["amplitudes"]])
    want = np.array([complex(re, im) for _, re, im in ref["amplitudes"]])
    fid = abs(np.vdot(want, got)) ** 2 / (np.vdot(want, want).real * np.vdot(got, got).real)
    err = relerr(got, want)
    print("fp8 7x7 d32 one slice: relerr", err, "fidelity", fid)
    assert 1e-4 < err < FP8_MAX_RELERR and fid > FP8_MIN_FIDELITY


FP8_MAX_RELERR, FP8_MIN_FIDELITY = 0.5, 0.75


@pytest.mark.parametrize("precision", ["split_fp16", "split_bf16", "fp32"])
def test_mapped_operands_bit_identical_to_reorders(precision):
    # Reading operands in stored order (no reorder) feeds the same values in
    # the same k order to the same kernels: amplitudes must match bit for bit.
    circ = circuits.random_circuit(7, 7, 16, 5)
    order = circuits.corner_inward_7x7()
    grid = circuits.grid_text(7, 7)
    try:
        qflex.set_mapped_operands(False)
        plain = run_engine(circ, order, grid, "0" * 49, "", precision)
        plain_stats = qflex.last_stats()
    finally:
        qflex.set_mapped_operands(True)
    mapped = run_engine(circ, order, grid, "0" * 49, "", precision)
    stats = qflex.last_stats()
    assert plain_stats["mapped_multiplies"] == 0
    if precision != "fp32":
        assert stats["mapped_multiplies"] > 0
        assert stats["reorder_elements"] < plain_stats["reorder_elements"]
    assert [s for s, _ in mapped] == [s for s, _ in plain]
    assert all(a == b for (_, a), (_, b) in zip(mapped, plain))


def test_memory_limit_error():
    circ = circuits.random_circuit(3, 3, 16, 1)
    with pytest.raises(RuntimeError, match=r"Required space \(.*\) exceeds memory limit \(1 kB\)"):
        qflex.evaluate({"circuit": circ.splitlines(), "ordering": circuits.snake_ordering(3, 3).splitlines(),
                        "grid": circuits.grid_text(3, 3).splitlines(), "memory_limit": 1024})


def test_simulate_with_files(tmp_path):
    k = KATS["simple_circuit"]
    for name in ("circuit", "ordering", "grid"):
        (tmp_path / name).write_text("\n".join(k[name]) + "\n")
    r = qflex.simulate({"circuit_filename": str(tmp_path / "circuit"),
                        "ordering_filename": str(tmp_path / "ordering"),
                        "grid_filename": str(tmp_path / "grid"),
                        "initial_state": "00000", "final_state": "1100x"})
    assert [s for s, _ in r] == ["11000", "11001"]
    assert abs(r[0][1] - complex(0.10669, 0.04419)) < 1e-5


def test_fp8_refused_by_the_engine():
    """Engine-level FP8 is switched off: simulate reports the error and
    returns no amplitudes (the reference's error convention)."""
    circ = circuits.random_circuit(3, 3, 8, 1)
    with pytest.raises(RuntimeError, match="disabled"):
        run_engine(circ, circuits.snake_ordering(3, 3), circuits.grid_text(3, 3), "0" * 9,
                   "0" * 9, "fp8")
