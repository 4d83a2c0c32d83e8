"""Pins the oracle (CPU): reference known answers, brute force, the
independent FP64 state vector, and a planted mistake that must be caught."""
import itertools
import json
import os

import numpy as np
import pytest

from oracle import network, statevector
from oracle.contraction import (contract_grid, evaluate_circuit, get_output_states,
                                plan, validate_ordering)
from oracle.formats import Circuit
from oracle.network import circuit_to_network, flatten
from oracle.tensor_ops import OTensor, OracleError, brute_force_reorder, multiply
from qflex_amd import circuits

KATS = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_kats.json")))


def _ops(raw):
    out = []
    for op in raw:
        if op[0] == "cut":
            out.append(("cut", tuple(tuple(p) for p in op[1]), list(op[2]) if len(op) > 2 else []))
        elif op[0] == "expand":
            out.append(("expand", op[1], tuple(op[2])))
        else:
            out.append(tuple(op))
    return out


def test_reorder_known_answer():
    k = KATS["simple_index_reordering"]
    t = OTensor(k["indices"], k["dims"], k["data_real"])
    t.reorder(k["new_indices"])
    assert t.data.reshape(-1).real.tolist() == k["expected_real"]


@pytest.mark.parametrize("case", [c for c in KATS["reorder_cases"] if not c.startswith("_")])
def test_reorder_matches_brute_force(case):
    c = KATS["reorder_cases"][case]
    old, new = list(c["indices"]), list(c["new"])
    n = len(old)
    vals = list(range(2 ** n))
    t = OTensor(old, [2] * n, vals)
    t.reorder(new)
    expected = brute_force_reorder(vals, n, old, new)
    assert t.data.reshape(-1).real.astype(int).tolist() == expected
    assert t.indices == new


def test_multiply_known_answer():
    k = KATS["multiply"]
    A = OTensor(k["a_indices"], [2, 2, 2], k["a_data_real"])
    B = OTensor(k["b_indices"], [2, 2, 2], k["b_data_real"])
    C = multiply(A, B)
    assert C.indices == k["c_indices"] and C.dims == k["c_dims"]
    assert C.data.reshape(-1).real.tolist() == k["expected_real"]


def test_multiply_closed_form():
    rng = np.random.default_rng(0)
    a = rng.standard_normal((2, 3, 4, 5)) + 1j * rng.standard_normal((2, 3, 4, 5))
    b = rng.standard_normal((5, 3, 6)) + 1j * rng.standard_normal((5, 3, 6))
    A = OTensor(list("pqrs"), a.shape, a)
    B = OTensor(list("sqt"), b.shape, b)
    C = multiply(A, B)
    assert C.indices == ["p", "r", "t"]
    assert np.allclose(C.data, np.einsum("pqrs,sqt->prt", a, b))


def test_simple_initialize_data_known_answer():
    # 3x2 grid of identities, one cut, one terminal cut -> amplitudes [1, 0]
    I4 = np.eye(4)
    I8 = np.eye(8)
    I2x4 = np.eye(2, 4)
    g = [[None] * 2 for _ in range(3)]
    g[0][0] = OTensor(["(0,0),(0,1)", "(0,0),(1,0)"], [4, 4], I4)
    g[0][1] = OTensor(["(0,1),(1,1)", "(0,0),(0,1)"], [4, 4], I4)
    g[1][0] = OTensor(["(0,0),(1,0)", "(1,0),(1,1)"], [4, 4], I4)
    g[1][1] = OTensor(["(0,1),(1,1)", "(1,1),(2,1)", "(1,0),(1,1)"], [4, 4, 4], I8)
    g[2][1] = OTensor(["(2,1),(o)", "(1,1),(2,1)"], [2, 4], I2x4)
    ops = [("cut", ((0, 1), (1, 1)), []), ("expand", "a", (0, 1)), ("expand", "a", (0, 0)),
           ("expand", "a", (1, 0)), ("cut", ((2, 1),), []), ("expand", "b", (2, 1)),
           ("expand", "b", (1, 1)), ("merge", "a", "b")]
    amps = np.zeros(2, dtype=np.complex128)
    contract_grid(ops, g, amps)
    exp = KATS["simple_initialize_data"]["expected"]
    assert np.allclose(amps, [complex(*e) for e in exp])


def test_simple_circuit_known_answer():
    k = KATS["simple_circuit"]
    for dt in (np.complex128, np.complex64):
        r = evaluate_circuit("\n".join(k["circuit"]), "\n".join(k["ordering"]),
                             "\n".join(k["grid"]), k["initial_state"], k["final_state"], dtype=dt)
        for (s, amp), (es, re, im) in zip(r, k["expected"]):
            assert s == es
            assert abs(amp.real - re) < k["tolerance"] and abs(amp.imag - im) < k["tolerance"]


@pytest.mark.parametrize("name", [n for n in KATS["get_output_states"] if not n.startswith("_")])
def test_get_output_states(name):
    c = KATS["get_output_states"][name]
    base, fq, states = get_output_states(c["I"], c["J"], [tuple(o) for o in c["off"]], "",
                                         c["final"], _ops(c["ops"]))
    assert [list(q) for q in fq] == c["final_qubits"]
    assert states == c["states"]


def test_condense_to_grid_index_orders():
    c = KATS["condense_to_grid"]
    ops = _ops(c["ordering"])
    circ = Circuit("\n".join(c["circuit"]))
    g3 = circuit_to_network(circ, c["I"], c["J"], c["initial"], c["final"],
                            [tuple(x) for x in c["final_region"]], [tuple(x) for x in c["off"]],
                            np.complex128)
    g2 = flatten(g3, c["I"], c["J"], [tuple(x) for x in c["final_region"]],
                 [tuple(x) for x in c["off"]], ops)
    for i in range(c["I"]):
        for j in range(c["J"]):
            got = g2[i][j].indices if g2[i][j] is not None else []
            assert got == c["expected"][i][j]


def test_two_qubit_splits_reproduce_textbook_gates():
    for name in ("cz", "cx"):
        q1, q2, dims = network.gate_arrays(name, [])
        a = np.array(q1).reshape(dims)  # (in, v, out)
        b = np.array(q2).reshape(dims)
        # U[(o1,o2),(i1,i2)] = sum_v a[i1,v,o1] b[i2,v,o2]
        u = np.einsum("avc,bvd->cdab", a, b).reshape(4, 4)
        assert np.allclose(u, statevector.two_qubit(name, []), atol=1e-12), name
    for th, ph in [(0.5, 1 / 6), (0.25, 0.1)]:
        q1, q2, dims = network.gate_arrays("fsim", [th, ph])
        a = np.array(q1).reshape(dims)
        b = np.array(q2).reshape(dims)
        u = np.einsum("avc,bvd->cdab", a, b).reshape(4, 4)
        assert np.allclose(u, statevector.two_qubit("fsim", [th, ph]), atol=1e-12)


def test_one_qubit_tables_are_transposed_textbook_gates():
    for name in ("h", "t", "x_1_2", "y_1_2", "hz_1_2"):
        t = np.array(network.gate_array(name, [])).reshape(2, 2)
        assert np.allclose(t.T, statevector.one_qubit(name, []), atol=1e-12), name
    t = np.array(network.gate_array("rz", [0.3])).reshape(2, 2)
    assert np.allclose(t.T, statevector.one_qubit("rz", [0.3]))


SV_CASES = [(3, 3, 8, "xyt", 1), (4, 4, 8, "uniform", 2), (3, 4, 16, "xyt", 3), (2, 3, 24, "xyt", 4)]


@pytest.mark.parametrize("I,J,depth,gates,seed", SV_CASES)
def test_tensor_network_matches_state_vector(I, J, depth, gates, seed):
    circ = circuits.random_circuit(I, J, depth, seed, gates=gates)
    c = Circuit(circ)
    psi = statevector.final_state_vector(c, I, J, [], "0" * (I * J))
    for b in circuits.sample_bitstrings(psi, 3, seed):
        r = evaluate_circuit(circ, circuits.snake_ordering(I, J), circuits.grid_text(I, J),
                             "0" * (I * J), b)
        ref = psi[tuple(int(x) for x in b)]
        assert abs(r[0][1] - ref) <= 1e-12 * max(1.0, abs(ref)) + 1e-13


def test_terminal_cuts_enumerate_all_outputs_and_normalise():
    # 2x3 grid, every qubit a terminal cut: 64 amplitudes = the full state,
    # which must match the state vector and have unit norm.
    I, J = 2, 3
    circ = circuits.random_circuit(I, J, 16, 9)
    region = [(i, j) for i in range(I) for j in range(J)]
    o = circuits.snake_ordering(I, J, final_region=region)
    r = evaluate_circuit(circ, o, circuits.grid_text(I, J))
    psi = statevector.final_state_vector(Circuit(circ), I, J, [], "0" * 6)
    amps = np.array([a for _, a in r])
    assert len(r) == 64
    assert abs(np.sum(np.abs(amps) ** 2) - 1.0) < 1e-12
    for s, a in r:
        assert abs(a - psi[tuple(int(x) for x in s)]) < 1e-12


def test_cut_ordering_matches_state_vector():
    # cuts and two patches, as in the reference's simulation tests
    I, J = 3, 3
    circ = circuits.random_circuit(I, J, 16, 5)
    order = "\n".join(["cut () 1 4", "cut () 3 4", "expand a 0", "expand a 1", "expand a 3",
                       "expand b 8", "expand b 7", "expand b 5", "expand b 2", "expand b 6",
                       "expand b 4", "merge a b"])
    psi = statevector.final_state_vector(Circuit(circ), I, J, [], "0" * 9)
    for b in circuits.sample_bitstrings(psi, 3, 1):
        r = evaluate_circuit(circ, order, circuits.grid_text(I, J), "0" * 9, b)
        assert abs(r[0][1] - psi[tuple(int(x) for x in b)]) < 1e-12


def test_planted_transposed_gate_is_caught(monkeypatch):
    # A plausible mistake (storing y_1_2 as U instead of U^T) must fail the
    # state-vector check.
    I, J = 3, 3
    circ = circuits.random_circuit(I, J, 8, 1, gates="uniform")
    assert "y_1_2" in circ
    good = np.array(network.GATES["y_1_2"]).reshape(2, 2)
    monkeypatch.setitem(network.GATES, "y_1_2", list(good.T.reshape(-1)))
    psi = statevector.final_state_vector(Circuit(circ), I, J, [], "0" * 9)
    bs = circuits.sample_bitstrings(psi, 4, 0)
    errs = []
    for b in bs:
        r = evaluate_circuit(circ, circuits.snake_ordering(I, J), circuits.grid_text(I, J),
                             "0" * 9, b)
        errs.append(abs(r[0][1] - psi[tuple(int(x) for x in b)]))
    assert max(errs) > 1e-3


def test_validate_ordering_errors():
    with pytest.raises(OracleError, match="contracted multiple times"):
        validate_ordering([("expand", "a", (0, 0)), ("expand", "b", (0, 0))])
    with pytest.raises(OracleError, match="after a cut"):
        validate_ordering([("expand", "a", (0, 0)), ("cut", ((0, 0), (0, 1)), []),
                           ("expand", "a", (0, 1))])
    with pytest.raises(OracleError, match="cut multiple times"):
        validate_ordering([("cut", ((0, 0), (0, 1)), []), ("cut", ((0, 0), (0, 1)), [])])


def test_memory_limit_message():
    circ = circuits.random_circuit(3, 3, 16, 1)
    with pytest.raises(OracleError, match=r"Required space \(.*\) exceeds memory limit \(1 kB\)"):
        evaluate_circuit(circ, circuits.snake_ordering(3, 3), circuits.grid_text(3, 3),
                         "0" * 9, "0" * 9, memory_limit=1024)


# ---------------------------------------------------------------- cost-model network

def test_bf16_rounding_ties_to_even():
    from oracle.mlp import bf16
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 1e-3], dtype=np.float32)
    got = bf16(x)
    assert got[0] == 1.0 and got[1] == 1.0 and got[2] == 1.0 + 4 * 2 ** -8
    assert got[3] == -2.5 and abs(got[4] - 1e-3) <= 2 ** -9 * 1e-3


def _mlp_case(seed, B=8):
    from oracle import mlp
    rng = np.random.default_rng(seed)
    p = mlp.init_params(seed)
    p[mlp.N_IN * mlp.N_HID:mlp.N_IN * mlp.N_HID + mlp.N_HID] = rng.standard_normal(mlp.N_HID) * 0.1
    X = rng.standard_normal((B, mlp.N_IN)).astype(np.float32)
    t = rng.standard_normal(B).astype(np.float32)
    return p, X, t


def test_mlp_oracle_gradient_matches_finite_differences():
    from oracle import mlp
    p, X, t = _mlp_case(1)
    loss, g = mlp.loss_and_grads(p, X, t, rounding=False)
    y, _ = mlp.forward_f64(p.astype(np.float64), X)
    assert abs(loss - float(np.mean((y - t) ** 2))) < 1e-12
    rng = np.random.default_rng(2)
    offsets = np.cumsum([0] + [int(np.prod(s)) for _, s in mlp.SIZES])
    idx = [int(rng.integers(offsets[b], offsets[b + 1])) for b in range(len(mlp.SIZES))
           for _ in range(6)]
    for i in idx:
        fd = mlp.finite_difference(p, X, t, i)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])), (i, fd, g[i])


def test_mlp_finite_differences_catch_a_missing_relu_mask():
    # planted mistake: layer-1 gradient without the ReLU mask
    from oracle import mlp
    p, X, t = _mlp_case(3)
    P = mlp.unpack(p)
    y, (Xb, H1, H2) = mlp.forward_f64(p.astype(np.float64), X)
    dy = 2.0 * (y - t) / len(t)
    dZ2 = (H2 > 0) * dy[:, None] * P["w3"][None, :]
    dW1_wrong = Xb.T @ (dZ2 @ P["W2"].astype(np.float64).T)
    worst = 0.0
    for i in range(0, mlp.N_IN * mlp.N_HID, 997):
        fd = mlp.finite_difference(p, X, t, i)
        worst = max(worst, abs(fd - dW1_wrong.reshape(-1)[i]))
    assert worst > 1e-4


# ---------------------------------------------------------------- FP8 (E4M3)

def test_e4m3_table_known_values():
    from oracle import fp8
    v = fp8.e4m3_values()
    assert v[0x7e] == 448.0 and v[0x01] == 2.0 ** -9 and v[0x08] == 2.0 ** -6
    assert v[0x38] == 1.0 and v[0xb8] == -1.0 and np.isnan(v[0x7f]) and np.isnan(v[0xff])
    pos = np.unique(v[np.isfinite(v) & (v > 0)])
    assert len(pos) == 126  # 7 subnormals + 15 binades x 8 - the NaN code
    # every finite value round-trips through the rounding
    fin = v[np.isfinite(v)]
    assert np.array_equal(fp8.e4m3_round(fin), fin)


def test_e4m3_round_matches_brute_force_nearest_even():
    """Brute force over the code table: nearest value, ties to the even code."""
    from oracle import fp8
    v = fp8.e4m3_values()
    codes = [c for c in range(256) if np.isfinite(v[c])]
    vals = np.array([v[c] for c in codes])
    pos = np.unique(vals[vals >= 0])
    mids = (pos[1:] + pos[:-1]) / 2
    rng = np.random.default_rng(5)
    x = np.concatenate([mids, -mids, rng.uniform(-448, 448, 2000),
                        rng.standard_normal(2000) * 2.0 ** rng.integers(-12, 8, 2000)])
    x = x[np.abs(x) <= 448]
    got = fp8.e4m3_round(x)
    for xi, gi in zip(x, got):
        d = np.abs(vals - xi)
        best = np.flatnonzero(d == d.min())
        if len(best) > 1:
            best = [b for b in best if codes[b] % 2 == 0]
        assert gi == vals[best[0]], (xi, gi)


def test_fp8_cgemm_oracle_exact_on_small_integers():
    """|x| <= 8 integers have at most 3 significant bits: exact at any tile
    scale, so the FP8 form reproduces the complex product (pins the real
    embedding, the 2x2 blocks of B and the tiling)."""
    from oracle import fp8
    rng = np.random.default_rng(6)
    a = rng.integers(-8, 9, (200, 150)) + 1j * rng.integers(-8, 9, (200, 150))
    b = rng.integers(-8, 9, (150, 90)) + 1j * rng.integers(-8, 9, (150, 90))
    a[:128, :64] *= 2.0 ** 9  # tiles at different scales (sums stay exact in float64)
    b[64:, :] *= 2.0 ** -7
    assert np.array_equal(fp8.cgemm_fp8(a, b), a @ b)


def test_fp8_quantize_error_bound_and_scale_range():
    from oracle import fp8
    rng = np.random.default_rng(7)
    x = rng.standard_normal((256, 384)) * 2.0 ** rng.integers(-20, 20, (256, 1))
    q, s = fp8.quantize(x)
    for ti in range(2):
        for tj in range(3):
            blk = x[128 * ti:128 * ti + 128, 128 * tj:128 * tj + 128]
            scaled = np.abs(blk).max() * 2.0 ** -s[ti, tj]
            assert 128 <= scaled < 256
            qb = q[128 * ti:128 * ti + 128, 128 * tj:128 * tj + 128]
            # relative half-quantum in the normal range, absolute 2^-10 2^s below it
            bound = np.maximum(np.abs(blk) * 2.0 ** -4, 2.0 ** (s[ti, tj] - 10))
            assert np.all(np.abs(qb - blk) <= bound)


def test_fp8_planted_sign_error_is_caught(monkeypatch):
    """A Bhat with the -bi block on the wrong row breaks the exact check."""
    from oracle import fp8
    good = fp8.real_bhat

    def bad(b):
        out = good(b)
        out[0::2, 1::2], out[1::2, 0::2] = out[1::2, 0::2].copy(), out[0::2, 1::2].copy()
        return out

    monkeypatch.setattr(fp8, "real_bhat", bad)
    rng = np.random.default_rng(8)
    a = rng.integers(-8, 9, (64, 40)) + 1j * rng.integers(-8, 9, (64, 40))
    b = rng.integers(-8, 9, (40, 30)) + 1j * rng.integers(-8, 9, (40, 30))
    assert not np.array_equal(fp8.cgemm_fp8(a, b), a @ b)
