"""Host-side logic of the C++ drop-in that needs no GPU: parsers, ordering
validation, output-state enumeration and their error texts (restating the
reference's tests/src/{circuit,grid,ordering,contraction_utils,
evaluate_circuit}_test.cpp cases)."""
import json
import os

import numpy as np
import pytest

from qflex_amd import qflex

KATS = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_kats.json")))


def ops_from_json(raw):
    out = []
    for op in raw:
        if op[0] == "cut":
            out.append(("cut", op[1], op[2] if len(op) > 2 else []))
        else:
            out.append(tuple(op))
    return out


@pytest.mark.parametrize("name", [n for n in KATS["get_output_states"] if not n.startswith("_")])
def test_get_output_states(name):
    c = KATS["get_output_states"][name]
    base, fq, states = qflex.get_output_states(c["I"], c["J"], c["off"], "", c["final"],
                                               ops_from_json(c["ops"]))
    assert fq == c["final_qubits"]
    assert states == c["states"]


def test_get_output_states_ignores_non_terminal_ops():
    # evaluate_circuit_test.cpp:24-49
    for ops in ([("expand", "a", [0, 0]), ("expand", "a", [0, 1])],
                [("merge", "a", "b"), ("merge", "b", "c")],
                [("cut", [[0, 0], [0, 1]], [1, 2]), ("cut", [[0, 0], [1, 0]], [3])]):
        base, fq, states = qflex.get_output_states(0, 0, [], "", "", ops)
        assert fq == [] and states == [""]


def test_parse_simple_ordering():
    # contraction_utils_test.cpp:376-428
    lines = ["# test comment", "cut (1,2) 1 3", "expand a 1", "expand a 0", "expand a 2",
             "cut () 5", "expand b 5", "expand b 3", "merge a b"]
    ops = qflex.parse_ordering(lines, ["11", "11", "01"])
    assert ops == [("cut", [[0, 1], [1, 1]], [1, 2]), ("expand", "a", [0, 1]),
                   ("expand", "a", [0, 0]), ("expand", "a", [1, 0]), ("cut", [[2, 1]], []),
                   ("expand", "b", [2, 1]), ("expand", "b", [1, 1]), ("merge", "a", "b")]


def test_parse_inverted_cut_and_spaces():
    assert qflex.parse_ordering(["cut (1,2) 1 0"], ["11", "11", "01"]) == \
        [("cut", [[0, 0], [0, 1]], [1, 2])]
    assert qflex.parse_ordering(["cut   (  1  , 2   )    2        3"], ["11", "11", "01"]) == \
        [("cut", [[1, 0], [1, 1]], [1, 2])]


@pytest.mark.parametrize("line,msg", [
    ("bad op 1 2", "Received an invalid operation in config."),
    ("expand a 8", "Index 8 must be within grid boundaries."),
    ("expand a -2", "Index -2 must be within grid boundaries."),
    ("cut () 1 7", "Index 7 must be within grid boundaries."),
    ("cut () -1 4", "Index -1 must be within grid boundaries."),
    ("cut 2 3", "Cut values must be comma-separated ints, e.g. (0,1,3)."),
    ("expand a 4", "Index must specify an active qubit."),
])
def test_parse_ordering_errors(line, msg):
    with pytest.raises(RuntimeError, match=msg.replace("(", r"\(").replace(")", r"\)")
                       .replace(".", r"\.")):
        qflex.parse_ordering([line], ["11", "11", "01"])


def test_validate_ordering_errors():
    cut = ("cut", [[1, 2], [1, 3]], [])
    cases = [
        ([("expand", "a", [1, 2]), ("expand", "b", [1, 2])], "Tensor (1,2) is contracted multiple times."),
        ([cut, cut], "Index (1,2),(1,3) is cut multiple times."),
        ([("merge", "a", "c"), ("merge", "a", "b")], "Patch a is merged multiple times."),
        ([("merge", "a", "b"), ("expand", "a", [1, 2])],
         "Tensor at (1,2) is added to non-empty patch a after a cut."),
        ([("expand", "a", [1, 2]), cut, ("expand", "a", [2, 2])],
         "Tensor at (2,2) is added to non-empty patch a after a cut."),
        ([("expand", "a", [1, 1]), ("expand", "b", [1, 2]), cut, ("merge", "a", "b")],
         "Patch a is merged into non-empty patch b after a cut."),
    ]
    for ops, msg in cases:
        with pytest.raises(RuntimeError) as e:
            qflex.validate_ordering(ops)
        assert msg in str(e.value), (ops, str(e.value))


BAD_CIRCUITS = [
    ("0 h 0\n0 h 1\n9 h 0\n9 h 1",
     "[1: 0 h 0] First line in circuit must be the number of active qubits."),
    ("6\n0 h 0\n0 h\n9 h 0\n9 h 1",
     "[3: 0 h] Gate must be specified as: cycle gate_name[(p1[,p2,...])] q1 [q2, ...]"),
    ("2\n0 h 0\n0 h 1\ncz 0 2\n9 h 1", "[4: cz 0 2] First token must be a valid cycle number."),
    ("2\n0 h 0\n0 h 1\n9 cz 0 2\n7 cz 1 3\n9 h 1", "[5: 7 cz 1 3] Cycle number can only increase."),
    ("4\n0 h 0\n0 h 1\n1 t 0\n1 t 2\n2 fsim(0.1, a) 0 1\n4 h 0\n4 h 1\n",
     "Params must be valid numbers."),
    ("2\n0 h 0\n0 h 1\n1 t 0\n1 t a\n4 h 0\n4 h 1", "[5: 1 t a] Qubit must be a valid number."),
    ("4\n0 h 0\n0 h 1\n4 t 2\n4 t 4 \n4 cz 2 4\n9 cz 0 2\n9 cz 1 3\n9 h 1",
     "[6: 4 cz 2 4] Qubits can only used once per cycle."),
    ("2\n0 h(0 0\n", "Not-matching parentheses."),
]


@pytest.mark.parametrize("text,msg", BAD_CIRCUITS)
def test_circuit_errors(text, msg):
    # tests/src/circuit_test.cpp:36-160
    with pytest.raises(RuntimeError) as e:
        qflex.load_circuit(text.split("\n"))
    assert msg in str(e.value)


def test_circuit_parse_and_depth():
    n, depth, gates = qflex.load_circuit(
        ["3 # qubits", "0 h 0", "0\th 1", "1 fsim(0.25, 0.1) 0 1", "2 cz 1 2 # c", "3 cz 0 1"])
    assert n == 3 and depth == 3
    assert gates[1] == (0, "h", [1], [])
    assert gates[2] == (1, "fsim", [0, 1], [0.25, 0.1])


def test_grid_parse_and_error():
    I, J, off = qflex.load_grid(["1 1 0", "# c", "1 0 1   ", ""])
    assert (I, J, off) == (2, 3, [[0, 2], [1, 1]])
    with pytest.raises(RuntimeError, match="Grid size at line 2 is inconsistent with a width of 4 "
                                           "instead of 3"):
        qflex.load_grid(["1 1 1", "1 1 1 1", "1 1 1"])


def test_simulate_swallows_errors_and_returns_empty(capfd):
    # src/pybind_main.cpp:127-137: errors are printed, [] is returned
    out = qflex.simulate({"circuit": ["2", "0 h 0", "0 h 1"], "grid": ["11"],
                          "ordering": ["expand a 0", "expand a 1"],
                          "initial_states": ["00", "11"]})
    assert out == []
    assert "Not yet supported" in capfd.readouterr().err


def test_simulate_rejects_bad_memory_limit(capfd):
    assert qflex.simulate({"memory_limit": -5}) == []
    assert "'memory_limit' must be positive." in capfd.readouterr().err


# ---------------------------------------------------------------- text formats in depth

def reference_line_rules(line, glue_parens):
    """The line normalisation that src/circuit.cpp:47-73 (circuit text, with
    glue_parens) and src/ordering.cpp:12-38 (ordering text) state as a
    sequence of regular-expression rewrites, applied here with Python's re
    in ASCII mode (the character classes of std::regex's ECMAScript
    grammar).  The C++ parser implements the same contract with its own
    scanner; this is the independent statement it is checked against."""
    import re
    A = re.ASCII
    line = re.sub(r"#.*", "", line, flags=A)
    line = re.sub(r"[^)(\s\ta-zA-Z0-9_.,-]", "", line, flags=A)
    line = re.sub(r"[\t]", " ", line, flags=A)
    line = re.sub(r"[\s]{2,}", " ", line, flags=A)
    line = re.sub(r"\s+$", "", line, flags=A)
    if glue_parens:
        line = re.sub(r"[\s]+[(]", "(", line, flags=A)
    line = re.sub(r"\s+(?=[^()]*\))", "", line, flags=A)
    return line


@pytest.mark.parametrize("glue", [True, False])
def test_canonical_lines_match_the_reference_rewrite_rules(glue):
    rng = np.random.default_rng(7 if glue else 8)
    alphabet = list(" \t\r()#,.-_aZx09$%;") + ["  ", "\t\t", " (", ") ", "fsim", "cz"]
    lines = ["  2\tfsim( 0.25 , 0.1 )  0 1  # c", "0 h 0", "\t\t", "cut   (  1  , 2   )  2  3",
             "a (b (c) d) e", "x ) y ( z", " (", "(  )", "1 rz (0.5) 3\r"]
    for _ in range(3000):
        n = int(rng.integers(0, 14))
        lines.append("".join(alphabet[int(i)] for i in rng.integers(0, len(alphabet), n)))
    got = qflex.canonical_lines(lines, glue)
    want = [reference_line_rules(l, glue) for l in lines]
    bad = [(l, g, w) for l, g, w in zip(lines, got, want) if g != w]
    assert not bad, bad[:5]


MORE_BAD_CIRCUITS = [
    ("2\n1.5 h 0\n", "[2: 1.5 h 0] First token must be a valid cycle number."),
    ("2\n1 h 0.5\n", "[2: 1 h 0.5] Qubit must be a valid number."),
    ("2\n1 rz(.5) 0\n", "Params must be valid numbers."),
    ("3\n1 ccz 0 1 2\n", "Depth calculation does not handle k-qubit gates with k > 2."),
    ("2\n1 h((0) 0\n", "[2: 1 h((0) 0] Not-matching parentheses."),
    ("2\n1 rz(0.1)(0.2) 0\n", "Not-matching parentheses."),
    ("x\n", "[1: x] First line in circuit must be the number of active qubits."),
    ("0\n", "[1: 0] First line in circuit must be the number of active qubits."),
    ("2 3\n", "[1: 2 3] First line in circuit must be the number of active qubits."),
]


@pytest.mark.parametrize("text,msg", MORE_BAD_CIRCUITS)
def test_more_circuit_errors(text, msg):
    with pytest.raises(RuntimeError) as e:
        qflex.load_circuit(text.split("\n"))
    assert msg in str(e.value)


def test_circuit_grammar_details():
    # spaces and tabs around parameters, a comment-only line numbered but
    # skipped, "3.0" as a valid integer, fsim counted twice in the depth
    n, depth, gates = qflex.load_circuit(
        ["4", "# layer 0", "0 h 0", "1  rz ( 0.5 )\t2", "1 fsim( 0.25 ,0.1 ) 0 1",
         "3.0 cz 2 3", "4 t 1 # trailing"])
    assert n == 4 and depth == 2
    assert gates == [(0, "h", [0], []), (1, "rz", [2], [0.5]), (1, "fsim", [0, 1], [0.25, 0.1]),
                     (3, "cz", [2, 3], []), (4, "t", [1], [])]
    # the qubit count must be on the first physical line: after a leading
    # comment line the count is read as a gate line
    with pytest.raises(RuntimeError, match=r"\[2: 4\] Gate must be specified as"):
        qflex.load_circuit(["# comment", "4", "0 h 0"])


def test_missing_files_raise(tmp_path):
    with pytest.raises(RuntimeError, match="Cannot open circuit file"):
        qflex.evaluate({"circuit_filename": str(tmp_path / "none"), "grid": ["1"],
                        "ordering": ["expand a 0"]})
    with pytest.raises(RuntimeError, match="Cannot open grid file: "):
        qflex.evaluate({"grid_filename": str(tmp_path / "none")})


def test_grid_details():
    # rows count non-comment lines (even rows that end up empty) for the
    # error text; a line without 0/1 characters is no row
    I, J, off = qflex.load_grid(["# c", "1 0", "abc", "0 1 # 2"])
    assert (I, J, off) == (2, 2, [[0, 1], [1, 0]])
    with pytest.raises(RuntimeError, match="Grid size at line 3 is inconsistent with a width of 1 "
                                           "instead of 2"):
        qflex.load_grid(["11", "x", "1"])


@pytest.mark.parametrize("line,msg", [
    ("cut (a) 1 2", "Cut values must be comma-separated ints, e.g. (0,1,3)."),
    ("cut 0,1 1 2", "Cut values must be comma-separated ints, e.g. (0,1,3)."),
    ("cut () 4 1", "Index 1 must specify an active qubit."),
    ("cut () 1 4", "Index 2 must specify an active qubit."),
    ("cut () 4", "Index 1 must specify an active qubit."),
    ("expand a", "Index -1 must be within grid boundaries."),
    ("cut ()", "Index -1 must be within grid boundaries."),
    ("cut () 1 6", "Index 6 must be within grid boundaries."),
])
def test_more_ordering_errors(line, msg):
    with pytest.raises(RuntimeError) as e:
        qflex.parse_ordering([line], ["11", "11", "01"])
    assert f"Parsing failed on line: '{line}' with error: {msg}" in str(e.value)


def test_ordering_grammar_details():
    grid = ["11", "11", "01"]
    # spaces inside the values are dropped by the line rules: "(1 2)" is (12)
    assert qflex.parse_ordering(["cut (1 2) 0 1"], grid) == [("cut", [[0, 0], [0, 1]], [12])]
    # empty pieces read as 0; a terminal cut has one qubit
    assert qflex.parse_ordering(["cut (1,,2) 5"], grid) == [("cut", [[2, 1]], [1, 0, 2])]
    # a merge with one name merges into the unnamed patch
    assert qflex.parse_ordering(["expand a 0", "merge a"], grid) == \
        [("expand", "a", [0, 0]), ("merge", "a", "")]
    # comment and blank lines are dropped by the loader
    assert qflex.parse_ordering(["# x", "", "expand b 3 # y"], grid) == [("expand", "b", [1, 1])]


def test_validate_ordering_collects_every_problem():
    ops = [("expand", "a", [0, 0]), ("merge", "a", "b"), ("expand", "a", [0, 1]),
           ("expand", "c", [0, 0]), ("merge", "a", "b")]
    with pytest.raises(RuntimeError) as e:
        qflex.validate_ordering(ops)
    msg = str(e.value)
    assert msg.count("\n") == 3
    assert "Tensor at (0,1) is added to non-empty patch a after a cut." in msg
    assert "Tensor (0,0) is contracted multiple times." in msg
    assert "Patch a is merged multiple times." in msg
    # a well-formed ordering passes
    qflex.validate_ordering([("cut", [[0, 0], [0, 1]], []), ("expand", "a", [0, 0]),
                             ("expand", "b", [0, 1]), ("merge", "a", "b")])


def test_simulate_input_type_errors(capfd):
    for opts, msg in [({"grid": "11"}, "'grid' must be a list of strings."),
                      ({"grid": ["1", 1]}, "grid must be a list of strings."),
                      ({"circuit_filename": 3}, "'circuit_filename' must be a string."),
                      ({"initial_states": 5}, "initial_states must be a list of strings."),
                      ({"final_states": ["0", 1]}, "states must be strings."),
                      ({"final_state": 0}, "states must be strings."),
                      ({"memory_limit": 1.5}, "'memory_limit' must be valid.")]:
        assert qflex.simulate(opts) == []
        assert msg in capfd.readouterr().err, opts


# ---------------------------------------------------------------- cost-model candidates

@pytest.mark.parametrize("I,J", [(4, 4), (5, 6), (6, 6), (7, 7)])
def test_costmodel_candidates_are_valid_orderings(I, J):
    # every candidate (and the plain fallback) parses and passes the
    # engine's ValidateOrdering; every site is expanded exactly once
    from qflex_amd import circuits, costmodel
    rng = np.random.default_rng(I * 10 + J)
    grid = circuits.grid_text(I, J).splitlines()
    cands = costmodel.candidate_set(I, J, rng, 40)
    assert cands[0] == costmodel.plain_ordering(I, J) and len(set(cands)) == len(cands)
    for order in cands:
        lines = order.splitlines()
        ops = qflex.parse_ordering(lines, grid)
        qflex.validate_ordering(ops)
        expanded = sorted(int(l.split()[2]) for l in lines if l.startswith("expand"))
        assert expanded == list(range(I * J))
        cuts = [l for l in lines if l.startswith("cut")]
        first_expand = next(n for n, l in enumerate(lines) if l.startswith("expand"))
        assert all(lines.index(c) < first_expand for c in cuts)


def test_costmodel_features_from_a_plan():
    from qflex_amd import costmodel
    plan = {"steps": [{"m": 256, "n": 16, "k": 16, "multiplicity": 1, "out_rank": 3,
                       "out_size": 4096},
                      {"m": 4096, "n": 256, "k": 256, "multiplicity": 16, "out_rank": 5,
                       "out_size": 1 << 20}],
            "patch_max_size": {"a": 1 << 20}, "max_grid": 65536}
    f = costmodel.features(plan, 6, 6, 24)
    assert f.shape == (64,) and f.dtype == np.float32 and np.all(np.isfinite(f))
    total = 8 * (256 * 16 * 16 + 4096 * 256 * 256 * 16)
    assert abs(f[0] - np.log2(total)) < 1e-5
    assert f[2] == 20 and f[3] == 5 and f[6] == 4
    assert abs(f[8:16].sum() - 1) < 1e-6 and abs(f[32:40].sum() - 1) < 1e-6


def test_several_final_states_of_unequal_length_are_refused():
    """simulate's final_states list (one network build for all states): a
    state of the wrong length is refused before any device work."""
    opts = {"circuit": ["2", "0 h 0", "0 h 1", "1 cz 0 1"], "grid": ["11"],
            "ordering": ["expand a 0", "expand a 1"], "initial_state": "00",
            "final_states": ["01", "011"]}
    with pytest.raises(RuntimeError, match="Final state '011' has 3 qubits instead of 2"):
        qflex.evaluate(opts)
