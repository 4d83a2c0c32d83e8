"""Circuit / grid / ordering text formats (happy path), restated.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

References: src/circuit.cpp:30-206, src/grid.cpp:11-45,
src/ordering.cpp:10-45, src/contraction_utils.cpp:358-475.
"""
import re

from .tensor_ops import OracleError

_SPECIAL = re.compile(r"[^)(\s\ta-zA-Z0-9_.,-]")


def _strip(line, space_before_paren):
    line = re.sub(r"#.*", "", line)
    line = _SPECIAL.sub("", line)
    line = line.replace("\t", " ")
    line = re.sub(r"\s{2,}", " ", line)
    line = re.sub(r"\s+$", "", line)
    if space_before_paren:
        line = re.sub(r"\s+\(", "(", line)
    line = re.sub(r"\s+(?=[^()]*\))", "", line)
    return line


class Gate:
    def __init__(self, cycle, name, qubits, params, raw):
        self.cycle, self.name, self.qubits, self.params, self.raw = \
            cycle, name, qubits, params, raw


class Circuit:
    """QflexCircuit::load (src/circuit.cpp:30-206), valid inputs."""

    def __init__(self, text):
        self.num_active_qubits = 0
        self.gates = []
        for counter, raw in enumerate(text.splitlines()):
            line = _strip(raw, True)
            if not line:
                continue
            toks = line.split()
            if counter == 0:
                self.num_active_qubits = int(toks[0])
                continue
            name = toks[1]
            params = []
            if "(" in name:
                beg = name.index("(")
                params = [float(p) for p in name[beg + 1:name.index(")")].split(",") if p]
                name = name[:beg]
            self.gates.append(Gate(int(toks[0]), name, [int(q) for q in toks[2:]],
                                   params, line))


class Grid:
    """QflexGrid::load (src/grid.cpp:11-45)."""

    def __init__(self, text):
        self.I = 0
        self.J = 0
        self.off = []
        for line in text.splitlines():
            if not line or line[0] == "#":
                continue
            bits = "".join(c for c in line if c in "01")
            if not bits:
                continue
            if self.J and self.J != len(bits):
                raise OracleError("Grid size is inconsistent.")
            self.J = len(bits)
            for q, c in enumerate(bits):
                if c == "0":
                    self.off.append((self.I, q))
            self.I += 1


def ordering_lines(text):
    """QflexOrdering::load (src/ordering.cpp:10-45)."""
    out = []
    for raw in text.splitlines():
        line = _strip(raw, False)
        if line:
            out.append(line)
    return out


# Operation tuples used by the oracle's interpreter:
#   ("expand", patch_id, (i, j))
#   ("cut", ((i1, j1), (i2, j2)) or ((i, j),), values_list)
#   ("merge", source_id, target_id)
def parse_ordering(lines, I, J, off):
    """ordering_data_to_contraction_ordering (src/contraction_utils.cpp:358-475)."""
    ops = []
    offset = set(map(tuple, off))

    def coords(q):
        q = int(q)
        if q < 0 or q >= I * J:
            raise OracleError(f"Index {q} must be within grid boundaries.")
        return (q // J, q % J)

    for line in lines:
        if not line or line[0] == "#":
            continue
        toks = line.split()
        op = toks[0]
        if op == "expand":
            pos = coords(toks[2])
            if pos in offset:
                raise OracleError("Index must specify an active qubit.")
            ops.append(("expand", toks[1], pos))
        elif op == "cut":
            vs = toks[1]
            if not re.fullmatch(r"\([0-9,]*\)", vs):
                raise OracleError("Cut values must be comma-separated ints, e.g. (0,1,3).")
            inner = vs[1:-1]
            values = [int(v) if v else 0 for v in inner.split(",")] if inner else []
            q1 = int(toks[2])
            p1 = coords(q1)
            if len(toks) == 3:
                ops.append(("cut", (p1,), values))
            else:
                q2 = int(toks[3])
                p2 = coords(q2)
                ops.append(("cut", (p1, p2) if q1 < q2 else (p2, p1), values))
        elif op == "merge":
            ops.append(("merge", toks[1], toks[2]))
        else:
            raise OracleError("Received an invalid operation in config.")
    return ops
