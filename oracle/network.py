"""Circuit -> 3D tensor network -> 2D grid (src/read_circuit.cpp), restated.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).
"""
import cmath
import functools
import math

import numpy as np

from .tensor_ops import OTensor, OracleError, multiply, bundle_between

_R = 1.0 / math.sqrt(2.0)

# Gate table, same numbers as src/read_circuit.cpp:27-79 (a table that must be
# alike for bit-level parity).  One-qubit gates are stored (input, output), i.e.
# the data is U^T in row-major; two-qubit gates are split into two rank-3
# (input, virtual, output) tensors whose contraction over `virtual` is the gate.
GATES = {
    "delta_0": [1.0, 0.0],
    "delta_1": [0.0, 1.0],
    "h": [_R, _R, _R, -_R],
    "hz_1_2": [0.5 + 0.5j, _R, -_R * 1j, 0.5 + 0.5j],
    "t": [1.0, 0.0, 0.0, complex(_R, _R)],
    "x_1_2": [0.5 + 0.5j, 0.5 - 0.5j, 0.5 - 0.5j, 0.5 + 0.5j],
    "y_1_2": [0.5 + 0.5j, 0.5 + 0.5j, -0.5 - 0.5j, 0.5 + 0.5j],
    "cz_q1": [-0.3446133714352872, 0., 1.1381806476131544, 0., 0.,
              -1.1381806476131544, 0., -0.3446133714352872],
    "cz_q2": [-1.0484937059720079, 0., 0.5611368023131075, 0., 0.,
              0.5611368023131075, 0., 1.0484937059720079],
    "cx_q1": [0.8408964152537143, 0., 0.8408964152537143, 0., 0.,
              -0.8408964152537143, 0., 0.8408964152537143],
    "cx_q2": [0.5946035575013604, -0.5946035575013604, 0.5946035575013604,
              0.5946035575013604, -0.5946035575013604, 0.5946035575013604,
              0.5946035575013604, 0.5946035575013604],
}


def gate_array(name, params):
    """src/read_circuit.cpp:82-95; rz angle in units of pi."""
    if name == "rz":
        ph = cmath.exp(1j * math.pi * params[0] / 2)
        return [ph.conjugate(), 0.0, 0.0, ph]
    if name not in GATES:
        raise OracleError(f"Invalid gate name provided: {name}")
    return GATES[name]


def fsim_arrays(theta, phi):
    """src/read_circuit.cpp:105-139 (theta, phi in radians)."""
    c1 = 0.5 * (cmath.exp(-1j * phi / 2) + math.cos(theta))
    c2 = 0.5 * (cmath.exp(-1j * phi / 2) - math.cos(theta))
    c3 = -0.5j * math.sin(theta)
    c4 = c3
    s1, s2, s3, s4 = (cmath.sqrt(c) for c in (c1, c2, c3, c4))
    d = cmath.exp(1j * phi / 4)
    arr = [d * s1, 0, d * s2, 0, 0, s3, 0, 1j * s4,
           0, d.conjugate() * s1, 0, -d.conjugate() * s2, s3, 0, -1j * s4, 0]
    return arr, list(arr)


def gate_arrays(name, params):
    """src/read_circuit.cpp:146-165."""
    if name == "cz":
        return GATES["cz_q1"], GATES["cz_q2"], [2, 2, 2]
    if name == "cx":
        return GATES["cx_q1"], GATES["cx_q2"], [2, 2, 2]
    if name == "fsim":
        a, b = fsim_arrays(math.pi * params[0], math.pi * params[1])
        return a, b, [2, 4, 2]
    raise OracleError(f"Invalid gate name provided: {name}")


def _pos_name(i, j):
    return f"({i},{j})"


def index_name(p1, p2=None):
    """src/contraction_utils.cpp:477-500."""
    if p2 is None or len(p2) == 0:
        return f"({p1[0]},{p1[1]}),(o)"
    if len(p1) == 2:
        return f"({p1[0]},{p1[1]}),({p2[0]},{p2[1]})"
    return f"({p1[0]},{p1[1]},{p1[2]}),({p2[0]},{p2[1]},{p2[2]})"


def circuit_to_network(circuit, I, J, initial, final, final_region, off, dtype):
    """circuit_data_to_tensor_network (src/read_circuit.cpp:296-544)."""
    offs = set(map(tuple, off))
    final_region = set(map(tuple, final_region))
    n_active = I * J - len(offs)
    if circuit.num_active_qubits != n_active:
        raise OracleError("The number of active qubits read from the file does "
                          "not match the number of active qubits read from the grid.")
    if len(initial) != n_active or len(final) != len(initial):
        raise OracleError("Size of initial_conf/final_conf mismatch.")
    grid = [[[] for _ in range(J)] for _ in range(I)]
    counters = [[0] * J for _ in range(I)]
    links = {}
    idx = 0
    for q in range(I * J):
        i, j = divmod(q, J)
        if (i, j) in offs:
            continue
        g = "delta_0" if initial[idx] == "0" else "delta_1"
        grid[i][j].append(OTensor([f"({i},{j}),(0)"], [2], GATES[g], dtype))
        idx += 1
    for gate in circuit.gates:
        if len(gate.qubits) == 1:
            i, j = divmod(gate.qubits[0], J)
            if (i, j) in offs:
                raise OracleError("Qubit must correspond to an active qubit.")
            a = f"({i},{j}),({counters[i][j]})"
            counters[i][j] += 1
            b = f"({i},{j}),({counters[i][j]})"
            grid[i][j].append(OTensor([a, b], [2, 2], gate_array(gate.name, gate.params), dtype))
        elif len(gate.qubits) == 2:
            p1 = divmod(gate.qubits[0], J)
            p2 = divmod(gate.qubits[1], J)
            if abs(p1[0] - p2[0]) + abs(p1[1] - p2[1]) != 1:
                raise OracleError("Qubits are not nearest neighbors.")
            g1, g2, dims = gate_arrays(gate.name, gate.params)
            link = index_name(p1, p2)
            links[link] = links.get(link, 0) + 1
            c = links[link]
            virt = (f"({min(p1[0], p2[0])},{min(p1[1], p2[1])},{c}),"
                    f"({max(p1[0], p2[0])},{max(p1[1], p2[1])},{c})")
            in1 = f"({p1[0]},{p1[1]}),({counters[p1[0]][p1[1]]})"
            in2 = f"({p2[0]},{p2[1]}),({counters[p2[0]][p2[1]]})"
            counters[p1[0]][p1[1]] += 1
            counters[p2[0]][p2[1]] += 1
            out1 = f"({p1[0]},{p1[1]}),({counters[p1[0]][p1[1]]})"
            out2 = f"({p2[0]},{p2[1]}),({counters[p2[0]][p2[1]]})"
            grid[p1[0]][p1[1]].append(OTensor([in1, virt, out1], dims, g1, dtype))
            grid[p2[0]][p2[1]].append(OTensor([in2, virt, out2], dims, g2, dtype))
        else:
            raise OracleError("k-qubit gates with k > 2 not yet implemented.")
    idx = 0
    for q in range(I * J):
        i, j = divmod(q, J)
        if (i, j) in offs:
            continue
        last = f"({i},{j}),({counters[i][j]})"
        if (i, j) in final_region:
            t = grid[i][j][-1]
            t.indices[t.indices.index(last)] = f"({i},{j}),(o)"
        else:
            g = "delta_0" if final[idx] == "0" else "delta_1"
            grid[i][j].append(OTensor([last], [2], GATES[g], dtype))
        idx += 1
    return grid


def order_cmp(ordering, local):
    """Comparator of src/read_circuit.cpp:191-291 (order_func), as cmp()."""
    local = tuple(local)

    def less(lhs, rhs):
        lp = (local, lhs) if (local[0] < lhs[0] or local[1] < lhs[1]) else (lhs, local)
        rp = (local, rhs) if (local[0] < rhs[0] or local[1] < rhs[1]) else (rhs, local)
        for op in ordering:
            if op[0] != "cut":
                continue
            if tuple(op[1]) == lp:
                return True
            if tuple(op[1]) == rp:
                return False

        def find(pos):
            n = 0
            for op in ordering:
                if op[0] != "expand":
                    continue
                if op[2] == pos:
                    return op[1], n
                n += 1
            return None, None
        lpatch, lpos = find(lhs)
        if lpos is None:
            raise OracleError("Left hand side of pair not found.")
        rpatch, rpos = find(rhs)
        if rpos is None:
            raise OracleError("Right hand side of pair not found.")
        if lpatch == rpatch:
            return lpos < rpos
        local_patch = None
        for op in ordering:
            if op[0] == "expand" and op[2] == local:
                local_patch = op[1]
                break
        if lpatch == local_patch:
            return True
        if rpatch == local_patch:
            return False
        for op in ordering:
            if op[0] != "merge":
                continue
            src, tgt = op[1], op[2]
            if local_patch == src:
                if lpatch == tgt:
                    return True
                if rpatch == tgt:
                    return False
                local_patch = tgt
            elif local_patch == tgt:
                if lpatch == src:
                    return True
                if rpatch == src:
                    return False
            elif lpatch == src:
                lpatch = tgt
            elif rpatch == src:
                rpatch = tgt
        raise OracleError("Failed to compare pairs.")

    def cmp(a, b):
        if less(a, b):
            return -1
        if less(b, a):
            return 1
        return 0
    return cmp


def _insertion_sort(items, cmp):
    """std::sort on <= 4 items is an insertion sort; keep its exact order."""
    out = list(items)
    for k in range(1, len(out)):
        x = out[k]
        m = k
        while m > 0 and cmp(x, out[m - 1]) < 0:
            out[m] = out[m - 1]
            m -= 1
        out[m] = x
    return out


def flatten(grid3d, I, J, final_region, off, ordering):
    """flatten_grid_of_tensors (src/read_circuit.cpp:548-689)."""
    offs = set(map(tuple, off))
    final_region = set(map(tuple, final_region))
    grid2d = [[None] * J for _ in range(I)]
    for i in range(I):
        for j in range(J):
            if (i, j) in offs:
                continue
            if not grid3d[i][j]:
                raise OracleError("Time-direction cannot be empty")
            A = grid3d[i][j][0].copy()
            for B in grid3d[i][j][1:]:
                A = multiply(A, B)
            grid2d[i][j] = A
    for i in range(I):
        for j in range(J):
            if (i, j) in offs:
                continue
            pairs = []
            if i < I - 1 and (i + 1, j) not in offs:
                pairs.append((i + 1, j))
            if j < J - 1 and (i, j + 1) not in offs:
                pairs.append((i, j + 1))
            for p in pairs:
                bundle_between(grid2d[i][j], grid2d[p[0]][p[1]], index_name((i, j), p))
    for i in range(I):
        for j in range(J):
            if (i, j) in offs:
                continue
            pairs = []
            if i > 0 and (i - 1, j) not in offs:
                pairs.append((i - 1, j))
            if j > 0 and (i, j - 1) not in offs:
                pairs.append((i, j - 1))
            if i < I - 1 and (i + 1, j) not in offs:
                pairs.append((i + 1, j))
            if j < J - 1 and (i, j + 1) not in offs:
                pairs.append((i, j + 1))
            order = []
            if (i, j) in final_region:
                order.append(index_name((i, j)))
            pairs = _insertion_sort(pairs, order_cmp(ordering, (i, j)))
            for p in pairs:
                if p[0] < i or p[1] < j:
                    order.append(index_name(p, (i, j)))
                else:
                    order.append(index_name((i, j), p))
            grid2d[i][j].reorder(order)
    return grid2d
