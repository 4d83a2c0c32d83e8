"""Grid contraction interpreter and EvaluateCircuit, restated.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

References: src/contraction_utils.cpp:16-354 (ContractionData),
:534-610 (ValidateOrdering), src/evaluate_circuit.cpp:22-253.
"""
import numpy as np

from .formats import Circuit, Grid, ordering_lines, parse_ordering
from .network import circuit_to_network, flatten, index_name
from .tensor_ops import OracleError, multiply, _sub, _inter, _union


def validate_ordering(ops):
    """ValidateOrdering (src/contraction_utils.cpp:534-610)."""
    state = {}
    cuts, used = set(), set()
    msg = ""

    def st(p):
        return state.setdefault(p, {"active": False, "used": False, "merged": False})
    for op in ops:
        if op[0] == "expand":
            s = st(op[1])
            if s["used"] or s["merged"]:
                msg += (f"\nTensor at ({op[2][0]},{op[2][1]}) is added to non-empty "
                        f"patch {op[1]} after a cut.")
            name = f"({op[2][0]},{op[2][1]})"
            if name in used:
                msg += f"\nTensor {name} is contracted multiple times."
            used.add(name)
            s["active"] = True
        elif op[0] == "cut":
            for s in state.values():
                if s["active"]:
                    s["used"] = True
            name = index_name(*op[1]) if len(op[1]) == 2 else index_name(op[1][0])
            if name in cuts:
                msg += f"\nIndex {name} is cut multiple times."
            cuts.add(name)
        else:
            if st(op[1])["merged"]:
                msg += f"\nPatch {op[1]} is merged multiple times."
            if st(op[2])["used"]:
                msg += (f"\nPatch {op[1]} is merged into non-empty patch {op[2]} "
                        f"after a cut.")
            st(op[1])["merged"] = True
            st(op[2])["active"] = True
    if msg:
        raise OracleError(msg)


def cut_name(cut_tensors):
    if len(cut_tensors) == 2:
        return index_name(cut_tensors[0], cut_tensors[1])
    return index_name(cut_tensors[0])


def readable_memory_string(x):
    """utils::readable_memory_string (src/utils.cpp:60-68)."""
    suffix = [" B", " kB", " MB", " GB"]
    s = 0
    while x >= 1024:
        s += 1
        x /= 1024
    return f"{int(x * 100) / 100.0:g}{suffix[s]}"


def plan(ops, grid, memory_limit=1 << 30):
    """Symbolic pass + space plan of ContractionData::Initialize
    (src/contraction_utils.cpp:16-221).  Returns (allocated elements,
    patch_max_size) and raises the reference's memory-limit error."""
    index_dim, gidx = {}, {}
    for i, row in enumerate(grid):
        for j, t in enumerate(row):
            if t is None:
                continue
            gidx[(i, j)] = list(t.indices)
            index_dim.update(t.index_to_dim())
    pidx, pmax, cut_copy = {}, {}, {}
    for op in ops:
        if op[0] == "expand":
            a = pidx.get(op[1], [])
            b = gidx[op[2]]
            pidx[op[1]] = _sub(_union(a, b), _inter(a, b))
        elif op[0] == "cut":
            name = cut_name(op[1])
            for side, pos in enumerate(op[1]):
                gidx[pos] = _sub(gidx[pos], [name])
                cut_copy[f"cut-{name}:size-{side}"] = grid[pos[0]][pos[1]].size
            continue
        else:
            a = pidx.get(op[2], [])
            b = pidx.get(op[1], [])
            pidx[op[2]] = _sub(_union(a, b), _inter(a, b))
        pid = op[1] if op[0] == "expand" else op[2]
        size = int(np.prod([index_dim[x] for x in pidx[pid]])) if pidx[pid] else 1
        pmax[pid] = max(pmax.get(pid, 0), size)
    max_size = max(pmax.values()) if pmax else 0
    total = max_size + sum(set(pmax.values())) + sum(pmax.values()) + sum(cut_copy.values())
    alloc = total * 8
    if alloc > memory_limit:
        raise OracleError(
            f"Required space ({readable_memory_string(alloc)}) exceeds memory "
            f"limit ({readable_memory_string(memory_limit)}). Cancelling simulation.")
    return total, pmax


def contract_grid(ops, grid, amplitudes, phase_seconds=None):
    """ContractionData::ContractGrid (src/contraction_utils.cpp:223-354):
    cut recursion, terminal-cut output indexing, FP64 accumulation.

    phase_seconds (optional dict): accumulates the time spent in expand /
    merge steps by cut depth (number of cuts enclosing the step)."""
    import time
    patches = {}

    def run(rest, out_idx, active, depth=0):
        output = None
        k = 0
        while k < len(rest):
            op = rest[k]
            k += 1
            t_op = time.perf_counter()
            if op[0] == "expand":
                nxt = grid[op[2][0]][op[2][1]]
                if not active.get(op[1], False):
                    patches[op[1]] = nxt.copy()
                    active[op[1]] = True
                    continue
                res = multiply(patches[op[1]], nxt)
                if phase_seconds is not None:
                    phase_seconds[depth] = phase_seconds.get(depth, 0.0) + time.perf_counter() - t_op
                if k == len(rest):
                    output = res
                    continue
                patches[op[1]] = res
            elif op[0] == "cut":
                name = cut_name(op[1])
                pa = op[1][0]
                copy_a = grid[pa[0]][pa[1]].copy()
                values = list(op[2]) or list(range(copy_a.index_to_dim()[name]))
                suffix = rest[k:]
                if len(op[1]) > 1:
                    pb = op[1][1]
                    copy_b = grid[pb[0]][pb[1]].copy()
                    for v in values:
                        grid[pa[0]][pa[1]] = copy_a.project(name, v)
                        grid[pb[0]][pb[1]] = copy_b.project(name, v)
                        run(suffix, out_idx, dict(active), depth + 1)
                    grid[pb[0]][pb[1]] = copy_b
                else:
                    out_idx *= len(values)
                    for v in values:
                        grid[pa[0]][pa[1]] = copy_a.project(name, v)
                        run(suffix, out_idx, dict(active), depth + 1)
                        out_idx += 1
                grid[pa[0]][pa[1]] = copy_a
                return
            else:
                src, tgt = op[1], op[2]
                if not active.get(tgt, False):
                    patches[tgt] = patches[src].copy()
                    active[tgt] = True
                    continue
                res = multiply(patches[src], patches[tgt])
                if phase_seconds is not None:
                    phase_seconds[depth] = phase_seconds.get(depth, 0.0) + time.perf_counter() - t_op
                if k == len(rest):
                    output = res
                    continue
                patches[tgt] = res
        if output is None or output.size != 1:
            raise OracleError("Contraction did not complete; final tensor is: "
                              + (str(output.indices) if output is not None else "none"))
        amplitudes[out_idx] += complex(output.data.reshape(-1)[0])

    run(list(ops), 0, {})


def find_output_pos(I, J, off, pos):
    """src/evaluate_circuit.cpp:22-36."""
    p = pos[0] * J + pos[1]
    for o in off:
        if o[0] < pos[0] or (o[0] == pos[0] and o[1] < pos[1]):
            p -= 1
    return p


def get_output_states(I, J, off, initial, final, ops):
    """src/evaluate_circuit.cpp:38-104."""
    base = final
    unspecified = False
    if not final:
        unspecified = True
        base = "0" * len(initial)
    base = list(base)
    pos_map, val_map, final_qubits = [], [], []
    for op in ops:
        if op[0] != "cut" or len(op[1]) != 1:
            continue
        p = find_output_pos(I, J, off, op[1][0])
        if unspecified:
            base[p] = "x"
        pos_map.append(p)
        val_map.append(list(op[2]) if op[2] else [0, 1])
        final_qubits.append(tuple(op[1][0]))
    states = ["".join(base)]
    for k in range(len(final_qubits)):
        nxt = []
        for s in states:
            for v in val_map[k]:
                t = list(s)
                t[pos_map[k]] = str(v)[0]
                nxt.append("".join(t))
        states = nxt
    for i, c in enumerate(states[0]):
        if c not in "01":
            raise OracleError(f"Final state has non-binary character {c} at index {i}")
    return "".join(base), final_qubits, states


def evaluate_circuit(circuit_text, ordering_text, grid_text, initial="", final="",
                     dtype=np.complex128, memory_limit=None, phase_seconds=None):
    """EvaluateCircuit (src/evaluate_circuit.cpp:106-253).  phase_seconds
    (optional dict) receives "build" (network + flatten) and the per-depth
    contraction times of contract_grid."""
    import time
    t0 = time.perf_counter()
    grid = Grid(grid_text)
    ops = parse_ordering(ordering_lines(ordering_text), grid.I, grid.J, grid.off)
    validate_ordering(ops)
    circuit = Circuit(circuit_text)
    n_active = grid.I * grid.J - len(grid.off)
    if not initial:
        initial = "0" * n_active
    base, final_qubits, states = get_output_states(
        grid.I, grid.J, grid.off, initial, final, ops)
    g3 = circuit_to_network(circuit, grid.I, grid.J, initial, base, final_qubits,
                            grid.off, dtype)
    g2 = flatten(g3, grid.I, grid.J, final_qubits, grid.off, ops)
    if phase_seconds is not None:
        phase_seconds["build"] = time.perf_counter() - t0
    if memory_limit is not None:
        plan(ops, g2, memory_limit)
    amps = np.zeros(len(states), dtype=np.complex128)
    contract_grid(ops, g2, amps, phase_seconds)
    return list(zip(states, amps))
