"""Greedy contraction ordering and cut search for grid circuits.

The reference's automatic ordering (qflexcirq/ordering/
order_circuit_simulation.py) works on a graph whose nodes are patches of
qubits and whose bonds carry the log2 dimension of the two-qubit gates
between them; it contracts, one bond at a time, the bond with the smallest
(heuristic, estimate, time) triple (`Bond.get_cost`, :87-106;
`get_steps_for_graph`, :261-288), tries each gate's qubit pair as a cut and
keeps the cut that lowers the contraction cost most, up to `max_cuts`
(:393-421), and narrows cut widths to a target fidelity (`match_fidelity`,
:291-332).  It needs cirq; this module states the same search on the
circuit text this package reads, with its own graph structures:

* nodes are integer ids with a qubit set; bonds are undirected node pairs
  with a log2 dimension (parallel bonds merge, their dimensions add);
* a gate adds log2 of its Schmidt rank to the bond of its qubit pair
  (cz, cx: 1; fsim: 2), one-qubit gates add nothing;
* step cost for a bond (u, v) of log2 dimension d, ranks ru, rv (sums of
  incident bond dimensions), remote bonds R (the other bonds of u and v) and
  remote nodes N (their other ends):
      heuristic = 2^sum(R) / (2^ru + 2^rv)     smaller result per input
      estimate  = |N| / 2^sum(R)               favour more, lighter bonds
      time      = 2^(ru + rv - d)              m * n * k of the product
  the smallest triple in lexicographic order wins; among equal triples the
  bond that has been in the graph longest (remote bonds are re-entered
  after each contraction) wins;
* the total time is the sum of the step times, i.e. m*n*k summed over the
  pairwise products (8x that is the complex-GEMM flop count the engine's
  planner reports, times the number of slices when cuts are made).

`greedy_ordering` returns the ordering text (cuts first, then expand/merge
lines) in the grammar of the engine's ordering parser, so every result goes
through the same validation as hand-written orderings.
"""
import itertools
import math

# log2 of the Schmidt rank of each two-qubit gate of the circuit format
GATE_BOND_BITS = {"cz": 1, "cx": 1, "fsim": 2}


def bond_bits_of_circuit(circuit_text):
    """{(q1, q2) with q1 < q2: log2 bond dimension} summed over the
    circuit's two-qubit gates, and the qubit pairs in first-gate order."""
    bits, order = {}, []
    lines = [l.split("#")[0].strip() for l in circuit_text.splitlines()]
    lines = [l for l in lines if l]
    for line in lines[1:]:
        words = line.replace("(", " (").split()
        name = words[1]
        qubits = [int(w) for w in words[2:] if not w.startswith("(")]
        if len(qubits) != 2:
            continue
        key = (min(qubits), max(qubits))
        b = GATE_BOND_BITS.get(name)
        if b is None:
            raise ValueError(f"no bond dimension known for two-qubit gate {name!r}")
        if key not in bits:
            bits[key] = 0
            order.append(key)
        bits[key] += b
    return bits, order


class BondGraph:
    """Patches (nodes) and the bonds between them."""

    def __init__(self, qubits, bonds):
        # node id -> frozenset of qubits; qubits start as their own nodes
        self.members = {q: frozenset([q]) for q in qubits}
        self.next_id = max(qubits) + 1 if qubits else 0
        # bond key (frozenset of two node ids) -> [log2 dim, origin qubit pair]
        self.bonds = {}
        self.adj = {q: set() for q in qubits}
        for (a, b), bits in bonds:
            self._add(a, b, bits, (a, b))

    def _add(self, u, v, bits, origin):
        key = frozenset((u, v))
        if key in self.bonds:
            self.bonds[key][0] += bits
        else:
            self.bonds[key] = [bits, origin]
            self.adj[u].add(v)
            self.adj[v].add(u)

    def rank(self, u):
        return sum(self.bonds[frozenset((u, w))][0] for w in self.adj[u])

    def cost(self, key):
        u, v = tuple(key)
        d = self.bonds[key][0]
        remote_bits, remote_nodes = 0, set()
        for x, y in ((u, v), (v, u)):
            for w in self.adj[x]:
                if w != y:
                    remote_bits += self.bonds[frozenset((x, w))][0]
                    remote_nodes.add(w)
        ru, rv = self.rank(u), self.rank(v)
        result = 2.0 ** remote_bits
        heuristic = result / (2.0 ** ru + 2.0 ** rv)
        estimate = len(remote_nodes) / result if remote_nodes else 0.0
        return (heuristic, estimate, 2 ** (ru + rv - d))

    def contract(self, key):
        """Merges the two nodes of the bond into a new node; their other
        bonds move to it (bonds to a common neighbour merge) and go to the
        end of the bond order."""
        u, v = tuple(key)
        new = self.next_id
        self.next_id += 1
        self.members[new] = self.members.pop(u) | self.members.pop(v)
        self.adj[new] = set()
        del self.bonds[key]
        moved = []
        for x in (u, v):
            for w in sorted(self.adj[x]):
                if w in (u, v):
                    continue
                bits, origin = self.bonds.pop(frozenset((x, w)))
                self.adj[w].discard(x)
                moved.append((w, bits, origin))
            del self.adj[x]
        for w, bits, origin in moved:
            self._add(new, w, bits, origin)
        return new


def greedy_steps(qubits, bonds):
    """(total time, [(origin qubit pair, u members, v members)]) of the
    greedy contraction of the bond graph."""
    g = BondGraph(qubits, bonds)
    total, steps = 0, []
    while g.bonds:
        best_key, best = None, None
        for key in g.bonds:
            c = g.cost(key)
            if best is None or c < best:
                best, best_key = c, key
        u, v = tuple(best_key)
        steps.append((g.bonds[best_key][1], g.members[u], g.members[v]))
        g.contract(best_key)
        total += best[2]
    return total, steps


def greedy_cuts(qubits, bonds, max_cuts, per_slice=True):
    """Greedy cut search.  Each round tries every remaining bond as a cut
    and keeps the one whose cut graph contracts cheapest, if that beats the
    current cost.  With per_slice (the reference's criterion) the cost of a
    cut graph is the cost of one slice; otherwise it is the cost of all
    slices (one slice times the product of the cut bonds' dimensions).
    Returns (cuts as [(pair, log2 dim)], time of one slice, steps)."""
    cuts = []
    best_cost, best_steps = greedy_steps(qubits, bonds)
    for _ in range(max_cuts):
        chosen = None
        cut_keys = {p for p, _ in cuts}
        slices = 2 ** sum(b for _, b in cuts)
        for pair, bits in bonds:
            if pair in cut_keys:
                continue
            rest = [(p, b) for p, b in bonds if p not in cut_keys and p != pair]
            cost, steps = greedy_steps(qubits, rest)
            score = cost if per_slice else cost * slices * 2 ** bits
            ref = best_cost if per_slice else best_cost * slices
            if score < ref and (chosen is None or score < chosen[0]):
                chosen = (score, pair, bits, cost, steps)
        if chosen is None:
            break
        _, pair, bits, best_cost, best_steps = chosen
        cuts.append((pair, bits))
    return cuts, best_cost, best_steps


def match_fidelity(target, dims):
    """Widths w_i <= dims[i] (values of each cut to evaluate) whose
    fidelity prod(w_i) / prod(dims) is the smallest value >= target, trying
    width sets in lexicographic order from the smallest admissible widths;
    (fidelity, widths).  target 1 (or no cuts) keeps every value."""
    if target >= 1.0 or not dims:
        return 1.0, list(dims)
    total = math.prod(dims)
    ranges = [range(min(d, math.ceil(d * target)), d + 1) for d in dims]
    best_f, best_w = 1.0, list(dims)
    for widths in itertools.product(*ranges):
        f = math.prod(widths) / total
        if target <= f < best_f:
            best_f, best_w = f, list(widths)
            if f == target:
                break
    return best_f, best_w


def ordering_from_steps(steps, all_qubits):
    """Expand / merge lines for greedy steps: a step whose two sides are
    fresh qubits opens a new patch with both; a fresh qubit next to a patch
    is expanded into it; two patches merge (the lower name into the higher).
    Qubits no step reaches are expanded into their own patches, and the
    patches left at the end are merged into one."""
    patch_of, lines, names = {}, [], iter(_patch_names())
    members = {}
    for (a, b), _, _ in steps:
        pa, pb = patch_of.get(a), patch_of.get(b)
        if pa is None and pb is None:
            p = next(names)
            members[p] = []
            for q in (a, b):
                lines.append(f"expand {p} {q}")
                patch_of[q] = p
                members[p].append(q)
        elif pa is None or pb is None:
            p, q = (pb, a) if pa is None else (pa, b)
            lines.append(f"expand {p} {q}")
            patch_of[q] = p
            members[p].append(q)
        elif pa != pb:
            lo, hi = sorted((pa, pb))
            lines.append(f"merge {lo} {hi}")
            for q in members.pop(lo):
                patch_of[q] = hi
                members[hi].append(q)
    for q in all_qubits:
        if q not in patch_of:
            p = next(names)
            lines.append(f"expand {p} {q}")
            patch_of[q] = p
            members[p] = [q]
    rest = sorted(members)
    for lo, hi in zip(rest, rest[1:]):
        lines.append(f"merge {lo} {hi}")
    return lines


def _patch_names():
    for n in itertools.count(1):
        for t in itertools.product("ABCDEFGHIJKLMNOPQRSTUVWXYZ", repeat=n):
            yield "".join(t)


def greedy_ordering(circuit_text, max_cuts=2, fidelity=1.0, per_slice=True):
    """Ordering text for the circuit: cut lines (with value lists when
    fidelity < 1 narrows them), then the greedy expand/merge sequence.
    Returns (text, info) with info = {"cuts", "slice_time", "fidelity",
    "slices"}."""
    bits, order = bond_bits_of_circuit(circuit_text)
    first = [l.split("#")[0].strip() for l in circuit_text.splitlines()]
    first = [l for l in first if l]
    qubits = sorted({q for pair in bits for q in pair} |
                    {int(w) for l in first[1:] for w in l.replace("(", " (").split()[2:]
                     if not w.startswith("(")})
    bonds = [(p, bits[p]) for p in order]
    cuts, slice_time, steps = greedy_cuts(qubits, bonds, max_cuts, per_slice)
    fid, widths = match_fidelity(fidelity, [2 ** b for _, b in cuts])
    lines = []
    for ((a, b), bb), w in zip(cuts, widths):
        values = "()" if w == 2 ** bb else "(" + ",".join(str(v) for v in range(w)) + ")"
        lines.append(f"cut {values} {a} {b}")
    lines += ordering_from_steps(steps, qubits)
    info = {"cuts": [[list(p), b] for p, b in cuts], "slice_time": slice_time,
            "fidelity": fid, "slices": math.prod(widths) if widths else 1}
    return "\n".join(lines) + "\n", info
