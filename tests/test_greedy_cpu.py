"""The greedy ordering and cut search (qflex_amd.greedy, after the reference's
qflexcirq/ordering/order_circuit_simulation.py:87-106,261-332,393-421):
valid orderings for the engine's parser and validator, step costs that agree
with an independent replay of the ordering text, the search's choices on
graphs small enough to decide by hand or by brute force, and match_fidelity
against exhaustive search."""
import itertools
import math

import numpy as np
import pytest

from qflex_amd import circuits, greedy, qflex


def replay_cost(order_text, circ_text):
    """Sum of m*n*k over the pairwise products of an ordering, from the
    ordering text alone: each patch is a set of bond names (the grid edges
    it has one end of) with dimensions 2^bits; cut bonds are removed; a
    product's m, n, k are the sizes of the indices only in the patch, only
    in the operand, and in both (products with a rank-0 side are free, as in
    the engine's planner).  Returns (cost of one slice, number of slices)."""
    bits, _ = greedy.bond_bits_of_circuit(circ_text)
    slices = 1
    cut = set()
    for line in order_text.splitlines():
        w = line.split()
        if w[0] == "cut":
            a, b = sorted((int(w[2]), int(w[3])))
            cut.add((a, b))
            vals = w[1].strip("()")
            slices *= len(vals.split(",")) if vals else 2 ** bits[(a, b)]
    site = {}
    for (a, b), bb in bits.items():
        if (a, b) in cut:
            continue
        site.setdefault(a, set()).add((a, b))
        site.setdefault(b, set()).add((a, b))
    dim = lambda idx: math.prod(2 ** bits[e] for e in idx)
    patches, cost = {}, 0

    def product(x, y):
        if x and y:
            return dim(x - y) * dim(y - x) * dim(x & y)
        return 0

    for line in order_text.splitlines():
        w = line.split()
        if w[0] == "expand":
            p, q = w[1], int(w[2])
            cur = patches.get(p, set())
            s = site.get(q, set())
            cost += product(cur, s)
            patches[p] = cur ^ s
        elif w[0] == "merge":
            a, b = w[1], w[2]
            cost += product(patches[a], patches[b])
            patches[b] = patches.pop(a) ^ patches[b]
    assert len(patches) == 1 and not next(iter(patches.values()))
    return cost, slices


CASES = [(3, 3, 8, 1, ()), (4, 4, 16, 2, ()), (5, 4, 16, 3, ()), (4, 5, 24, 4, ((1, 1),)),
         (6, 6, 24, 5, ()), (7, 7, 32, 0, ())]


@pytest.mark.parametrize("I,J,depth,seed,off", CASES)
@pytest.mark.parametrize("max_cuts,per_slice", [(0, True), (1, True), (2, True), (2, False)])
def test_greedy_orderings_are_valid_and_costed(I, J, depth, seed, off, max_cuts, per_slice):
    circ = circuits.random_circuit(I, J, depth, seed, off=off)
    text, info = greedy.greedy_ordering(circ, max_cuts=max_cuts, per_slice=per_slice)
    lines = text.splitlines()
    ops = qflex.parse_ordering(lines, circuits.grid_text(I, J, off).splitlines())
    qflex.validate_ordering(ops)
    active = sorted(i * J + j for i in range(I) for j in range(J) if (i, j) not in off)
    assert sorted(int(l.split()[2]) for l in lines if l.startswith("expand")) == active
    # cuts come before any patch exists
    ncut = sum(l.startswith("cut") for l in lines)
    assert all(l.startswith("cut") for l in lines[:ncut]) and ncut == len(info["cuts"])
    assert ncut <= max_cuts
    cost, slices = replay_cost(text, circ)
    assert cost == info["slice_time"] and slices == info["slices"]


def test_cuts_lower_the_criterion_they_are_chosen_by():
    for I, J, depth, seed, off in CASES[2:]:
        circ = circuits.random_circuit(I, J, depth, seed, off=off)
        _, none = greedy.greedy_ordering(circ, max_cuts=0)
        _, per = greedy.greedy_ordering(circ, max_cuts=2, per_slice=True)
        _, tot = greedy.greedy_ordering(circ, max_cuts=2, per_slice=False)
        assert per["slice_time"] <= none["slice_time"]
        assert tot["slice_time"] * tot["slices"] <= none["slice_time"]


def test_step_choice_on_a_path_by_hand():
    # path 0 -(1)- 1 -(3)- 2 -(1)- 3: contracting the middle bond first
    # leaves a result of 2^2 from inputs 2^4 + 2^4 (ratio 1/8); an end bond
    # gives 2^3 from 2^1 + 2^4 (ratio 8/18).  Then the two end bonds (now
    # equal), the one entered first first.
    total, steps = greedy.greedy_steps([0, 1, 2, 3], [((0, 1), 1), ((1, 2), 3), ((2, 3), 1)])
    assert [s[0] for s in steps] == [(1, 2), (0, 1), (2, 3)]
    # times: 2^(4+4-3) = 32, then {0}(rank 1) with {1,2}(rank 2): 2^(1+2-1) = 4,
    # then the rank-1 patch with {3}: 2^(1+1-1) = 2
    assert total == 32 + 4 + 2


def test_greedy_matches_brute_force_rule_on_small_graphs():
    # every step is the lexicographic minimum of (heuristic, estimate, time)
    # recomputed from scratch on the current graph (an independent,
    # quadratic restatement: node sets and bond dims as plain dicts)
    rng = np.random.default_rng(3)
    for trial in range(30):
        n = int(rng.integers(3, 8))
        pairs = [(a, b) for a in range(n) for b in range(a + 1, n) if rng.random() < 0.5]
        if not pairs:
            continue
        bonds = [(p, int(rng.integers(1, 4))) for p in pairs]
        _, steps = greedy.greedy_steps(list(range(n)), bonds)
        nodes = {frozenset([q]) for q in range(n)}
        edge = {}
        for (a, b), d in bonds:
            edge[frozenset([frozenset([a]), frozenset([b])])] = d
        for origin, u_members, v_members in steps:
            def cost(key):
                u, v = tuple(key)
                rank = lambda x: sum(d for k, d in edge.items() if x in k)
                rem = [(k, d) for k, d in edge.items() if (u in k or v in k) and k != key]
                rb = sum(d for _, d in rem)
                rn = {x for k, _ in rem for x in k} - {u, v}
                res = 2.0 ** rb
                return (res / (2.0 ** rank(u) + 2.0 ** rank(v)), len(rn) / res if rn else 0.0,
                        2 ** (rank(u) + rank(v) - edge[key]))
            best = min(cost(k) for k in edge)
            key = frozenset([u_members, v_members])
            assert cost(key) == best, (trial, origin)
            u, v = u_members, v_members
            new = u | v
            merged = {}
            for k, d in edge.items():
                if k == key:
                    continue
                k2 = frozenset(new if x in (u, v) else x for x in k)
                merged[k2] = merged.get(k2, 0) + d
            edge = merged
            nodes = (nodes - {u, v}) | {new}


@pytest.mark.parametrize("target,dims,want", [(0.5, [4, 4], (0.5, [2, 4])),
                                              (0.3, [16], (0.3125, [5])),
                                              (1.0, [16, 16], (1.0, [16, 16])),
                                              (0.01, [], (1.0, []))])
def test_match_fidelity_known_answers(target, dims, want):
    assert greedy.match_fidelity(target, dims) == want


def test_match_fidelity_is_the_smallest_fidelity_above_target():
    rng = np.random.default_rng(1)
    for _ in range(200):
        dims = [int(2 ** rng.integers(1, 5)) for _ in range(int(rng.integers(1, 4)))]
        target = float(rng.uniform(0.05, 0.999))
        f, widths = greedy.match_fidelity(target, dims)
        brute = min(math.prod(w) / math.prod(dims)
                    for w in itertools.product(*[range(1, d + 1) for d in dims])
                    if math.prod(w) / math.prod(dims) >= target)
        assert abs(f - brute) < 1e-15 and f >= target
        assert abs(math.prod(widths) / math.prod(dims) - f) < 1e-15


def test_low_fidelity_cut_lines():
    circ = circuits.random_circuit(5, 4, 16, 3)
    text, info = greedy.greedy_ordering(circ, max_cuts=2, fidelity=0.5)
    cuts = [l for l in text.splitlines() if l.startswith("cut")]
    assert len(cuts) == 2 and info["fidelity"] == 0.5
    widths = [len(c.split()[1].strip("()").split(",")) if c.split()[1] != "()" else 4
              for c in cuts]
    assert math.prod(widths) == 8 and info["slices"] == 8
    ops = qflex.parse_ordering(text.splitlines(), circuits.grid_text(5, 4).splitlines())
    assert [o[2] for o in ops if o[0] == "cut"] in ([[0, 1], []], [[], [0, 1]])


def test_bond_bits_count_gates_per_edge():
    bits, order = greedy.bond_bits_of_circuit(
        "3\n0 h 0\n1 cz 0 1\n2 fsim(0.1,0.2) 1 2\n3 cz 1 0\n4 t 2\n")
    assert bits == {(0, 1): 2, (1, 2): 2} and order == [(0, 1), (1, 2)]
    with pytest.raises(ValueError):
        greedy.bond_bits_of_circuit("2\n0 iswap 0 1\n")
