"""Seeded random-circuit and ordering generators for grid workloads.

Writes circuits, grids and orderings in qFlex's text formats
(reference docs/input_formats.md; parsers src/circuit.cpp, src/grid.cpp,
src/ordering.cpp), so every workload here is also a valid input for the
reference.

Circuit family ("8-pattern", the reference's rectangular_IxJ_1-d-1 family,
e.g. reference config/circuits/rectangular_7x7_1-32-1_0.txt):
  * cycle 0 and cycle d+1: H on every active qubit;
  * cycle c in 1..d applies CZ on the edges of pattern (c-1) mod 8 of
    [H0, V1, H1, V0, H2, V3, H3, V2], where H_s = horizontal edges
    (i,j)-(i,j+1) with j = 2i+s (mod 4) and V_s = vertical edges
    (i,j)-(i+1,j) with i = 2j+s (mod 4);
  * gates="uniform": a qubit idle in cycle c that was in a CZ in cycle c-1
    gets a one-qubit gate drawn uniformly (seeded) from {x_1_2, y_1_2, t};
  * gates="xyt": as in the reference's circuit files, that qubit gets
    x_1_2 or y_1_2 (seeded coin), and a T in the next cycle if it is still
    idle; the first idle cycle after the opening H gets a T.
Each edge is active once per 8 cycles, so after bundling every grid bond has
dimension 2^(d/8): 16 at depth 32.

Circuit family ("abcd"): patterns A, B, C, D = horizontal j even, horizontal
j odd, vertical i even, vertical i odd in the order ABCDCDAB, with a
one-qubit layer on every qubit before each CZ layer (the literal
"ABCDCDAB" rule is degenerate without it: interior qubits are never idle).
"""
import random

_ONE_Q = ("x_1_2", "y_1_2", "t")


def _pattern_edges(I, J, kind, s, off):
    e = []
    for i in range(I):
        for j in range(J):
            if kind == "H" and j + 1 < J and (j - 2 * i - s) % 4 == 0:
                e.append(((i, j), (i, j + 1)))
            if kind == "V" and i + 1 < I and (i - 2 * j - s) % 4 == 0:
                e.append(((i, j), (i + 1, j)))
    return [(a, b) for a, b in e if a not in off and b not in off]


_EIGHT = [("H", 0), ("V", 1), ("H", 1), ("V", 0), ("H", 2), ("V", 3), ("H", 3), ("V", 2)]


def _abcd_edges(I, J, letter, off):
    e = []
    for i in range(I):
        for j in range(J):
            if letter in "AB" and j + 1 < J and j % 2 == "AB".index(letter):
                e.append(((i, j), (i, j + 1)))
            if letter in "CD" and i + 1 < I and i % 2 == "CD".index(letter):
                e.append(((i, j), (i + 1, j)))
    return [(a, b) for a, b in e if a not in off and b not in off]


def random_circuit(I, J, depth, seed, off=(), layout="8pattern", gates="xyt"):
    """Returns the circuit text (qFlex format) of a seeded grid circuit."""
    rng = random.Random(seed)
    off = set(map(tuple, off))
    active = [(i, j) for i in range(I) for j in range(J) if (i, j) not in off]
    q = lambda p: p[0] * J + p[1]
    lines = [str(len(active))]
    for p in active:
        lines.append(f"0 h {q(p)}")
    if layout == "8pattern":
        prev_cz = set()
        prev_xy = set()
        for c in range(1, depth + 1):
            kind, s = _EIGHT[(c - 1) % 8]
            edges = _pattern_edges(I, J, kind, s, off)
            busy = set()
            for a, b in edges:
                lines.append(f"{c} cz {q(a)} {q(b)}")
                busy.update((a, b))
            now_xy = set()
            for p in active:
                if p in busy:
                    continue
                if gates == "uniform":
                    if p in prev_cz:
                        lines.append(f"{c} {rng.choice(_ONE_Q)} {q(p)}")
                elif p in prev_cz:
                    lines.append(f"{c} {rng.choice(_ONE_Q[:2])} {q(p)}")
                    now_xy.add(p)
                elif p in prev_xy or c == 1:
                    lines.append(f"{c} t {q(p)}")
            prev_cz, prev_xy = busy, now_xy
        last = depth + 1
    elif layout == "abcd":
        seq = "ABCDCDAB"
        cyc = 1
        for c in range(depth):
            for p in active:
                lines.append(f"{cyc} {rng.choice(_ONE_Q)} {q(p)}")
            cyc += 1
            for a, b in _abcd_edges(I, J, seq[c % 8], off):
                lines.append(f"{cyc} cz {q(a)} {q(b)}")
            cyc += 1
        last = cyc
    else:
        raise ValueError(layout)
    for p in active:
        lines.append(f"{last} h {q(p)}")
    return "\n".join(lines) + "\n"


def grid_text(I, J, off=()):
    off = set(map(tuple, off))
    return "\n".join(" ".join("0" if (i, j) in off else "1" for j in range(J))
                     for i in range(I)) + "\n"


def snake_ordering(I, J, off=(), final_region=()):
    """One patch grown row by row, alternating direction.  Terminal cuts for
    the final region come first so the (o) index is projected before use."""
    off = set(map(tuple, off))
    lines = []
    for p in final_region:
        lines.append(f"cut () {p[0] * J + p[1]}")
    for i in range(I):
        cols = range(J) if i % 2 == 0 else range(J - 1, -1, -1)
        for j in cols:
            if (i, j) not in off:
                lines.append(f"expand a {i * J + j}")
    return "\n".join(lines) + "\n"


# Corner-inward ordering for the 7x7 grid.  Corner patches A (top-left 3x3),
# pB (bottom-left 4x2) and ppD (bottom-right 3x3) are built once; two cuts
# on the (6,2)-(6,3) and (2,6)-(3,6) bonds give 256 slices; the final region
# C = rows 0-2, cols 3-4 carries six terminal cuts (64 outputs per slice).
# This follows the shape of the reference's 7x7 example
# (tests/src/contraction_utils_test.cpp:306-374), with the second cut placed
# before any patch touches (2,6) (that example expands (2,6) into pC before
# cutting it, which the reference's CUT cannot project).  After the big
# B x D merge, column 3 of the final region joins the big patch and column 4
# joins pC, and the two meet in the last step: 2.44e14 flop per 64 outputs
# (an optimal-tree search over the per-slice items finds nothing cheaper than
# 2.31e14 with these cuts).
CORNER_7X7_FINAL = [(0, 3), (0, 4), (1, 3), (1, 4), (2, 3), (2, 4)]
CORNER_7X7_CUTS = [((6, 2), (6, 3)), ((2, 6), (3, 6))]


def corner_inward_7x7(cut_values=(None, None)):
    """The ordering text.  cut_values optionally restricts the two bond
    cuts to subsets of their 16 values (a sub-job over fewer slices)."""
    J = 7
    q = lambda i, j: i * J + j
    L = []
    ex = lambda p, sites: L.extend(f"expand {p} {q(*s)}" for s in sites)
    vals = ["(" + ",".join(str(v) for v in cv) + ")" if cv is not None else "()"
            for cv in cut_values]
    ex("A", [(0, 0), (0, 1), (1, 0), (1, 1), (0, 2), (2, 0), (1, 2), (2, 1), (2, 2)])
    ex("pB", [(6, 0), (5, 0), (4, 0), (3, 0), (6, 1), (5, 1), (4, 1), (3, 1)])
    ex("ppD", [(6, 6), (6, 5), (5, 6), (5, 5), (6, 4), (4, 6), (5, 4), (4, 5), (4, 4)])
    L.append(f"cut {vals[0]} {q(6, 2)} {q(6, 3)}")
    L.append("merge pB B")
    ex("B", [(6, 2), (5, 2), (4, 2), (3, 2)])
    L.append("merge A B")
    L.append("merge ppD pD")
    ex("pD", [(6, 3), (5, 3), (4, 3)])
    L.append(f"cut {vals[1]} {q(2, 6)} {q(3, 6)}")
    L.append("merge pD D")
    ex("D", [(3, 6), (3, 5), (3, 4), (3, 3)])
    L.append("merge B D")
    ex("pC", [(0, 6), (1, 6), (2, 6), (0, 5), (1, 5), (2, 5)])
    for s in CORNER_7X7_FINAL:
        L.append(f"cut () {q(*s)}")
    L.append("merge D E")
    ex("E", [(2, 3), (1, 3), (0, 3)])
    L.append("merge pC F")
    ex("F", [(0, 4), (1, 4), (2, 4)])
    L.append("merge F E")
    return "\n".join(L) + "\n"


def sample_bitstrings(psi, n, seed):
    """Draws n bitstrings from |psi|^2 (psi shaped (2,)*q), seeded."""
    import numpy as np
    p = np.abs(psi.reshape(-1)) ** 2
    p = p / p.sum()
    rng = np.random.default_rng(seed)
    idx = rng.choice(p.size, size=n, replace=False, p=p)
    nq = psi.ndim
    return [format(int(k), f"0{nq}b") for k in idx]
