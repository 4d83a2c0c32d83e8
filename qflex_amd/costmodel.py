"""Learned cost model for contraction orderings of grid circuits.

Scores candidate (ordering, cut placement) pairs with a small MLP trained on
timings the engine itself produces; amplitudes always come from the exact
contraction of whichever ordering is chosen, so the model only affects speed.
The plain (row snake) ordering is the fallback.

* candidates    -- seeded orderings of an I x J grid: one patch swept
                   along rows or columns (snake or raster, from any corner),
                   or two patches split along a row/column, each swept, then
                   merged, with 0-2 bond cuts across the split; plus the
                   greedy bond-contraction orderings with 0-2 greedy cuts
                   (qflex_amd.greedy, the search of the reference's
                   qflexcirq/ordering/order_circuit_simulation.py:261-421).
* features      -- 64 floats from the engine's symbolic plan (qflex.plan: the
                   same pass that plans device memory): log2 flops, peak
                   tensor rank and size, per-step GEMM shapes pooled.
* the network   -- MLP 64 -> 256 -> 256 -> 1 in BF16 with FP32
                   accumulation (qfx_mlp_* kernels), MSE on log time, Adam.

Command line (one GPU): ``python -m qflex_amd.costmodel --train-circuits 24
--candidates 48 --heldout 200`` prints a JSON report: on held-out circuits,
the measured time of the chosen ordering vs the plain ordering's.
"""
import argparse
import json
import math
import sys

import numpy as np

from . import capi, circuits

N_FEATURES = capi.MLP_IN


# ---------------------------------------------------------------- candidates

def _sweep(sites, axis, snake, flip_rows, flip_cols):
    """Visit order of a rectangle of sites, line by line along `axis`."""
    rows = sorted({i for i, _ in sites}, reverse=flip_rows)
    cols = sorted({j for _, j in sites}, reverse=flip_cols)
    lines = [[(i, j) for j in cols] for i in rows] if axis == 0 else \
            [[(i, j) for i in rows] for j in cols]
    out = []
    for n, line in enumerate(lines):
        out.extend(reversed(line) if snake and n % 2 else line)
    return [s for s in out if s in sites]


def plain_ordering(I, J):
    """The fallback: one patch, rows in snake order."""
    return circuits.snake_ordering(I, J)


def random_candidate(I, J, rng):
    """One seeded candidate ordering (text) for an I x J grid."""
    q = lambda s: s[0] * J + s[1]
    sites = {(i, j) for i in range(I) for j in range(J)}
    rnd = lambda: bool(rng.integers(2))
    if rng.random() < 0.25:
        order = _sweep(sites, int(rng.integers(2)), rnd(), rnd(), rnd())
        return "".join(f"expand a {q(s)}\n" for s in order)
    axis = int(rng.integers(2))  # 0: split between rows, 1: between columns
    size = I if axis == 0 else J
    cut_at = int(rng.integers(1, size))
    first = {s for s in sites if s[axis] < cut_at}
    second = sites - first
    lines = []
    # bond cuts across the split, before any site is expanded
    boundary = sorted(((i, cut_at - 1), (i, cut_at)) if axis == 1 else ((cut_at - 1, i), (cut_at, i))
                      for i in range(J if axis == 0 else I))
    ncut = int(rng.choice([0, 0, 1, 2]))
    for k in rng.choice(len(boundary), size=min(ncut, len(boundary)), replace=False):
        a, b = boundary[int(k)]
        lines.append(f"cut () {q(a)} {q(b)}")
    for name, part, toward in (("a", first, False), ("b", second, True)):
        # sweep each part from its far side toward the split
        if axis == 0:
            order = _sweep(part, int(rng.integers(2)), rnd(), toward, rnd())
        else:
            order = _sweep(part, int(rng.integers(2)), rnd(), rnd(), toward)
        lines.extend(f"expand {name} {q(s)}" for s in order)
    lines.append("merge a b")
    return "\n".join(lines) + "\n"


# ---------------------------------------------------------------- features

def _hist(values, weights, edges):
    h = np.zeros(len(edges) + 1)
    idx = np.searchsorted(edges, values, side="right")
    np.add.at(h, idx, weights)
    s = h.sum()
    return h / s if s > 0 else h


def features(plan, I, J, depth):
    """64 floats from qflex.plan(...) of one candidate."""
    st = plan["steps"]
    f = np.zeros(N_FEATURES, dtype=np.float64)
    if not st:
        return f.astype(np.float32)
    m = np.array([s["m"] for s in st], dtype=np.float64)
    n = np.array([s["n"] for s in st], dtype=np.float64)
    k = np.array([s["k"] for s in st], dtype=np.float64)
    mult = np.array([s["multiplicity"] * s.get("batch", 1) for s in st], dtype=np.float64)
    out = np.array([s["out_size"] for s in st], dtype=np.float64)
    rank = np.array([s["out_rank"] for s in st], dtype=np.float64)
    flops = 8.0 * m * n * k * mult
    moved = 8.0 * (m * k + k * n + m * n) * mult
    lg = lambda x: np.log2(np.maximum(x, 1.0))
    f[0] = lg(flops.sum())
    f[1] = lg(flops.max())
    f[2] = lg(out.max())
    f[3] = rank.max()
    f[4] = lg(len(st) + 1.0)
    f[5] = lg(moved.sum())
    f[6] = lg(mult.max())
    f[7] = lg(sum(plan["patch_max_size"].values()) + 1.0)
    edges = [4, 6, 8, 10, 12, 14, 16]
    f[8:16] = _hist(lg(m), flops, edges)
    f[16:24] = _hist(lg(n), flops, edges)
    f[24:32] = _hist(lg(k), flops, edges)
    f[32:40] = _hist(lg(flops), np.ones_like(flops), [10, 14, 18, 22, 26, 30, 34])
    top = np.sort(lg(flops))[::-1][:8]
    f[40:40 + len(top)] = top
    topo = np.sort(lg(out))[::-1][:8]
    f[48:48 + len(topo)] = topo
    f[56], f[57], f[58] = I, J, depth / 8.0
    f[59] = lg(plan["max_grid"])
    f[60] = lg(len(st) * mult.mean())
    small = np.minimum(np.minimum(m, n), k) < 64
    f[61] = flops[small].sum() / flops.sum()
    f[62] = flops[k < 64].sum() / flops.sum()
    f[63] = lg(moved.max())
    return f.astype(np.float32)


# ---------------------------------------------------------------- the model

class CostModel:
    """The MLP on the GPU plus feature / label normalisation."""

    def __init__(self, ctx, seed=0):
        self.ctx = ctx
        rng = np.random.default_rng(seed)
        p = np.zeros(capi.MLP_PARAMS, dtype=np.float32)
        i, h = capi.MLP_IN, capi.MLP_HIDDEN
        p[:i * h] = rng.standard_normal(i * h) * math.sqrt(2.0 / i)
        o = i * h + h
        p[o:o + h * h] = rng.standard_normal(h * h) * math.sqrt(2.0 / h)
        o += h * h + h
        p[o:o + h] = rng.standard_normal(h) * math.sqrt(1.0 / h)
        self.params = ctx.upload(p)
        self.mu = np.zeros(N_FEATURES, np.float32)
        self.sd = np.ones(N_FEATURES, np.float32)
        self.y_mu, self.y_sd = 0.0, 1.0

    def _norm(self, X):
        return ((np.asarray(X, np.float32) - self.mu) / self.sd).astype(np.float32)

    def fit(self, X, log_t, steps=600, lr=1e-3, batch=512, seed=0):
        """MSE on standardised log time, Adam, minibatches; returns losses."""
        X = np.asarray(X, np.float32)
        self.mu = X.mean(axis=0)
        self.sd = np.maximum(X.std(axis=0), 1e-3)
        y = np.asarray(log_t, np.float64)
        self.y_mu, self.y_sd = float(y.mean()), float(max(y.std(), 1e-6))
        Xn = self._norm(X)
        yn = ((y - self.y_mu) / self.y_sd).astype(np.float32)
        ctx = self.ctx
        G = ctx.alloc(4 * capi.MLP_PARAMS)
        M = ctx.upload(np.zeros(capi.MLP_PARAMS, np.float32))
        V = ctx.upload(np.zeros(capi.MLP_PARAMS, np.float32))
        L = ctx.alloc(4)
        rng = np.random.default_rng(seed)
        losses = []
        bsz = min(batch, len(yn))
        DX, DT = ctx.alloc(4 * bsz * N_FEATURES), ctx.alloc(4 * bsz)
        for step in range(1, steps + 1):
            idx = rng.choice(len(yn), size=bsz, replace=False)
            xb, tb = np.ascontiguousarray(Xn[idx]), np.ascontiguousarray(yn[idx])
            capi.check(capi.load().qfx_copy_h2d(ctx.handle, DX.ptr, xb.ctypes.data, xb.nbytes))
            capi.check(capi.load().qfx_copy_h2d(ctx.handle, DT.ptr, tb.ctypes.data, tb.nbytes))
            ctx.mlp_backward(bsz, DX, DT, self.params, G, L)
            ctx.mlp_adam(self.params, G, M, V, step, lr)
            if step % 50 == 0 or step == 1:
                losses.append(float(ctx.download(L, np.float32, (1,))[0]))
        for b in (G, M, V, L, DX, DT):
            b.free()
        return losses

    def predict_log_time(self, X):
        Xn = self._norm(X)
        ctx = self.ctx
        DX, DY = ctx.upload(Xn), ctx.alloc(4 * len(Xn))
        ctx.mlp_forward(len(Xn), DX, self.params, DY)
        y = ctx.download(DY, np.float32, (len(Xn),)).astype(np.float64)
        DX.free()
        DY.free()
        return y * self.y_sd + self.y_mu


# ---------------------------------------------------------------- driver

def options(circ_text, order_text, I, J, precision):
    return {"circuit": circ_text.splitlines(), "ordering": order_text.splitlines(),
            "grid": circuits.grid_text(I, J).splitlines(), "initial_state": "0" * (I * J),
            "final_state": "0" * (I * J), "memory_limit": 1 << 36, "precision": precision}


def time_contraction(qflex, opts, repeats=2):
    """Seconds of the contraction (ContractGrid, device work finished) of one
    EvaluateCircuit, best of `repeats`; the network build and flatten, the
    same for every ordering of a circuit, are not included.  An ordering whose
    tensors the engine refuses to allocate under the memory limit times as
    infinity."""
    best = float("inf")
    for _ in range(repeats):
        try:
            qflex.evaluate(opts)
        except RuntimeError as e:
            if "exceeds memory limit" in str(e):
                return float("inf")
            raise
        best = min(best, qflex.last_stats()["seconds"])
    return best


def candidate_set(I, J, rng, n):
    """n distinct candidates plus the plain ordering first."""
    seen = {plain_ordering(I, J)}
    out = [plain_ordering(I, J)]
    tries = 0
    while len(out) < n + 1 and tries < 20 * n:
        c = random_candidate(I, J, rng)
        tries += 1
        if c not in seen:
            seen.add(c)
            out.append(c)
    return out


def greedy_candidates(circ_text):
    """Orderings from the greedy bond-contraction search (qflex_amd.greedy):
    no cut, and 1 or 2 greedy cuts chosen by the cost of one slice (the
    reference's criterion) or of all slices; duplicates dropped."""
    from . import greedy
    out = []
    for max_cuts, per_slice in ((0, True), (1, True), (2, True), (1, False), (2, False)):
        text, _ = greedy.greedy_ordering(circ_text, max_cuts=max_cuts, per_slice=per_slice)
        if text not in out:
            out.append(text)
    return out


def feasible(plan, limit_elems):
    return all(s["out_size"] <= limit_elems for s in plan["steps"])


def plan_flops(plan):
    return sum(8.0 * s["m"] * s["n"] * s["k"] * s.get("batch", 1) * s["multiplicity"]
               for s in plan["steps"])


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def run(args):
    from . import qflex
    ctx = capi.Context(0)
    rng = np.random.default_rng(args.seed)
    # (I, J, depth): bond dimension 2^(depth/8) per grid edge
    shapes = [tuple(int(x) for x in s.split("x")) for s in args.shapes.split(",")]
    limit = 1 << args.max_log2_elems

    def ok(plan):
        return feasible(plan, limit) and plan_flops(plan) <= args.max_flops

    # training data: timed candidates of seeded circuits
    X, T = [], []
    for c in range(args.train_circuits):
        I, J, depth = shapes[c % len(shapes)]
        circ = circuits.random_circuit(I, J, depth, 1000 + c)
        for order in candidate_set(I, J, rng, args.candidates) + greedy_candidates(circ):
            opts = options(circ, order, I, J, args.precision)
            plan = qflex.plan(opts)
            if not ok(plan):
                continue
            t = time_contraction(qflex, opts, repeats=1)
            if not math.isfinite(t):
                continue
            X.append(features(plan, I, J, depth))
            T.append(math.log(t))
        log(f"[costmodel] train circuit {c + 1}/{args.train_circuits}: {len(T)} samples")
    model = CostModel(ctx, args.seed)
    losses = model.fit(np.array(X), np.array(T), steps=args.steps)
    log(f"[costmodel] fit: loss {losses[0]:.3f} -> {losses[-1]:.3f}")

    # held-out circuits: the model's pick vs the plain ordering (both timed)
    rows = []
    for c in range(args.heldout):
        I, J, depth = shapes[c % len(shapes)]
        circ = circuits.random_circuit(I, J, depth, 500000 + c)
        cands = candidate_set(I, J, rng, args.heldout_candidates)
        cands += [g for g in greedy_candidates(circ) if g not in cands]
        feats, usable, flops = [], [], []
        for order in cands:
            plan = qflex.plan(options(circ, order, I, J, args.precision))
            if ok(plan):
                feats.append(features(plan, I, J, depth))
                usable.append(order)
                flops.append(feats[-1][0])
        pred = model.predict_log_time(np.array(feats))
        chosen = usable[int(np.argmin(pred))]
        by_flops = usable[int(np.argmin(flops))]
        # the plain ordering is the baseline: timed whatever its size
        plain = cands[0]
        t_plain = time_contraction(qflex, options(circ, plain, I, J, args.precision))
        t_model = time_contraction(qflex, options(circ, chosen, I, J, args.precision))
        t_flops = time_contraction(qflex, options(circ, by_flops, I, J, args.precision))
        rows.append({"grid": [I, J], "depth": depth, "plain_s": t_plain, "model_s": t_model,
                     "min_flops_s": t_flops, "model_is_plain": chosen == plain})
        if (c + 1) % 10 == 0:
            log(f"[costmodel] held-out {c + 1}/{args.heldout}")
    t_plain = np.array([r["plain_s"] for r in rows])
    t_model = np.array([r["model_s"] for r in rows])
    t_flops = np.array([r["min_flops_s"] for r in rows])
    fin = np.isfinite(t_plain) & np.isfinite(t_model) & np.isfinite(t_flops)
    report = {
        "train_samples": len(T), "train_circuits": args.train_circuits,
        "candidates_per_circuit": args.candidates, "heldout_circuits": args.heldout,
        "heldout_candidates": args.heldout_candidates, "precision": args.precision,
        "fit_loss_first_last": [losses[0], losses[-1]],
        "geomean_speedup_vs_plain": float(np.exp(np.mean(np.log(t_plain[fin] / t_model[fin])))),
        "geomean_speedup_vs_min_flops": float(np.exp(np.mean(np.log(t_flops[fin] / t_model[fin])))),
        "total_plain_s": float(t_plain[fin].sum()), "total_model_s": float(t_model[fin].sum()),
        "total_min_flops_s": float(t_flops[fin].sum()),
        "model_not_slower_than_plain": float(np.mean(t_model[fin] <= t_plain[fin] * 1.05)),
        "model_chose_plain": float(np.mean([r["model_is_plain"] for r in rows])),
        "per_shape": {},
    }
    for I, J, depth in shapes:
        sel = [r for r in rows if r["grid"] == [I, J] and r["depth"] == depth]
        if sel:
            tp = np.array([r["plain_s"] for r in sel])
            tm = np.array([r["model_s"] for r in sel])
            keep = np.isfinite(tp) & np.isfinite(tm)
            tp, tm = tp[keep], tm[keep]
            if not len(tp):
                continue
            report["per_shape"][f"{I}x{J}x{depth}"] = {
                "circuits": len(sel), "mean_plain_ms": 1e3 * float(tp.mean()),
                "mean_model_ms": 1e3 * float(tm.mean()),
                "geomean_speedup": float(np.exp(np.mean(np.log(tp / tm))))}
    ctx.close()
    return report


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--train-circuits", type=int, default=24)
    ap.add_argument("--candidates", type=int, default=48)
    ap.add_argument("--heldout", type=int, default=200)
    ap.add_argument("--heldout-candidates", type=int, default=24)
    ap.add_argument("--steps", type=int, default=600)
    ap.add_argument("--precision", default="split_fp16")
    ap.add_argument("--shapes", default="5x6x32,6x7x24,6x6x24,7x7x24,5x5x32,6x5x32",
                    help="grids and depths IxJxDEPTH, cycled over the circuits")
    ap.add_argument("--max-log2-elems", type=int, default=28,
                    help="skip candidates whose largest tensor exceeds 2^N elements")
    ap.add_argument("--max-flops", type=float, default=3e13,
                    help="skip candidates above this many flops")
    ap.add_argument("--seed", type=int, default=0)
    print(json.dumps(run(ap.parse_args(argv))))


if __name__ == "__main__":
    main()
