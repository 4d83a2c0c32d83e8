#!/usr/bin/env python3
"""Amplitudes/s of qFlex's 7x7, depth 1+32+1 grid-circuit workload on one MI355X.

One step = one EvaluateCircuit through the native engine (module qflex:
C++ host + HIP kernels): build and flatten the network on the GPU, then
contract all 256 cut slices of the corner-inward ordering and deliver the
64 amplitudes of the open 2x3 output region (FP64 sums across slices).

Precision variants of the complex GEMMs: fp32 (FP32 MFMA), split_fp16
(operands scaled by powers of two, FP16 hi+lo, 3 products, FP32
accumulate), split_bf16x3 (BF16 hi+mid+lo, 6 products), split_bf16
(hi+mid, 3 products).  Each is timed; its norm-wise relative
error (and fidelity |<ref|psi>|^2 / (|ref|^2 |psi|^2)) is measured on the
full job against committed complex128 oracle amplitudes.  `value` is the
fastest variant whose error is <= 1e-5.

Printed: ONE JSON line (rank 0).  --gpus N > 1 is refused: this build runs
on one GPU.
"""
import argparse
import glob
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "amplitudes/s, 7x7 depth-32 grid: plain FP32 vs split-BF16 at rel.err≤1e-5"
PEAK_TFLOPS = {"fp32": 157.3, "split_fp16": 2500.0, "split_bf16x3": 2500.0, "split_bf16": 2500.0,
               "fp8": 5000.0}
HBM_GBPS = 8000.0
# The metric names split-BF16 (hi+mid, 3 products); that form misses the
# 1e-5 bar on this workload (~1.9e-5), so the headline may be another split.
VARIANT_NOTE = {
    "split_fp16": "deviation from the metric's split-BF16: the same 3-product hi+lo split "
                  "with FP16 planes of power-of-two-scaled operands (22 significant bits); "
                  "the literal split_bf16 is reported under variants",
    "split_bf16x3": "deviation from the metric's split-BF16: hi+mid+lo BF16, 6 products",
    "split_bf16": "the metric's split-BF16 form (hi+mid, 3 products)",
    "fp32": "plain FP32 MFMA",
}
ERR_BAR = 1e-5


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def workload(seed):
    from qflex_amd import circuits
    circ = circuits.random_circuit(7, 7, 32, seed)
    return circ, circuits.corner_inward_7x7(), circuits.grid_text(7, 7)


AUTOTUNE = True  # tile autotuner on (the warmup step does the timing of new products)
# Values of each bond cut contracted together (simulate's "slice_batch"):
# one value at a time.  (Batching the inner cut's 16 values measured no
# faster: its stacked operands of 2-8 GiB no longer stay in the 256 MB
# last-level cache between consecutive products, as the per-slice ones do.)
SLICE_BATCH = [1]


def options(circ, order, grid, precision):
    return {"circuit": circ.splitlines(), "ordering": order.splitlines(),
            "grid": grid.splitlines(), "initial_state": "0" * 49, "final_state": "",
            "memory_limit": "200GB", "precision": precision, "autotune": AUTOTUNE,
            "slice_batch": SLICE_BATCH}


def relerr(got, ref):
    import numpy as np
    g = np.array(got)
    r = np.array(ref)
    return float(np.linalg.norm(g - r) / np.linalg.norm(r))


def fidelity(got, ref):
    import numpy as np
    g = np.array(got, dtype=complex)
    r = np.array(ref, dtype=complex)
    return float(abs(np.vdot(r, g)) ** 2 / (np.vdot(r, r).real * np.vdot(g, g).real))


def measure_error(qflex, circ, grid, precision):
    """Norm-wise relative error and fidelity vs the committed complex128
    oracle amplitudes of the largest committed (sub-)job (same circuit; a
    sub-job restricts the ordering's cuts to a subset of their values)."""
    files = sorted(glob.glob(os.path.join(ROOT, "tests", "golden", "headline_7x7_d32_seed0_slices*.json")),
                   key=lambda p: json.load(open(p))["slices"])
    if not files:
        return None, 0, None
    from qflex_amd import circuits
    ref = json.load(open(files[-1]))
    order = circuits.corner_inward_7x7(tuple(ref["cut_values"]))
    r = qflex.evaluate(options(circ, order, grid, precision))
    got = {s: a for s, a in r}
    refs = [complex(re, im) for _, re, im in ref["amplitudes"]]
    vals = [got[s] for s, _, _ in ref["amplitudes"]]
    return relerr(vals, refs), ref["slices"], fidelity(vals, refs)


def time_variant(qflex, opts, steps, warmup):
    """Times `steps` steps with nothing but the work on the stream, then runs
    one more step with a HIP event pair around every launch (qfx_profile_*),
    which costs about 8% of a step, for the per-kernel-class figures."""
    for _ in range(warmup):
        qflex.evaluate(opts)
    qflex.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        r = qflex.evaluate(opts)
        log(f"[bench]   step {k + 1}/{steps} done")
    qflex.synchronize()
    dt = time.perf_counter() - t0
    stats = qflex.last_stats()
    qflex.profile_enable(True)
    qflex.profile_reset()
    qflex.synchronize()
    t1 = time.perf_counter()
    qflex.evaluate(opts)
    qflex.synchronize()
    dt_prof = time.perf_counter() - t1
    prof = qflex.profile_read()
    qflex.profile_enable(False)
    log("[bench]   profiled step done")
    return dt, len(r), prof, stats, dt_prof


def cpu_baseline(circ, grid):
    """The oracle in complex64 (host reorders + numpy/OpenBLAS cgemm, the
    reference's algorithm, terminal cuts by recursion) on a bounded sample:
    the pre-cut work and one of the 256 slices with all 64 outputs.  Each
    phase is timed by cut depth and scaled by how many more times the full
    job runs it (x1 before the first cut, x16 after it, x256 after the
    second), i.e. the full-job time is extrapolated."""
    import numpy as np
    from oracle.contraction import evaluate_circuit
    from qflex_amd import circuits
    try:
        from threadpoolctl import threadpool_info
        cores = max([i.get("num_threads", 1) for i in threadpool_info()
                     if i.get("internal_api") == "openblas"] or [os.cpu_count()])
    except Exception:
        cores = os.cpu_count()
    lines = circuits.corner_inward_7x7(([0], [0])).splitlines()
    # value count of each cut in the full job / in the sample, in order
    scale = [16 if len(l.split()) == 4 else 1 for l in lines if l.startswith("cut ")]
    phases = {}
    t0 = time.perf_counter()
    evaluate_circuit(circ, "\n".join(lines) + "\n", grid, "0" * 49, "", dtype=np.complex64,
                     phase_seconds=phases)
    wall = time.perf_counter() - t0
    total = phases.pop("build", 0.0)
    for depth, sec in phases.items():
        mult = 1
        for c in scale[:depth]:
            mult *= c
        total += sec * mult
    return {"value": 64.0 / total, "unit": "amplitudes/s", "cores": int(cores), "kind": "port",
            "sample": (f"oracle complex64 (numpy/OpenBLAS cgemm + host reorders, the reference "
                       f"algorithm incl. terminal-cut recursion): pre-cut work + 1 of 256 slices "
                       f"with all 64 outputs, {wall:.1f} s wall; per-depth times extrapolated to "
                       f"the full 256-slice job = {total:.0f} s"),
            "extrapolated": True}


def run_queue(qflex, states, precision):
    """The uneven queue of qflex_amd.jobqueue: each job (one circuit, `states`
    base bitstrings, 64 amplitudes each) run once untimed (tile autotuning),
    then timed; amplitudes/s per job and for the whole queue."""
    from qflex_amd import jobqueue
    jobs = jobqueue.queue(states=states)
    out = {"metric": "amplitudes/s, uneven queue (6x6 d24, 7x7 d32, 8x7 d24; 64-amplitude open "
                     "region per base bitstring)", "precision": precision, "jobs": []}
    total_amps, total_s = 0, 0.0
    for job in jobs:
        opts = dict(job["options"], precision=precision, autotune=AUTOTUNE)
        warm = dict(opts, final_states=opts["final_states"][:1])
        qflex.evaluate(warm)
        qflex.synchronize()
        t0 = time.perf_counter()
        r = qflex.evaluate(opts)
        qflex.synchronize()
        dt = time.perf_counter() - t0
        st = qflex.last_stats()
        total_amps += len(r)
        total_s += dt
        out["jobs"].append({"job": job["name"], "base_bitstrings": job["states"],
                            "amplitudes": len(r), "seconds": dt, "amplitudes_per_s": len(r) / dt,
                            "flops": st["flops"], "slices": st["slices"],
                            "slice_batch": opts["slice_batch"]})
        log(f"[bench] queue {job['name']}: {len(r)} amplitudes in {dt:.2f} s")
    out["value"] = total_amps / total_s
    out["unit"] = "amplitudes/s"
    out["seconds"] = total_s
    return out


def microbench(args):
    """Times qfx_cgemm_c64 on fixed shapes (random complex N(0,1) operands)."""
    import numpy as np
    # the library calibration runs first, in child processes, before this
    # process initializes the GPU
    library = []
    if args.library_fp16:
        for shape in args.microbench.split(";"):
            if not shape.startswith("P:"):
                m, n, k = (int(x) for x in shape.split(","))
                library.append(library_fp16_child(m, n, k, args.warmup, args.steps))
    from qflex_amd import capi
    ctx = capi.Context(0)
    rng = np.random.default_rng(0)
    out = []
    for shape in args.microbench.split(";"):
        if shape.startswith("P:"):
            # permutation: "P:d0,d1,...:p0,p1,..." (complex64 tensor)
            _, dims_s, perm_s = shape.split(":")
            dims = [int(x) for x in dims_s.split(",")]
            perm = [int(x) for x in perm_s.split(",")]
            nel = int(np.prod(dims))
            src, dst = ctx.alloc(nel * 8), ctx.alloc(nel * 8)
            for _ in range(args.warmup):
                ctx.reorder(src, dst, dims, perm)
            ctx.profile(True)
            ctx.profile_reset()
            for _ in range(args.steps):
                ctx.reorder(src, dst, dims, perm)
            prof = ctx.profile_read()
            prof = prof["permute"] if prof["permute"]["launches"] else prof["copy"]
            ctx.profile(False)
            ms = prof["ms"] / max(prof["launches"], 1)
            out.append({"permute": dims, "perm": perm, "avg_ms": ms,
                        "gbps": 16.0 * nel / ms / 1e6, "frac_of_hbm": 16.0 * nel / ms / 1e6 / HBM_GBPS})
            log(f"[bench] permute {dims} {perm}: {ms:.3f} ms, {out[-1]['gbps']:.0f} GB/s")
            src.free()
            dst.free()
            continue
        m, n, k = (int(x) for x in shape.split(","))
        a = (rng.standard_normal((m, k)) + 1j * rng.standard_normal((m, k))).astype(np.complex64)
        b = (rng.standard_normal((k, n)) + 1j * rng.standard_normal((k, n))).astype(np.complex64)
        A, B, C = ctx.upload(a), ctx.upload(b), ctx.alloc(m * n * 8)
        for prec in args.variants.split(","):
            p = capi.PRECISIONS[prec]
            for _ in range(args.warmup):
                ctx.cgemm(p, m, n, k, A, B, C)
            ctx.profile(True)
            ctx.profile_reset()
            for _ in range(args.steps):
                ctx.cgemm(p, m, n, k, A, B, C)
            prof = ctx.profile_read()
            ctx.profile(False)
            cls = "cgemm" if prof["cgemm"]["launches"] else ("gemv" if prof["gemv"]["launches"]
                                                             else "cgemm_small")
            g = prof[cls]
            ms = g["ms"] / max(g["launches"], 1)
            out.append({"m": m, "n": n, "k": k, "variant": prec, "class": cls, "avg_ms": ms,
                        "tflops": 8.0 * m * n * k / ms / 1e9,
                        "frac_of_peak": 8.0 * m * n * k / ms / 1e9 / PEAK_TFLOPS[prec]
                        if prec in PEAK_TFLOPS else None})
            log(f"[bench] {m}x{n}x{k} {prec}: {ms:.3f} ms, {out[-1]['tflops']:.1f} TF/s")
        for buf in (A, B, C):
            buf.free()
    ctx.close()
    return {"microbench": out + library}


def library_fp16_child(m, n, k, warmup, steps):
    """library_fp16 in a child process: torch brings its own HIP runtime,
    which does not share a process with the one libqfx.so uses."""
    import subprocess
    r = subprocess.run([sys.executable, os.path.abspath(__file__), "--library-shape", f"{m},{n},{k}",
                        "--steps", str(steps), "--warmup", str(warmup)],
                       capture_output=True, text=True, check=True)
    sys.stderr.write(r.stderr)
    return json.loads(r.stdout.strip().splitlines()[-1])


def library_fp16(m, n, k, warmup, steps):
    """Calibration: the vendor library's FP16 GEMM (torch.matmul, hipBLASLt)
    on the real product one split-FP16 term of an m x n x k complex product
    performs, A (m x 2k) times the 2 x 2 expansion of B (2k x 2n), FP32
    accumulation; the split form runs three such products."""
    import torch
    g = torch.Generator(device="cuda").manual_seed(0)
    a = torch.randn(m, 2 * k, device="cuda", dtype=torch.float16, generator=g)
    b = torch.randn(2 * k, 2 * n, device="cuda", dtype=torch.float16, generator=g)
    for _ in range(warmup):
        torch.matmul(a, b)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        torch.matmul(a, b)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    real_flops = 2.0 * m * (2 * k) * (2 * n)
    log(f"[bench] {m}x{n}x{k} library FP16 real product {m}x{2 * n}x{2 * k}: {ms:.3f} ms, "
        f"{real_flops / ms / 1e9:.1f} TF/s; three of them {3 * ms:.3f} ms")
    return {"m": m, "n": n, "k": k, "variant": "library_fp16_real_product",
            "real_shape": [m, 2 * n, 2 * k], "avg_ms": ms, "tflops_real": real_flops / ms / 1e9,
            "three_products_ms": 3 * ms}


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "cgemm_traffic_7x7_d32.json")
MFMA_GEMM_KERNELS = ("cgemm_split_big_kernel", "cgemm_split_chain_kernel", "cgemm_fp16x2_mfma_kernel",
                     "cgemm_f32_mfma_kernel", "cgemm_bf16_mfma_kernel")
GEMM_HELPER_KERNELS = ("amax_partial_kernel",)


def traffic_for(variant):
    """HBM bytes per MFMA cGEMM launch of this variant from the committed
    rocprofv3 summary (made by --summarize-pmc), or None."""
    try:
        t = json.load(open(TRAFFIC_FILE))
    except (OSError, ValueError):
        return None
    v = t.get("variants", {}).get(variant)
    return v["traffic_bytes_per_launch"] if v else None


def summarize_pmc(fetch_dir, write_dir, copy_mib):
    """HBM bytes per MFMA cGEMM launch from two rocprofv3 --pmc runs
    (FETCH_SIZE, WRITE_SIZE, KiB per dispatch), each scaled by what the same
    counter reads for qfx_copy_kernel moving copy_mib MiB."""
    import csv

    def load(d, counter):
        rows = []
        for f in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
            rows += [r for r in csv.DictReader(open(f)) if r["Counter_Name"] == counter]
        return rows

    out = {}
    for name, d, counter in (("fetch", fetch_dir, "FETCH_SIZE"), ("write", write_dir, "WRITE_SIZE")):
        rows = load(d, counter)
        copy = [float(r["Counter_Value"]) for r in rows if "qfx_copy_kernel" in r["Kernel_Name"]]
        gemm = [float(r["Counter_Value"]) for r in rows
                if any(k in r["Kernel_Name"] for k in MFMA_GEMM_KERNELS)]
        helper = [float(r["Counter_Value"]) for r in rows
                  if any(k in r["Kernel_Name"] for k in GEMM_HELPER_KERNELS)]
        if not copy or not gemm:
            raise SystemExit(f"{d}: no copy or no cGEMM dispatches with {counter}")
        factor = (copy_mib << 20) / (max(copy) * 1024.0)
        out[name] = {"counter": counter, "kib_copy": max(copy), "scale": factor,
                     "gemm_dispatches": len(gemm),
                     "bytes_per_launch": factor * 1024.0 * (sum(gemm) + sum(helper)) / len(gemm)}
    out["traffic_bytes_per_launch"] = out["fetch"]["bytes_per_launch"] + out["write"]["bytes_per_launch"]
    return out


def main():
    global AUTOTUNE, SLICE_BATCH
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--variants", default="fp32,split_fp16,split_bf16x3,split_bf16")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-autotune", action="store_true",
                    help="fixed tile rule instead of the tile autotuner")
    ap.add_argument("--slice-batch", default=",".join(str(x) for x in SLICE_BATCH),
                    help="values of each bond cut contracted together, one per cut in ordering "
                         "order ('1' = one slice at a time, the reference's loop)")
    ap.add_argument("--library-fp16", action="store_true",
                    help="with --microbench: also time the vendor library's FP16 GEMM (torch.matmul) "
                         "on the real product one split-FP16 term performs, for calibration")
    ap.add_argument("--library-shape", default="", help=argparse.SUPPRESS)
    ap.add_argument("--microbench", default="",
                    help="kernel mode: 'M,N,K' complex GEMM shapes separated by ';' (timed "
                         "with HIP events through the C-ABI, each --steps times per variant); "
                         "prints one JSON line and skips the workload")
    ap.add_argument("--no-error-check", action="store_true",
                    help="skip the full-job error measurement against the complex128 "
                         "amplitudes (for profiler runs that only need the timed steps)")
    ap.add_argument("--queue", type=int, default=0, metavar="STATES",
                    help="queue mode: the uneven queue of qflex_amd.jobqueue (6x6, 7x7, 8x7 "
                         "grids, a six-qubit open region, STATES base bitstrings each) in "
                         "split_fp16; prints one JSON line and skips the workload")
    ap.add_argument("--summarize-pmc", nargs=2, metavar=("FETCH_DIR", "WRITE_DIR"),
                    help="summarize two rocprofv3 --pmc output folders (FETCH_SIZE, WRITE_SIZE) "
                         "of a run with --calibrate-copy-mib into HBM bytes per MFMA cGEMM "
                         "launch; prints JSON and skips the workload")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="first run one device copy of this many MiB (qfx_copy_kernel), the "
                         "known-size reference for HBM byte counters")
    args = ap.parse_args()
    AUTOTUNE = not args.no_autotune
    SLICE_BATCH = [int(x) for x in args.slice_batch.split(",")]
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}))
        sys.exit(2)

    if args.summarize_pmc:
        print(json.dumps(summarize_pmc(args.summarize_pmc[0], args.summarize_pmc[1],
                                       args.calibrate_copy_mib or 1024), indent=1))
        return

    if args.library_shape:
        m, n, k = (int(x) for x in args.library_shape.split(","))
        print(json.dumps(library_fp16(m, n, k, args.warmup, args.steps)))
        return

    from qflex_amd import qflex

    if args.calibrate_copy_mib:
        from qflex_amd import capi
        ctx = capi.Context(0)
        n = args.calibrate_copy_mib << 20
        a, b = ctx.alloc(n), ctx.alloc(n)
        ctx.copy(b, a, n)
        ctx.sync()
        log(f"[bench] calibration copy: {n} bytes read, {n} bytes written")
        ctx.close()

    if args.microbench:
        print(json.dumps(microbench(args)))
        return

    if args.queue:
        print(json.dumps(run_queue(qflex, args.queue, args.variants.split(",")[0]
                                   if args.variants != ap.get_default("variants") else "split_fp16")))
        return

    circ, order, grid = workload(args.seed)
    variants = {}
    for prec in args.variants.split(","):
        log(f"[bench] variant {prec}: warmup {args.warmup}, steps {args.steps}")
        dt, n_amps, prof, stats, dt_prof = time_variant(qflex, options(circ, order, grid, prec),
                                                        args.steps, args.warmup)
        err, err_slices, fid = ((None, 0, None) if args.no_error_check
                                else measure_error(qflex, circ, grid, prec))
        g = prof["cgemm"]
        p = prof["permute"]
        variants[prec] = {
            "amplitudes_per_s": n_amps * args.steps / dt,
            "ms_per_step": 1e3 * dt / args.steps,
            "relerr_vs_complex128": err,
            "relerr_slices": err_slices,
            "fidelity": fid,
            "profiled_ms_per_step": 1e3 * dt_prof,
            # per-class HIP-event totals of the profiled step
            "cgemm": {"launches": g["launches"], "ms": g["ms"],
                      "tflops": g["flops"] / max(g["ms"], 1e-9) / 1e9},
            "permute": {"launches": p["launches"], "ms": p["ms"],
                        "gbps": p["bytes"] / max(p["ms"], 1e-9) / 1e6},
            "class_ms": {c: round(v["ms"], 3) for c, v in prof.items() if v["launches"]},
            "flops_per_step": stats["flops"], "slices": stats["slices"],
            # ContractGrid alone (device-synchronized wall time of the last
            # timed step); the rest of a step is the network build and flatten
            "contraction_ms": 1e3 * stats["seconds"],
            "_prof": prof,
        }
        log(f"[bench] {prec}: {variants[prec]['amplitudes_per_s']:.2f} amp/s, "
            f"relerr {err}, cgemm {variants[prec]['cgemm']['tflops']:.1f} TF/s")

    ok = {k: v for k, v in variants.items()
          if v["relerr_vs_complex128"] is not None and v["relerr_vs_complex128"] <= ERR_BAR}
    # the headline is the fastest variant that meets the error bar; when none
    # does, value is null (the fastest variant still fills the rest)
    pick = max(ok or variants, key=lambda k: variants[k]["amplitudes_per_s"])
    best = variants[pick]
    if not ok:
        log(f"[bench] no variant meets the {ERR_BAR} error bar: value is null")
    # reorders: the variant that moves the most bytes through them (fp32,
    # which reorders every operand; the split variants read operands in their
    # stored order and keep only a few small reorders)
    perm_variant = max(variants, key=lambda k: variants[k]["_prof"]["permute"]["bytes"])
    p = variants[perm_variant]["_prof"]["permute"]
    p_ms = max(p["ms"], 1e-9)
    prof = best.pop("_prof")
    for v in variants.values():
        v.pop("_prof", None)
    g = prof["cgemm"]
    achieved = g["flops"] / g["ms"] / 1e9
    out = {
        "metric": METRIC,
        "value": best["amplitudes_per_s"] if ok else None,
        "unit": "amplitudes/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": best["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": pick,
        "data": (f"synthetic: seeded 7x7 random circuit (seed {args.seed}) of the reference's "
                 "8-pattern family, depth 1+32+1, bond dimension 16"),
        "config": {"workload": "7x7 grid, depth 1+32+1, bond 16, corner-inward ordering, "
                               "2 cuts = 256 slices, 6-qubit open region = 64 amplitudes per step",
                   "variant": pick, "precision_bar": ERR_BAR, "autotune": AUTOTUNE,
                   "slice_batch": SLICE_BATCH,
                   "passes_bar": pick in ok,
                   "variant_note": VARIANT_NOTE.get(pick, "")},
        "variants": variants,
        "roofline": {"bound": "mfma", "kernel": "cgemm_" + pick,
                     "achieved": achieved, "peak": PEAK_TFLOPS[pick], "unit": "TFLOP/s",
                     "frac": achieved / PEAK_TFLOPS[pick], "traffic": traffic_for(pick),
                     "algorithmic_bytes_per_launch": g["bytes"] / max(g["launches"], 1),
                     "avg_launch_ms": g["ms"] / max(g["launches"], 1),
                     "flops_per_launch": g["flops"] / max(g["launches"], 1)},
        "roofline_permute": {"bound": "hbm", "variant": perm_variant,
                             "achieved": p["bytes"] / p_ms / 1e6,
                             "peak": HBM_GBPS, "unit": "GB/s",
                             "frac": p["bytes"] / p_ms / 1e6 / HBM_GBPS,
                             "avg_launch_ms": p["ms"] / max(p["launches"], 1),
                             "launches_per_step": p["launches"]},
    }
    if not args.no_cpu_baseline:
        log("[bench] cpu baseline (bounded sample)")
        out["cpu_baseline"] = cpu_baseline(circ, grid)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
