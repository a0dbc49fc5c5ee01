#!/usr/bin/env python3
"""Benchmark: batched mixed-precision randomized TT rounding on one MI355X.

Workload (default): 256 TT vectors, each a linear combination of 8 seeded
N(0,1) TTs with d=40, n=4 and ranks drawn uniformly from 16..256, coefficients
0.5^j.  One step rounds the whole batch: Gaussian TT sketch (FP32 contractions,
block rank 40 = d, sketch ranks from the rank heuristic), FP64 Householder QR
and projection sweep, then the FP64 SVD cleanup to tol 1e-6.  Inputs are
generated on the device and resident in HBM before timing starts.

Prints one JSON line.  Besides the metric it reports the FP64 deterministic
SVD rounding of the same TTs (on a subset, because 256 materialised sums do not
fit in HBM), the relative error of a sample item against the FP64 oracle, the
dominant kernel's achieved rate against its roofline, and the oracle's CPU time
on a bounded sample.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tensor-train-julia_amd"))

import numpy as np  # noqa: E402

PEAKS = {"gemm_f64": ("mfma", 78.6e12), "gemm_f32": ("mfma", 157.3e12), "gemm_bf16": ("mfma", 2516e12),
         "qr_panel": ("hbm", 8.0e12), "jacobi_svd": ("mfma", 78.6e12)}
# device kernel behind each profiled class (for matching rocprofv3 summaries)
KERNEL_SYMBOL = {"gemm_f64": "gemm_scalar_kernel<double, double, double, double>",
                 "gemm_f32": "gemm_scalar_kernel<double, float, float, float>",
                 "gemm_bf16": "gemm_bf16_kernel", "qr_panel": "qr_panel_reg_kernel",
                 "jacobi_svd": "jacobi_round_kernel"}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def measured_traffic(args, kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summary of this workload (FETCH_SIZE and WRITE_SIZE runs, each scaled by what
    the counter reads for a plain copy of known size), or None."""
    path = os.path.join(ROOT, "profiles", "c3_batched_rounding_hbm_traffic.json")
    try:
        with open(path) as f:
            rec = json.load(f)
    except (OSError, ValueError):
        return None
    cfg = rec.get("config", {})
    if any(cfg.get(k) != getattr(args, k) for k in ("items", "d", "n", "terms", "sketch_prec")):
        return None
    k = rec.get("kernels", {}).get(kernel)
    return None if k is None else k.get("hbm_bytes_per_launch")


def term_ranks(dims, r):
    import ttround as T
    return T.r_and_d_to_rks([r] * (len(dims) + 1), dims, rmax=r)


def build_workload(args):
    import ttround as T
    rng = np.random.default_rng(3)
    dims = (args.n,) * args.d
    items = []
    for b in range(args.items):
        terms, alphas, rk = [], [], []
        for j in range(1, args.terms + 1):
            r = int(rng.integers(args.rmin, args.rmax + 1))
            t = T.DeviceTT.randn(dims, r, seed=1000 * 3 + j + 100000 * b)
            terms.append(t)
            alphas.append(0.5 ** j)
            rk.append(t.rks)
        rks = list(rk[0])
        for q in rk[1:]:
            rks = [max(a, c) for a, c in zip(rks, q)]
        ovs = int(math.ceil(0.5 * max(rks)))
        for k in range(1, args.d):
            rks[k] += ovs
        items.append((alphas, terms, rks))
    return dims, items


def run_step(items, args):
    import ttround as T
    outs = T.round_rand_batch(items, orthogonal=False, seed=1234, block_rks=None, sketch_prec=args.sketch_prec)
    if args.cleanup_tol > 0:
        clean = T.round_svd_batch(outs, tol=args.cleanup_tol)
        for o in outs:
            o.free()
        outs = clean
    return outs


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--items", type=int, default=256)
    ap.add_argument("--d", type=int, default=40)
    ap.add_argument("--n", type=int, default=4)
    ap.add_argument("--terms", type=int, default=8)
    ap.add_argument("--rmin", type=int, default=16)
    ap.add_argument("--rmax", type=int, default=256)
    ap.add_argument("--sketch-prec", default="f32")
    ap.add_argument("--cleanup-tol", type=float, default=1e-6)
    ap.add_argument("--det-items", type=int, default=48,
                    help="items for the FP64 deterministic comparison (0: skip); their materialised sums must fit in HBM")
    ap.add_argument("--cpu-items", type=int, default=1, help="items for the oracle CPU baseline / parity (0: skip)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="first copy BYTES of device memory 3 times (a plain copy kernel of known size, "
                         "the reference point for hardware traffic counters)")
    ap.add_argument("--no-events", action="store_true",
                    help="do not record the library's per-launch hipEvents (for runs under an external profiler)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    if args.copy_calibration > 0:
        import torch
        n = args.copy_calibration // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"[bench] copied {args.copy_calibration} bytes x3")
        del src, dst

    import ttround as T
    h = T.default_handle()
    log(f"[bench] building workload: {args.items} items x {args.terms} terms, d={args.d}, n={args.n}")
    t0 = time.time()
    dims, items = build_workload(args)
    log(f"[bench] workload ready in {time.time() - t0:.1f}s")

    for w in range(args.warmup):
        outs = run_step(items, args)
        log(f"[bench] warmup {w} done; out ranks (item 0, max) = {max(outs[0].rks)}")
        for o in outs:
            o.free()

    h.profile(not args.no_events)
    h.profile_reset()
    t0 = time.perf_counter()
    for s in range(args.steps):
        outs = run_step(items, args)
        if s < args.steps - 1:
            for o in outs:
                o.free()
        log(f"[bench] step {s} done")
    dt = time.perf_counter() - t0
    h.profile(False)
    stats = h.profile_stats()
    ms_per_step = 1e3 * dt / args.steps
    value = args.items * args.steps / dt

    sweeps = stats.pop("jacobi_svd:sweeps", None)
    jacobi = ({"svds": sweeps["launches"], "mean_sweeps": round(sweeps["flops"] / max(sweeps["launches"], 1), 2),
               "max_sweeps": sweeps["bytes"]} if sweeps else None)
    if jacobi:
        for ph in ("visits", "ticks_gram", "ticks_inner", "ticks_apply"):
            st = stats.pop("jacobi_svd:" + ph, None)
            if st:
                jacobi[ph] = st["flops"]
    sites = {k[5:]: {"ms_per_step": round(v["ms"] / args.steps, 3), "launches_per_step": v["launches"] // args.steps,
                     "tflops": round(v["flops"] / max(v["ms"], 1e-9) / 1e9, 3)}
             for k, v in stats.items() if k.startswith("site:")}
    kern = {k: v for k, v in stats.items() if not k.startswith("phase:") and not k.startswith("site:")}
    phases = {k[6:]: round(v["ms"] / args.steps, 3) for k, v in stats.items() if k.startswith("phase:")}
    dom = max(kern, key=lambda k: kern[k]["ms"]) if kern else None
    roof = None
    if dom:
        st = kern[dom]
        bound, peak = PEAKS.get(dom, ("hbm", 8.0e12))
        avg_s = st["ms"] / st["launches"] / 1e3
        if bound == "mfma":
            per_launch = st["flops"] / st["launches"]
            ach, unit, pk = per_launch / avg_s / 1e12, "TFLOP/s", peak / 1e12
        else:
            per_launch = st["bytes"] / st["launches"]
            ach, unit, pk = per_launch / avg_s / 1e9, "GB/s", peak / 1e9
        roof = {"kernel": dom, "device_kernel": KERNEL_SYMBOL.get(dom), "bound": bound, "achieved": round(ach, 3),
                "peak": pk, "unit": unit, "frac": round(ach / pk, 5), "traffic": measured_traffic(args, dom),
                "algorithmic_bytes_per_launch": round(st["bytes"] / st["launches"]),
                "avg_launch_us": round(avg_s * 1e6, 2),
                "launches_per_step": st["launches"] / args.steps,
                "share_of_kernel_time": round(st["ms"] / sum(v["ms"] for v in kern.values()), 4)}

    kernels = {k: {"ms_per_step": round(v["ms"] / args.steps, 3), "launches_per_step": v["launches"] // args.steps,
                   "tflops": round(v["flops"] / max(v["ms"], 1e-9) / 1e9, 3)} for k, v in kern.items()}

    # ------------------------------------------------------------ parity sample + CPU baseline
    rel_err = None
    cpu = None
    if args.cpu_items > 0:
        import oracle as O
        errs, tcpu = [], 0.0
        for i in range(args.cpu_items):
            al, terms, rks = items[i]
            ys = [O.TT([t.core(k) for k in range(t.d)], t.dims, t.rks, t.ot) for t in terms]
            blocks = {}

            def prov(k, rks=rks):
                if k not in blocks:
                    blocks[k] = T.sketch_blocks(dims, rks, k, orthogonal=False, seed=1234, sketch_prec=args.sketch_prec)
                return blocks[k]
            for k in range(args.d):
                prov(k)
            tc = time.perf_counter()
            ref = O.ttrand_rounding_lincomb(al, ys, rks, prov, False)
            if args.cleanup_tol > 0:
                ref = O.tt_rounding(ref, tol=args.cleanup_tol)
            tcpu += time.perf_counter() - tc
            got = outs[i]
            g = O.TT([got.core(k) for k in range(got.d)], got.dims, got.rks, got.ot)
            errs.append(O.rel_err(ref, g))
            log(f"[bench] item {i}: rel err vs FP64 oracle {errs[-1]:.3e} (oracle {tcpu:.1f}s)")
        rel_err = max(errs)
        cpu = {"value": round(args.cpu_items / tcpu, 4), "unit": "rounded TTs/s", "cores": os.cpu_count(),
               "kind": "port",
               "sample": f"{args.cpu_items} item(s) of the same batch, FP64 numpy/OpenBLAS oracle "
                         f"(randomized rounding + SVD cleanup), BLAS threads = {os.environ.get('OPENBLAS_NUM_THREADS', 'default')}"}
    for o in outs:
        o.free()

    # ------------------------------------------------------------ FP64 deterministic comparison
    det = None
    if args.det_items > 0:
        sub = items[: args.det_items]
        sums = [T.lincomb_device(al, terms) for al, terms, _ in sub]
        # the materialised sums (1.28 GB each at these ranks) need the room of the terms
        for _, terms, _ in items:
            for t in terms:
                t.free()
        items = None
        r = T.round_svd_batch(sums, tol=args.cleanup_tol if args.cleanup_tol > 0 else 1e-12)
        for o in r:
            o.free()
        tdet = time.perf_counter()
        r = T.round_svd_batch(sums, tol=args.cleanup_tol if args.cleanup_tol > 0 else 1e-12)
        tdet = time.perf_counter() - tdet
        det = {"items": args.det_items, "rounded_tts_per_s": round(args.det_items / tdet, 4),
               "ms": round(tdet * 1e3, 2), "out_rank_max": max(max(o.rks) for o in r)}
        log(f"[bench] FP64 deterministic on {args.det_items} items: {tdet:.2f}s")
        for o in r + sums:
            o.free()

    res = {
        "metric": "rounded TTs/sec, mixed-prec batched vs FP64 plain, at stated rel. error; HBM GB/s",
        "value": round(value, 4),
        "unit": "rounded TTs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": f"{args.sketch_prec}-sketch+f64",
        "data": "synthetic (seeded N(0,1) TT cores drawn on the device)",
        "config": {"workload": f"batched randomized rounding: {args.items} TTs, each sum of {args.terms} N(0,1) TTs "
                               f"d={args.d} n={args.n} r~U{{{args.rmin}..{args.rmax}}} coeff 0.5^j; "
                               f"{args.sketch_prec} Gaussian sketch, block_rks={args.d}, heuristic ranks; "
                               f"FP64 QR/projection; SVD cleanup tol={args.cleanup_tol}",
                   "items": args.items, "d": args.d, "n": args.n, "terms": args.terms,
                   "sketch_prec": args.sketch_prec, "cleanup_tol": args.cleanup_tol},
        "rel_err_vs_oracle": rel_err,
        "fp64_deterministic": det,
        "speedup_vs_fp64_deterministic": (round(value / det["rounded_tts_per_s"], 3) if det else None),
        "phases_ms_per_step": phases,
        "jacobi": jacobi,
        "kernels": kernels,
        "gemm_sites": sites,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
