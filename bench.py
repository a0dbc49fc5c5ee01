#!/usr/bin/env python3
"""Benchmark: batched mixed-precision randomized TT rounding on one MI355X.

Default workload (`--workload c3`): 256 TT vectors, each a linear combination
of 8 seeded N(0,1) TTs with d=40, n=4 and ranks drawn uniformly from 16..256,
coefficients 0.5^j.  One step rounds the whole batch: Gaussian TT sketch (FP32
contractions, block rank 40 = d, sketch ranks from the rank heuristic), FP64
Householder QR and projection sweep, then the FP64 SVD cleanup to tol 1e-6.
Inputs are generated on the device and resident in HBM before timing starts.

Other workloads (one JSON line each, same fields):
  c1  6 N(0,1) TTs (d=32, n=4, r=100), weights 2^-k; FP64 Gaussian TT sketch of
      rank 144 (block rank 144), cleanup to rmax 128; one TT.
  c2  quantics: 8 N(0,1) TTs (d=64, n=2, r=128), weights 10^(-k/2); BF16 sketch
      of rank 272 (block rank 272), cleanup to rmax 256; one TT.
  c4  1D Hubbard MPO (L=32, local dim 4, t=1, U=4, bond dim 6) applied to 64
      N(0,1) TTs of rank 128 without forming the products; BF16 sketch of rank
      144 (block rank 144), cleanup to rmax 128.
  dot    one Gram-Schmidt pass: 32 inner products of rank-128 TTs (d=40, n=4)
         with one rank-128 TT, FP64.
  dotop  16 Hubbard energies <psi|H|psi> (L=32, rank-128 states), FP64.

Prints one JSON line.  Besides the metric it reports the FP64 deterministic SVD
rounding of the same inputs (materialised sums / products, on a subset where
they do not fit in HBM), the relative error of a sample item against the FP64
oracle run on the same sketch blocks, the dominant kernel's achieved rate
against its roofline, and the oracle's CPU time on a bounded sample.
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
         "qr_panel": ("hbm", 8.0e12), "qr_block": ("hbm", 8.0e12), "jacobi_svd": ("mfma", 78.6e12),
         "op_apply": ("hbm", 8.0e12)}
# device kernel behind each profiled class (for matching rocprofv3 summaries)
KERNEL_SYMBOL = {"gemm_f64": "gemm_scalar_kernel<double, double, double, double>",
                 "gemm_f32": "gemm_scalar_kernel<double, float, float, float>",
                 "gemm_bf16": "gemm_bf16_kernel", "qr_panel": "qr_panel_kernel", "qr_block": "qr_block_kernel",
                 "jacobi_svd": "jacobi_round_kernel", "op_apply": "op_apply_kernel"}

WORKLOADS = {
    "c3": dict(items=256, d=40, n=4, terms=8, rmin=16, rmax=256, weight="0.5^j", seed_base=3000, sketch="f32",
               accum="f64", gemm_dominant=True,
               block_rks=None, rks=None, cleanup=dict(tol=1e-6), op=False, det_items=48,
               text="batched randomized rounding: 256 TTs, each sum of 8 N(0,1) TTs d=40 n=4 r~U{16..256} "
                    "coeff 0.5^j; FP32 Gaussian sketch, block_rks=40, heuristic ranks; FP64 QR/projection; "
                    "SVD cleanup tol=1e-6"),
    "c1": dict(items=1, d=32, n=4, terms=6, rmin=100, rmax=100, weight="2^-k", seed_base=1000, sketch="f64",
               accum="sketch",
               block_rks=144, rks=144, cleanup=dict(rmax=128), op=False, det_items=1,
               text="randomized rounding of sum_k 2^-k y_k, 6 N(0,1) TTs d=32 n=4 r=100 (formal rank 600); "
                    "FP64 Gaussian TT sketch rank 144 (block_rks=144); SVD cleanup rmax=128"),
    "c2": dict(items=1, d=64, n=2, terms=8, rmin=128, rmax=128, weight="10^(-k/2)", seed_base=2000, sketch="bf16",
               accum="sketch",
               block_rks=272, rks=272, cleanup=dict(rmax=256), op=False, det_items=1,
               text="quantics TT: sum_k 10^(-k/2) y_k, 8 N(0,1) TTs d=64 n=2 r=128 (formal rank 1024); "
                    "sketch rank 272 (block_rks=272); SVD cleanup rmax=256"),
    "c4": dict(items=64, d=32, n=4, terms=1, rmin=128, rmax=128, weight="1", seed_base=4000, sketch="bf16",
               accum="sketch",
               block_rks=144, rks=144, cleanup=dict(rmax=128), op=True, det_items=16,
               text="1D Hubbard MPO (L=32, t=1, U=4, bond dim 6) times 64 N(0,1) TTs d=32 n=4 r=128, "
                    "implicit product; sketch rank 144 (block_rks=144); SVD cleanup rmax=128"),
}

# inner-product workloads: one step evaluates the whole batch
INNER = {
    "dot": dict(pairs=32, d=40, n=4, r=128, seed_base=5000,
                text="Gram-Schmidt pass: <V_i, w> for a basis of 32 N(0,1) TTs d=40 n=4 r=128 against one TT "
                     "r=128 (dot, FP64)"),
    "dotop": dict(pairs=16, d=32, n=4, r=128, seed_base=6000,
                  text="energies <psi_i|H|psi_i>: 1D Hubbard MPO (L=32, t=1, U=4, bond dim 6), 16 N(0,1) TTs "
                       "d=32 n=4 r=128 (dot_operator, FP64, H psi never formed)"),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def weight(wl, j):
    return {"0.5^j": 0.5 ** j, "2^-k": 2.0 ** -j, "10^(-k/2)": 10.0 ** (-j / 2), "1": 1.0}[wl["weight"]]


def measured_traffic(args, kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summary of this workload (FETCH_SIZE and WRITE_SIZE runs, each scaled by what
    the counter reads for a plain copy of known size), or None."""
    path = os.path.join(ROOT, "profiles", f"{args.workload}_batched_rounding_hbm_traffic.json")
    try:
        with open(path) as f:
            rec = json.load(f)
    except (OSError, ValueError):
        return None
    cfg = rec.get("config", {})
    if any(cfg.get(k) != getattr(args, k) for k in ("items", "sketch_prec", "sketch_accum")):
        return None
    k = rec.get("kernels", {}).get(kernel)
    return None if k is None else k.get("hbm_bytes_per_launch")


def build_workload(wl, args):
    """Device-resident inputs: per item (alphas, [DeviceTT], target ranks)."""
    import ttround as T
    rng = np.random.default_rng(3)
    d, n = wl["d"], wl["n"]
    dims = (n,) * d
    op = T.DeviceTTO(T.hubbard_1d_mpo(d, 1.0, 4.0)) if wl["op"] else None
    items, specs = [], []
    for b in range(args.items):
        terms, alphas, rk, seeds = [], [], [], []
        for j in range(1, wl["terms"] + 1):
            r = int(rng.integers(wl["rmin"], wl["rmax"] + 1))
            seeds.append(wl["seed_base"] + j + 100000 * b)
            t = T.DeviceTT.randn(dims, r, seed=seeds[-1])
            terms.append(t)
            alphas.append(weight(wl, j))
            rk.append(t.rks)
        specs.append((seeds, [list(r) for r in rk]))
        if wl["rks"] is None:
            rks = list(rk[0])
            for q in rk[1:]:
                rks = [max(a, c) for a, c in zip(rks, q)]
            ovs = int(math.ceil(0.5 * max(rks)))
            for k in range(1, d):
                rks[k] += ovs
        else:
            rks = [1] + [wl["rks"]] * (d - 1) + [1]
        items.append((alphas, terms, rks))
    return dims, op, items, specs


def run_step(items, op, wl, args, keep_raw=False):
    """One pass over the batch; returns (rounded TTs, randomized-stage TTs or None)."""
    import ttround as T
    batch = [(al, op, terms, rks) if op is not None else (al, terms, rks) for al, terms, rks in items]
    raw = T.round_rand_batch(batch, orthogonal=False, seed=1234, block_rks=wl["block_rks"],
                             sketch_prec=args.sketch_prec, sketch_accum=args.sketch_accum)
    if args.no_cleanup:
        return raw, None
    clean = T.round_svd_batch(raw, **wl["cleanup"])
    if not keep_raw:
        for o in raw:
            o.free()
        raw = None
    return clean, raw


def _oracle_item(task):
    """One item through the FP64 oracle in a worker process (one BLAS thread): the
    terms are regenerated from their seeds (the device's N(0,1) stream, bit for
    bit), the device's sketch blocks are given, and the errors of the device's
    results are returned with the oracle's time."""
    import oracle as O
    from oracle import philox as P
    dims, seeds, term_rks, al, rks, blocks, block_rks, hub, cleanup, got, got_raw = task
    ys = [O.TT(P.rand_tt_cores(dims, r, s), dims, list(r), [0] * len(dims)) for s, r in zip(seeds, term_rks)]
    H = O.hubbard_1d_mpo(len(dims), 1.0, 4.0) if hub else None
    tc = time.perf_counter()
    if H is not None:
        ref = O.ttrand_rounding_op(al, H, ys, rks, blocks.__getitem__, False, block_rks)
    else:
        ref = O.ttrand_rounding_lincomb(al, ys, rks, blocks.__getitem__, False, block_rks)
    e_raw = O.rel_err(ref, got_raw) if got_raw is not None else None
    if cleanup is not None:
        ref = O.tt_rounding(ref, **cleanup)
    t = time.perf_counter() - tc
    return O.rel_err(ref, got), e_raw, t


def parity_and_cpu(items, specs, op, outs, raw, dims, wl, args):
    """The first cpu_items items of the batch through the FP64 oracle on the
    device's own sketch blocks (and the same cleanup), which is also the CPU
    baseline: the items run batch-parallel in worker processes with one BLAS
    thread each (the oracle is not BLAS-bound: one C3 item takes about as long on
    one thread as on eight), and item 0 runs again in this process with every
    BLAS thread (the serial loop of the reference's benchmarks).  Returns the
    largest relative errors (final result, randomized stage) and the baseline."""
    import multiprocessing as mp

    import oracle as O
    import ttround as T

    def host(dx):
        return O.TT([dx.core(k) for k in range(dx.d)], dx.dims, dx.rks, dx.ot)
    cleanup = None if args.no_cleanup else wl["cleanup"]
    tasks = []
    for i in range(args.cpu_items):
        al, terms, rks = items[i]
        blocks = {k: T.sketch_blocks(dims, rks, k, orthogonal=False, seed=1234, block_rks=wl["block_rks"],
                                     sketch_prec=args.sketch_prec) for k in range(wl["d"])}
        seeds, trks = specs[i]
        tasks.append((dims, seeds, trks, al, rks, blocks, wl["block_rks"], op is not None, cleanup, host(outs[i]),
                      host(raw[i]) if raw is not None else None))
    nproc = max(1, min(args.cpu_items, args.cpu_procs))
    os.environ["OPENBLAS_NUM_THREADS"] = "1"  # the workers' BLAS (this process keeps its own)
    walls = []
    with mp.get_context("spawn").Pool(nproc) as pool:
        t0 = time.perf_counter()
        res = pool.map(_oracle_item, tasks, chunksize=1)
        walls.append(time.perf_counter() - t0)
        # more trials of the same loop (at most 5 in all) while they keep the leg near 30 s
        while len(walls) < 5 and sum(walls) + walls[0] <= 30.0:
            t0 = time.perf_counter()
            pool.map(_oracle_item, tasks, chunksize=1)
            walls.append(time.perf_counter() - t0)
    wall = sorted(walls)[len(walls) // 2]  # median trial
    for i, (e, er, t) in enumerate(res):
        log(f"[bench] item {i}: rel err vs FP64 oracle {e:.3e}"
            f"{f' (randomized stage {er:.3e})' if er is not None else ''} (oracle {t:.1f}s, 1 thread)")
    errs = [r[0] for r in res]
    errs_raw = [r[1] for r in res if r[1] is not None]
    item_s = sorted(r[2] for r in res)
    med = item_s[len(item_s) // 2]
    # serial: item 0 in this process with all BLAS threads
    try:
        from threadpoolctl import threadpool_info
        threads = str(max(i["num_threads"] for i in threadpool_info() if i.get("internal_api") == "openblas"))
    except (ImportError, ValueError):
        threads = os.environ.get("OMP_NUM_THREADS", str(os.cpu_count()))
    tc = time.perf_counter()
    _oracle_item(tasks[0])
    t_serial = time.perf_counter() - tc
    log(f"[bench] batch-parallel oracle: {args.cpu_items} items on {nproc} processes in {wall:.1f}s "
        f"(median of {len(walls)} trial(s)) "
        f"(median item {med:.1f}s); serial item 0 with all BLAS threads {t_serial:.1f}s")
    par = args.cpu_items / wall
    ser = 1.0 / t_serial
    cpu = {"value": round(max(par, ser), 4), "unit": "rounded TTs/s", "cores": nproc if par >= ser else None,
           "kind": "port",
           "sample": f"{args.cpu_items} item(s) of the same batch through the FP64 numpy/OpenBLAS oracle "
                     f"(randomized rounding{'' if args.no_cleanup else ' + SVD cleanup'}) on the device's sketch "
                     f"blocks; value = the faster of the batch-parallel loop ({nproc} processes x 1 BLAS thread, median "
                     f"of {len(walls)} trial(s)) "
                     f"and the serial loop (all BLAS threads)",
           "batch_parallel": {"items": args.cpu_items, "processes": nproc, "blas_threads": 1, "wall_s": round(wall, 2),
                              "trials": len(walls), "trial_wall_s": [round(w, 2) for w in walls],
                              "value": round(par, 4), "median_item_s": round(med, 2),
                              "item_s": [round(v, 2) for v in item_s]},
           "serial": {"items": 1, "blas_threads": threads, "item_s": round(t_serial, 2), "value": round(ser, 4)}}
    if cpu["cores"] is None:
        cpu["cores"] = int(threads) if threads.isdigit() else os.cpu_count()
    return max(errs), (max(errs_raw) if errs_raw else None), cpu, errs


def input_errors(items, op, outs, raw, n):
    """Relative Frobenius error of the first n results against the input they round
    (sum_j alpha_j y_j, or H y for the operator workload), for the final result and
    the randomized stage, from inner products on the device (one batched ttr_dot):
    ||r - x||^2 = <r, r> - 2 sum_j alpha_j <r, y_j> + sum_jl alpha_j alpha_l <y_j, y_l>.
    The expansion resolves errors down to ~1e-7 relative; the errors against the
    input here are set by the truncation (1e-2 and above)."""
    import ttround as T
    n = min(n, len(items))
    xs, ys, made = [], [], []
    for i in range(n):
        al, terms, _ = items[i]
        base = [T.mpo_apply_device(op, terms[0])] if op is not None else list(terms)
        if op is not None:
            made += base
        for j in range(len(base)):
            for l in range(j, len(base)):
                xs.append(base[j]); ys.append(base[l])
        for res in (outs, raw):
            if res is not None:
                xs += [res[i]] * (len(base) + 1)
                ys += base + [res[i]]
    d = T.dot_batch(xs, ys)
    for m in made:
        m.free()
    fin, rnd, p = [], [], 0
    for i in range(n):
        al = [1.0] if op is not None else items[i][0]
        nb = len(al)
        G = np.zeros((nb, nb))
        for j in range(nb):
            for l in range(j, nb):
                G[j, l] = G[l, j] = d[p]
                p += 1
        xx = float(np.dot(al, G @ np.asarray(al)))
        for res, dst in ((outs, fin), (raw, rnd)):
            if res is None:
                continue
            rx = float(np.dot(al, d[p:p + nb]))
            rr = float(d[p + nb])
            p += nb + 1
            dst.append(math.sqrt(max(rr - 2.0 * rx + xx, 0.0) / xx))
    log(f"[bench] error against the input on {n} item(s): final {max(fin):.3e}"
        + (f", randomized stage {max(rnd):.3e}" if rnd else ""))
    return (max(fin) if fin else None), (max(rnd) if rnd else None)


def deterministic(items, op, wl, args):
    """Plain FP64 SVD rounding (tt_rounding with the cleanup's tol/rmax) of the
    materialised sums / products of the first det_items items."""
    import ttround as T
    nd = min(args.det_items, len(items))
    sub = items[:nd]
    if op is not None:
        mats = [T.mpo_apply_device(op, terms[0]) for _, terms, _ in sub]
    else:
        mats = [T.lincomb_device(al, terms) for al, terms, _ in sub]
    # the materialised inputs need the room of the terms
    for _, terms, _ in items:
        for t in terms:
            t.free()
    # the batch's workspace arena and freed storage go back to the runtime first
    T.default_handle().trim()
    r = T.round_svd_batch(mats, **wl["cleanup"])
    for o in r:
        o.free()
    tdet = time.perf_counter()
    r = T.round_svd_batch(mats, **wl["cleanup"])
    tdet = time.perf_counter() - tdet
    det = {"items": nd, "rounded_tts_per_s": round(nd / tdet, 4), "ms": round(tdet * 1e3, 2),
           "in_rank_max": max(max(m.rks) for m in mats), "out_rank_max": max(max(o.rks) for o in r)}
    log(f"[bench] FP64 deterministic on {nd} items: {tdet:.2f}s")
    for o in r + mats:
        o.free()
    return det


def summarize(stats, args):
    cert = stats.pop("svd_certify:bonds", None)
    svdb = stats.pop("svd_sweep:svd_bonds", None)
    cleanup = None
    if cert or svdb:
        cleanup = {"bonds_certified_per_step": round((cert or {}).get("flops", 0.0) / args.steps, 1),
                   "bonds_tried_per_step": round((cert or {}).get("launches", 0) / args.steps, 1),
                   "svd_bonds_per_step": round((svdb or {}).get("launches", 0) / args.steps, 1),
                   "svd_bonds_truncating_per_step": round((svdb or {}).get("flops", 0.0) / args.steps, 1)}
    sweeps = stats.pop("jacobi_svd:sweeps", None)
    jacobi = ({"svds": sweeps["launches"], "mean_sweeps": round(sweeps["flops"] / max(sweeps["launches"], 1), 2),
               "max_sweeps": sweeps["bytes"]} if sweeps else None)
    for ph in ("visits", "ticks_gram", "ticks_inner", "ticks_apply"):
        st = stats.pop("jacobi_svd:" + ph, None)
        if st and jacobi:
            jacobi[ph] = st["flops"]
    # executed MFMA work of the Jacobi launches (Gram + apply of the visits made) next to
    # the algorithmic work of the same SVDs (21 m^3 each, Golub-Kahan-Reinsch count)
    alg = stats.pop("jacobi_svd:alg_flops", None)
    if jacobi and "jacobi_svd" in stats:
        js = stats["jacobi_svd"]
        jacobi["executed_gflop_per_step"] = round(js["flops"] / args.steps / 1e9, 3)
        if alg:
            jacobi["algorithmic_gflop_per_step"] = round(alg["flops"] / args.steps / 1e9, 3)
            jacobi["algorithmic_tflops"] = round(alg["flops"] / max(js["ms"], 1e-9) / 1e9, 3)
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
            ach, unit, pk = st["flops"] / st["launches"] / avg_s / 1e12, "TFLOP/s", peak / 1e12
        else:
            ach, unit, pk = st["bytes"] / st["launches"] / avg_s / 1e9, "GB/s", peak / 1e9
        roof = {"kernel": dom, "device_kernel": KERNEL_SYMBOL.get(dom), "bound": bound, "achieved": round(ach, 3),
                "peak": pk, "unit": unit, "frac": round(ach / pk, 5), "traffic": measured_traffic(args, dom),
                "algorithmic_bytes_per_launch": round(st["bytes"] / st["launches"]),
                "avg_launch_us": round(avg_s * 1e6, 2),
                "launches_per_step": st["launches"] / args.steps,
                "share_of_kernel_time": round(st["ms"] / sum(v["ms"] for v in kern.values()), 4)}
    kernels = {k: {"ms_per_step": round(v["ms"] / args.steps, 3), "launches_per_step": v["launches"] // args.steps,
                   "tflops": round(v["flops"] / max(v["ms"], 1e-9) / 1e9, 3),
                   "gbps": round(v["bytes"] / max(v["ms"], 1e-9) / 1e6, 1)} for k, v in kern.items()}
    return jacobi, sites, phases, roof, kernels, cleanup


def inner_main(args):
    """Inner-product workloads: value = inner products per second."""
    import oracle as O
    import ttround as T
    wl = INNER[args.workload]
    h = T.default_handle()
    d, n, r = wl["d"], wl["n"], wl["r"]
    dims = (n,) * d
    m = args.items or wl["pairs"]
    op = T.DeviceTTO(T.hubbard_1d_mpo(d, 1.0, 4.0)) if args.workload == "dotop" else None
    xs = [T.DeviceTT.randn(dims, r, seed=wl["seed_base"] + i) for i in range(m)]
    if op is None:
        w = T.DeviceTT.randn(dims, r, seed=wl["seed_base"] + 99999)
        ys = [w] * m

        def step():
            return T.dot_batch(xs, ys)
    else:
        ys = xs

        def step():
            return T.dot_operator_batch(xs, op, xs)
    log(f"[bench] workload {args.workload}: {m} inner products, d={d}, n={n}, r={r}")
    for _ in range(args.warmup):
        step()
    h.profile(not args.no_events)
    h.profile_reset()
    t0 = time.perf_counter()
    for s in range(args.steps):
        got = step()
        log(f"[bench] step {s} done")
    dt = time.perf_counter() - t0
    h.profile(False)
    args.steps = max(args.steps, 1)
    _, sites, _, roof, kernels, _ = summarize(h.profile_stats(), args)
    value = m * args.steps / dt
    rel, cpu = None, None
    if args.cpu_items > 0:
        args.cpu_items = m  # the whole batch: the oracle takes seconds at these sizes
        H = O.hubbard_1d_mpo(d, 1.0, 4.0) if op is not None else None
        host = [O.TT([t.core(k) for k in range(d)], dims, t.rks, t.ot) for t in xs[:args.cpu_items]]
        wy = O.TT([ys[0].core(k) for k in range(d)], dims, ys[0].rks, ys[0].ot)
        tc = time.perf_counter()
        refs = [O.dot(x, wy) if H is None else O.dot_operator(x, H, x) for x in host]
        tcpu = time.perf_counter() - tc
        if H is None:
            # relative to |x| |w|: a product of independent draws cancels
            nx, nw = T.norm_batch(xs[:args.cpu_items]), T.norm_batch([ys[0]])[0]
            rel = max(abs(g - f) / (a * nw) for g, f, a in zip(got, refs, nx))
        else:
            rel = max(abs(g - f) / abs(f) for g, f in zip(got, refs))
        log(f"[bench] {args.cpu_items} oracle products in {tcpu:.2f}s, max rel err {rel:.3e}")
        cpu = {"value": round(args.cpu_items / tcpu, 4), "unit": "inner products/s", "cores": os.cpu_count(),
               "kind": "port",
               "sample": f"{args.cpu_items} of the batch's products, FP64 numpy/OpenBLAS oracle "
                         f"({'dot' if H is None else 'dot_operator'}), "
                         f"BLAS threads = {os.environ.get('OPENBLAS_NUM_THREADS', 'default')}"}
    res = {
        "metric": "TT inner products/sec" if op is None else "TT operator expectation values/sec",
        "value": round(value, 4), "unit": "inner products/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(1e3 * dt / args.steps, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded N(0,1) TT cores drawn on the device)",
        "config": {"workload": wl["text"], "name": args.workload, "items": m, "d": d, "n": n, "r": r},
        "rel_err_vs_oracle": rel, "kernels": kernels, "gemm_sites": sites, "roofline": roof, "cpu_baseline": cpu,
    }
    for t in xs + ([ys[0]] if op is None else []):
        t.free()
    if op is not None:
        op.free()
    print(json.dumps(res), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=sorted(WORKLOADS) + sorted(INNER), default="c3")
    ap.add_argument("--items", type=int, default=None, help="batch size (default: the workload's)")
    ap.add_argument("--sketch-prec", default=None, help="f64 | f32 | bf16 (default: the workload's)")
    ap.add_argument("--sketch-accum", default=None, choices=["sketch", "f64"],
                    help="sketch: contractions in the sketch precision (FP32 accumulation on FP32/BF16 MFMA); "
                         "f64: the blocks rounded to the sketch precision, contractions in FP64 (default: the "
                         "workload's; C3 uses f64, the mode that meets its 1e-6 error bar on every item)")
    ap.add_argument("--no-cleanup", action="store_true", help="skip the SVD cleanup after the randomized rounding")
    ap.add_argument("--det-items", type=int, default=None,
                    help="items for the FP64 deterministic comparison (0: skip; default: the workload's); their "
                         "materialised inputs must fit in HBM")
    ap.add_argument("--cpu-items", type=int, default=8,
                    help="items checked against the oracle, which is also the CPU baseline (0: skip)")
    ap.add_argument("--cpu-procs", type=int, default=16,
                    help="worker processes of the batch-parallel CPU oracle (one BLAS thread each)")
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
    if args.workload in INNER:
        inner_main(args)
        return
    wl = WORKLOADS[args.workload]
    args.items = args.items or wl["items"]
    args.sketch_prec = args.sketch_prec or wl["sketch"]
    args.sketch_accum = args.sketch_accum or wl["accum"]
    args.det_items = wl["det_items"] if args.det_items is None else args.det_items

    if args.copy_calibration > 0:
        import torch
        n = args.copy_calibration // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):  # an elementwise kernel that reads and writes `BYTES` once each
            torch.mul(src, 2.0, out=dst)
        torch.cuda.synchronize()
        log(f"[bench] copied {args.copy_calibration} bytes x3")
        del src, dst

    import ttround as T
    h = T.default_handle()
    log(f"[bench] workload {args.workload}: {args.items} items, d={wl['d']}, n={wl['n']}")
    t0 = time.time()
    dims, op, items, specs = build_workload(wl, args)
    log(f"[bench] workload ready in {time.time() - t0:.1f}s")

    for w in range(args.warmup):
        outs, _ = run_step(items, op, wl, args)
        log(f"[bench] warmup {w} done; out ranks (item 0, max) = {max(outs[0].rks)}")
        for o in outs:
            o.free()

    # Each timed launch costs ~3 us of host time for its two hipEvents.  Where the
    # dominant kernel class is the FP64 GEMM (C3: long launches), the timed steps
    # record events around the GEMM classes and the phases only; the other classes
    # come from one more, fully profiled, untimed step afterwards.  The other
    # workloads are chains of short dependent launches, where events on every launch
    # would slow the steps themselves: their timed steps record no events and all
    # kernel statistics come from that extra step.
    gemm_only = wl.get("gemm_dominant", False) and not args.no_events
    h.profile(gemm_only, level=2)
    h.profile_reset()
    t0 = time.perf_counter()
    for s in range(args.steps):
        last = s == args.steps - 1
        outs, raw = run_step(items, op, wl, args, keep_raw=last and args.cpu_items > 0)
        if not last:
            for o in outs:
                o.free()
        log(f"[bench] step {s} done")
    dt = time.perf_counter() - t0
    h.profile(False)
    jacobi, sites, phases, roof, kernels, cleanup_stats = summarize(h.profile_stats(), args)
    kernel_source = "timed steps"
    if not args.no_events:
        h.profile(True, level=1)
        h.profile_reset()
        extra_outs, _ = run_step(items, op, wl, args)
        for o in extra_outs:
            o.free()
        h.profile(False)
        one = argparse.Namespace(**vars(args))
        one.steps = 1
        if not gemm_only:
            jacobi, sites, phases, roof, kernels, cleanup_stats = summarize(h.profile_stats(), one)
            kernel_source = "one fully profiled step after the timed steps (which record no events)"
        else:
            jacobi, _, _, _, kx, cleanup_stats = summarize(h.profile_stats(), one)
            kernels = {**{k: v for k, v in kx.items() if not k.startswith("gemm_")}, **kernels}
            kernel_source = "GEMM classes: timed steps; other classes: one more fully profiled step after them"
        if roof and gemm_only:
            tot = sum(v["ms_per_step"] for v in kernels.values())
            roof["share_of_kernel_time"] = round(kernels[roof["kernel"]]["ms_per_step"] / tot, 4)
            top = max(kernels, key=lambda k: kernels[k]["ms_per_step"])
            if top != roof["kernel"]:
                log(f"[bench] warning: {top} takes more time than {roof['kernel']} on this run")
    ms_per_step = 1e3 * dt / args.steps
    value = args.items * args.steps / dt
    out_rks = max(outs[0].rks)

    rel_err, rel_err_raw, cpu, errs = (None, None, None, None)
    if args.cpu_items > 0:
        args.cpu_items = min(args.cpu_items, args.items)
        rel_err, rel_err_raw, cpu, errs = parity_and_cpu(items, specs, op, outs, raw, dims, wl, args)
    # with the parity leg (checked items only; --cpu-items 0 turns both off)
    err_in, err_in_raw = (input_errors(items, op, outs, raw, args.cpu_items) if args.cpu_items > 0
                          else (None, None))
    for o in outs + (raw or []):
        o.free()
    det = deterministic(items, op, wl, args) if args.det_items > 0 else None

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
        "dtype": ("f64" if args.sketch_accum == "f64" or args.sketch_prec == "f64"
                  else f"{args.sketch_prec}-sketch+f64"),
        "data": "synthetic (seeded N(0,1) TT cores drawn on the device)",
        "config": {"workload": wl["text"], "name": args.workload, "items": args.items, "d": wl["d"], "n": wl["n"],
                   "terms": wl["terms"], "sketch_prec": args.sketch_prec,
                   "sketch_accum": args.sketch_accum,
                   "cleanup": None if args.no_cleanup else wl["cleanup"]},
        "rel_err_vs_oracle": rel_err,
        "rel_err_vs_oracle_randomized_stage": rel_err_raw,
        "rel_err_vs_oracle_items": None if errs is None else [float(f"{e:.3e}") for e in errs],
        # approximation quality (largest over the same items): against the input itself
        "rel_err_vs_input": err_in,
        "rel_err_vs_input_randomized_stage": err_in_raw,
        "out_rank_max_item0": out_rks,
        "fp64_deterministic": det,
        "speedup_vs_fp64_deterministic": (round(value / det["rounded_tts_per_s"], 3) if det else None),
        "phases_ms_per_step": phases,
        "svd_sweep_bonds": cleanup_stats,
        "jacobi": jacobi,
        "kernels": kernels,
        "kernel_stats_source": kernel_source,
        "gemm_sites": sites,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
