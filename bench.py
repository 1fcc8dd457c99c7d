#!/usr/bin/env python3
"""PatchTable benchmark on one MI355X: query Mpatch/s at matched mean NNF SSD (4MP pair) and
table build ms.

Workload: the plain PatchTable at speed 1 (PatchTableParams() + set_speed(1)) on a seeded
synthetic 2048x2048 query vs 2048x2048 database pair (seed 3). A step is one lookup() of the
whole query -- query-side colour conversion + GCK, table candidates, k-coherence, spatial search
and the propagation rounds -- with both images already resident in HBM. Table build time is the
whole constructor (GCK, quantiser, populate, DT, k-coherence) from a resident database image.

Prints one JSON line on stdout; progress goes to stderr. The CPU baseline is the project's CPU
restatement of the reference (oracle/liboracle.so, one thread, raster-order lookup) on a bounded
1024x768 pair from the same generator, and the GPU's mean NNF SSD on that same pair is compared
against it ("matched mean NNF SSD").
"""
import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "query Mpatch/s at matched mean NNF SSD (4MP pair); table build ms"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def lookup_alg_bytes(ah, aw, bh, bw, K):
    """Algorithmic bytes of one lookup: read a and b once (12 B/px) and, per query patch, one
    table cell (4 B), one k-coherence list (8 B per entry) and the NNF output (3 doubles)."""
    nq = (ah - 7) * (aw - 7)
    return 12 * ah * aw + 12 * bh * bw + nq * (4 + 8 * K + 24)


def propagate_alg_bytes(ah, aw, bh, bw):
    """Algorithmic bytes of one propagation round (k_propagate): read the current field and write
    the next one (x, y, dist: 12 B each per query patch) and read a and b once (12 B/px)."""
    nq = (ah - 7) * (aw - 7)
    return 24 * nq + 12 * ah * aw + 12 * bh * bw


TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "k_propagate_2048x2048_traffic.md")


def measured_traffic(H, W):
    """HBM bytes per k_propagate launch from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE
    summary for this workload (already scaled by the copy-kernel calibration), or None."""
    if (H, W) != (2048, 2048) or not os.path.exists(TRAFFIC_SUMMARY):
        return None
    for line in open(TRAFFIC_SUMMARY):
        if line.startswith("traffic_bytes_per_launch:"):
            return int(line.split(":")[1])
    return None


COUNTER_SUMMARY = os.path.join(ROOT, "profiles", "lookup_counters_2048x2048_speed1.md")


def measured_l1_rate(H, W):
    """k_propagate's L1 tag lookups per CU per cycle (the TCP handles at most one), from the
    committed rocprofv3 counter summary for this workload, or None."""
    if (H, W) != (2048, 2048) or not os.path.exists(COUNTER_SUMMARY):
        return None
    for line in open(COUNTER_SUMMARY):
        if line.startswith("| L1 tag lookups (TCP_TOTAL_CACHE_ACCESSES"):
            cols = [c.strip() for c in line.strip().strip("|").split("|")]
            return float(cols[2])   # columns: label, k_candidates, k_propagate, k_select
    return None


def copy_probe(torch, mib=1024, reps=3):
    """A plain device-to-device copy of known size, for calibrating HBM byte counters: each
    dispatch reads and writes exactly `mib` MiB."""
    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    log(f"copy probe: {reps} copies of {mib} MiB")
    del src, dst


def one_vs_many(args):
    """One 1024x1024 query (seed 4) against 32 databases of sizes drawn uniformly from 640..1920 x
    480..1080 (numpy PCG64 seed 999) with seeds 1000..1031, speed 1: all tables built, then each
    step looks the query up in every table (shared query descriptor, one stream per table, HIP
    graph replays). Prints one JSON line with the whole-batch throughput and p50/p99 per-pair
    device latency."""
    import numpy as np
    import torch
    from patchtable_amd import PatchTableBatch, PatchTableParams, synth_pair
    rng = np.random.Generator(np.random.PCG64(999))
    n = args.databases
    sizes = [(int(rng.integers(480, 1081)), int(rng.integers(640, 1921))) for _ in range(n)]
    log(f"one-vs-many: generating 1 query and {n} databases")
    a = torch.from_numpy(synth_pair(4, 1024, 1024)[0]).cuda()
    dbs = [torch.from_numpy(synth_pair(1000 + i, h, w)[1]).cuda() for i, (h, w) in enumerate(sizes)]
    p = PatchTableParams()
    p.set_speed(1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    batch = PatchTableBatch(p, dbs, threads=4)
    torch.cuda.synchronize()
    build_s = time.perf_counter() - t0
    log(f"built {n} tables in {build_s * 1e3:.1f} ms")
    anns = None
    for _ in range(args.warmup):
        anns, ms, secs = batch.lookup(a, anns)
    per_pair, walls = [], []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        anns, ms, secs = batch.lookup(a, anns)
        per_pair += ms
        walls.append(secs)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    nq = (1024 - 7) * (1024 - 7) * n
    mean_ssd = float(np.mean([x[..., 2].mean().item() for x in anns]))
    out = {
        "metric": "one-vs-many query Mpatch/s (1024x1024 query vs 32 databases 640x480..1920x1080)",
        "value": round(nq / (ms_per_step / 1e3) / 1e6, 3), "unit": "Mpatch/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"one-vs-many: 1024x1024 query (seed 4) vs {n} databases (seeds 1000.., sizes from seed 999)",
                   "params": "PatchTableParams()+set_speed(1)", "databases": sizes},
        "pair_latency_ms": {"p50": round(float(np.percentile(per_pair, 50)), 3),
                            "p99": round(float(np.percentile(per_pair, 99)), 3),
                            "max": round(float(np.max(per_pair)), 3)},
        "build_all_ms": round(build_s * 1e3, 1),
        "mean_ssd": round(mean_ssd, 6),
    }
    print(json.dumps(out), flush=True)


def timed_lookups(torch, table, a, ann, warmup, steps):
    """Wall ms per lookup over `steps` calls (after `warmup`) and the last lookup's stage times."""
    for _ in range(warmup):
        table.lookup(a, ann)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        table.lookup(a, ann)
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) * 1e3 / steps, table.stage_ms()


def learned_key(args):
    """Learned key: a 192 -> 64 -> 8 key network trained on the GPU (triplets mined from the
    library's own lookups of 64 synthetic 1024x1024 pairs, seeds 2000..2063, 31250 per pair:
    ~2M triplets; 2000 Adam steps of 4096), then the held-out 2048x2048 pair (seed 3) looked up in
    the key table (with its GCK fallback) and in the plain GCK table, at speed 1 and at fewer
    propagation rounds. Reports each point's query Mpatch/s and mean NNF SSD, and the fastest
    keyed point whose SSD is at or below the plain table's at speed 1."""
    import numpy as np
    import torch
    from patchtable_amd import PatchKey, PatchTable, PatchTableParams, synth_pair

    def params(extra=""):
        p = PatchTableParams()
        p.set_speed(1)
        if extra:
            p.set_from_switches("-x 0 " + extra)
        return p

    npairs, S, tpp, steps, batch = 64, 1024, 31250, 2000, 4096
    log(f"learned key: generating {npairs} training pairs of {S}x{S}")
    pairs = [tuple(torch.from_numpy(x).cuda() for x in synth_pair(2000 + i, S, S)) for i in range(npairs)]
    key = PatchKey(1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    losses = key.train(params(), pairs, tpp=tpp, steps=steps, batch=batch, lr=2e-3, margin=1.0, seed=9)
    torch.cuda.synchronize()
    train_s = time.perf_counter() - t0
    log(f"trained in {train_s:.2f} s: loss {losses[:10].mean():.4f} -> {losses[-50:].mean():.4f}")
    del pairs
    H = W = 2048
    a_np, b_np = synth_pair(3, H, W)
    a, b = torch.from_numpy(a_np).cuda(), torch.from_numpy(b_np).cuda()
    nq = (H - 7) * (W - 7)
    ann = torch.empty((H - 7, W - 7, 3), dtype=torch.float64, device="cuda")
    points = []
    for extra in ["", "-jf_rounds 4", "-jf_rounds 2", "-jf_rounds 1", "-do_prop 0"]:
        for kind in ("gck", "key"):
            p = params(extra)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            t = PatchTable(p, b) if kind == "gck" else PatchTable.keyed(p, key, b)
            torch.cuda.synchronize()
            build_ms = (time.perf_counter() - t0) * 1e3
            ms, st = timed_lookups(torch, t, a, ann, args.warmup, args.steps)
            ssd = float(ann[..., 2].mean().item())
            points.append({"table": kind, "switches": extra or "(speed 1)", "ms": round(ms, 3),
                           "mpatch_s": round(nq / ms / 1e3, 2), "mean_ssd": round(ssd, 6),
                           "build_ms": round(build_ms, 1), "query_key_ms": round(st["query_gck"], 3)})
            log(json.dumps(points[-1]))
            t.close()
    base = points[0]
    ok = [q for q in points if q["table"] == "key" and q["mean_ssd"] <= base["mean_ssd"]]
    best = max(ok, key=lambda q: q["mpatch_s"]) if ok else None
    key_default = points[1]
    out = {
        "metric": "learned-key query Mpatch/s (2048x2048 pair, speed 1) beside the plain GCK table",
        "value": key_default["mpatch_s"], "unit": "Mpatch/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": key_default["ms"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "bf16 key network, f32 SSD", "data": "synthetic",
        "config": {"workload": "learned key: train on 64 1024x1024 pairs (seeds 2000..2063), evaluate the held-out "
                               "2048x2048 pair (seed 3)", "triplets": npairs * tpp, "adam_steps": steps,
                   "batch": batch},
        "train_s": round(train_s, 2),
        "loss_first_last": [round(float(losses[:10].mean()), 5), round(float(losses[-50:].mean()), 5)],
        "points": points,
        "plain_speed1": {"mpatch_s": base["mpatch_s"], "mean_ssd": base["mean_ssd"]},
        "fastest_key_at_or_below_plain_ssd": best,
        "speedup_at_equal_ssd": round(best["mpatch_s"] / base["mpatch_s"], 3) if best else None,
    }
    print(json.dumps(out), flush=True)


def key_forward(args):
    """The learned key's forward pass over the 2048x2048 query (seed 3): 4,165,681 patches x
    25,600 FLOP (192x64 + 64x8 multiply-adds), BF16 layer 1 and FP32 layer 2 on MFMA. Times
    `steps` calls of pt_key_forward_dev into a preallocated plane buffer."""
    import ctypes
    import torch
    from patchtable_amd import PatchKey, _ffi, synth_pair
    H = W = 2048
    a = torch.from_numpy(synth_pair(3, H, W)[0]).cuda()
    key = PatchKey(1)
    planes = torch.zeros((8, H + 7, W + 7), dtype=torch.float32, device="cuda")
    lib = _ffi.load()

    def call():
        _ffi.check(lib.pt_key_forward_dev(key._k, ctypes.c_void_p(a.data_ptr()), H, W,
                                          ctypes.c_void_p(planes.data_ptr())))
    for _ in range(args.warmup):
        call()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        call()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    n = (H - 7) * (W - 7)
    tf = n * 25600 / (ms / 1e3) / 1e12
    out = {"metric": "learned-key forward Mpatch/s (2048x2048 image)", "value": round(n / ms / 1e3, 2),
           "unit": "Mpatch/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 4),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "bf16 (layer 2 f32)",
           "data": "synthetic", "config": {"workload": "key forward over a 2048x2048 image (seed 3), 192-64-8 MLP"},
           "roofline": {"bound": "mfma", "achieved": round(tf, 1), "peak": 2500.0, "unit": "TFLOP/s",
                        "frac": round(tf / 2500.0, 4), "traffic": None,
                        "note": "wall time per call including its stream synchronisation"}}
    print(json.dumps(out), flush=True)


def brute(args):
    """Exhaustive search on the 2048x2048 pair (seed 3): 65,536 query patches (every 8th in x and
    y) against all 4,165,681 database patches, candidate scoring |a|^2+|b|^2-2a.b on MFMA in BF16
    and in FP8 (e4m3, per-patch scale), 8 best rescored in exact FP32. The FP64 exhaustive search on
    4,096 of those queries is the exact reference for both and for the PatchTable lookup of the
    same pair at speed 1."""
    import numpy as np
    import torch
    from patchtable_amd import PatchTable, PatchTableParams, brute_exact_f64, brute_lookup, synth_pair
    H = W = 2048
    a_np, b_np = synth_pair(3, H, W)
    a, b = torch.from_numpy(a_np).cuda(), torch.from_numpy(b_np).cuda()
    gx, gy = np.meshgrid(np.arange(0, W - 7, 8), np.arange(0, H - 7, 8))
    qpos_np = ((gy.ravel() << 16) | gx.ravel()).astype(np.int32)
    qpos = torch.from_numpy(qpos_np).cuda()
    nq, nd = len(qpos_np), (H - 7) * (W - 7)
    sub = torch.from_numpy(qpos_np[::16].copy()).cuda()          # 4,096 queries for the FP64 reference
    log(f"brute: {nq} queries x {nd} database patches; FP64 reference on {sub.numel()} queries")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    epos, edist = brute_exact_f64(a, b, sub)
    torch.cuda.synchronize()
    f64_s = time.perf_counter() - t0
    epos, edist = epos.cpu().numpy(), edist.cpu().numpy()
    log(f"FP64 reference: {f64_s:.2f} s")
    res = {}
    for prec, peak in (("bf16", 2500.0), ("fp8", 5000.0)):
        for _ in range(args.warmup):
            brute_lookup(a, b, qpos, precision=prec)
        ms, search = [], []
        for _ in range(args.steps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            r = brute_lookup(a, b, qpos, precision=prec)
            torch.cuda.synchronize()
            ms.append((time.perf_counter() - t0) * 1e3)
            search.append(r["stage_ms"]["search"])
        pos, dist = r["pos"].cpu().numpy(), r["dist"].cpu().numpy()
        ps, ds = pos[::16], dist[::16]
        flops = 2.0 * 192 * nq * nd
        tf = flops / (statistics.median(search) / 1e3) / 1e12
        res[prec] = {"ms": round(statistics.median(ms), 2), "stage_ms": {k: round(v, 3) for k, v in r["stage_ms"].items()},
                     "mpatch_s": round(nq / statistics.median(ms) / 1e3, 3),
                     "roofline": {"bound": "mfma", "achieved": round(tf, 1), "peak": peak, "unit": "TFLOP/s",
                                  "frac": round(tf / peak, 4), "traffic": None},
                     "exact_nn_fraction": round(float(np.mean(ps == epos)), 5),
                     "mean_ssd_ratio_vs_f64": round(float(ds.mean() / edist.mean()), 6),
                     "max_ssd_ratio_vs_f64": round(float((ds / np.maximum(edist, 1e-30)).max()), 4)}
        log(prec, json.dumps(res[prec]))
    # the PatchTable lookup of the same pair against the exact nearest neighbours
    p = PatchTableParams()
    p.set_speed(1)
    t = PatchTable(p, b)
    ann = t.lookup(a).cpu().numpy()
    sx, sy = sub.cpu().numpy() & 0xffff, sub.cpu().numpy() >> 16
    tab = ann[sy, sx]
    tab_pos = ((tab[:, 1].astype(np.int64) << 16) | tab[:, 0].astype(np.int64)).astype(np.int32)
    table_vs = {"exact_nn_fraction": round(float(np.mean(tab_pos == epos)), 5),
                "mean_ssd_ratio_vs_f64": round(float(tab[:, 2].mean() / edist.mean()), 6)}
    best = res["bf16"]
    out = {
        "metric": "exhaustive search query Mpatch/s (65,536 queries vs 4,165,681 database patches, BF16 MFMA "
                  "scoring + exact FP32 rescoring)",
        "value": best["mpatch_s"], "unit": "Mpatch/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": best["ms"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16 (fp8 beside it), f32 rescoring", "data": "synthetic",
        "config": {"workload": "2048x2048 query vs 2048x2048 database (seed 3), every 8th query patch in x and y",
                   "queries": nq, "database_patches": nd, "f64_reference_queries": int(sub.numel())},
        "roofline": best["roofline"],
        "bf16": res["bf16"], "fp8": res["fp8"], "f64_reference_s": round(f64_s, 3),
        "patchtable_speed1_vs_f64": table_vs,
    }
    print(json.dumps(out), flush=True)


def cpu_baseline(h=768, w=1024, seed=2):
    """Oracle (CPU restatement of the reference, 1 thread) on a bounded pair; returns its lookup
    Mpatch/s, build ms and raster mean SSD, plus the GPU mean SSD on the same pair."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_ffi as O
    from patchtable_amd import PatchTable, PatchTableParams, synth_pair
    a, b = synth_pair(seed, h, w)
    p = PatchTableParams()
    p.set_speed(1)
    t0 = time.perf_counter()
    ot = O.OracleTable(p.raw, b)
    t1 = time.perf_counter()
    ref = ot.lookup_raster(a)
    t2 = time.perf_counter()
    nq = (h - 7) * (w - 7)
    gt = PatchTable(p, b)
    ann = gt.lookup(a)
    return dict(mpatch_s=nq / (t2 - t1) / 1e6, build_ms=(t1 - t0) * 1e3, lookup_s=t2 - t1,
                cpu_mean_ssd=float(ref[..., 2].mean()), gpu_mean_ssd=float(ann[..., 2].mean()),
                cpu_mean_dist=O.mean_dist(ref), gpu_mean_dist=O.mean_dist(ann), nq=nq)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--size", default="2048x2048", help="HxW of query and database")
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--build-reps", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--workload", default="pair",
                    choices=["pair", "one_vs_many", "learned_key", "key_forward", "brute"],
                    help="pair: the 4MP pair of the headline metric (default); one_vs_many: one query vs many "
                         "databases; learned_key: train the key network, then key table vs GCK table; "
                         "key_forward: the key network's forward pass alone; brute: exhaustive MFMA search "
                         "against the FP64 exact nearest neighbours")
    ap.add_argument("--databases", type=int, default=32)
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run a 1 GiB device copy (counter calibration for the traffic figure)")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: the PatchTable hot path is built and timed on a single "
              "MI355X; multi-GPU sharding is not run.", file=sys.stderr)
        sys.exit(2)

    import torch
    from patchtable_amd import PatchTable, PatchTableParams, synth_pair

    if args.workload == "one_vs_many":
        one_vs_many(args)
        return
    if args.workload == "learned_key":
        learned_key(args)
        return
    if args.workload == "brute":
        brute(args)
        return
    if args.workload == "key_forward":
        key_forward(args)
        return

    H, W = (int(v) for v in args.size.lower().split("x"))
    log(f"generating {H}x{W} pair seed {args.seed}")
    a_np, b_np = synth_pair(args.seed, H, W)
    torch.cuda.set_device(0)
    a = torch.from_numpy(a_np).cuda()
    b = torch.from_numpy(b_np).cuda()
    p = PatchTableParams()
    p.set_speed(1)

    # --- table build ------------------------------------------------------------------------
    build_ms, stages = [], None
    table = None
    for r in range(args.build_reps + 1):
        if table is not None:
            table.close()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        table = PatchTable(p, b)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        if r > 0:
            build_ms.append((t1 - t0) * 1e3)
        stages = table.stage_ms()
        log(f"build rep {r}: {(t1 - t0) * 1e3:.1f} ms  stages {json.dumps({k: round(v, 3) for k, v in stages.items()})}")
    info = table.info()

    # --- lookup steps -----------------------------------------------------------------------
    ann = torch.empty((H - 7, W - 7, 3), dtype=torch.float64, device="cuda")
    for _ in range(args.warmup):
        table.lookup(a, ann)
    torch.cuda.synchronize()
    prop_ms, cand_ms, total_ms = [], [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        table.lookup(a, ann)
        st = table.stage_ms()
        prop_ms.append(st["query_propagation"])
        cand_ms.append(st["query_candidates"])
        total_ms.append(st["query_total"])
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    nq = (H - 7) * (W - 7)
    value = nq / (ms_per_step / 1e3) / 1e6
    mean_ssd = float(ann[..., 2].mean().item())
    mean_dist = float(ann[..., 2].clamp(min=0).sqrt().mean().item())
    log(f"lookup: {ms_per_step:.3f} ms/step, {value:.1f} Mpatch/s, mean SSD {mean_ssd:.6f}")

    # --- roofline of the dominant kernel (k_propagate) -----------------------------------------
    launches = max(1, int(st["propagation_launches"]))
    prop_launch_ms = statistics.mean(prop_ms) / launches
    prop_bytes = propagate_alg_bytes(H, W, H, W)
    achieved = prop_bytes / (prop_launch_ms / 1e3) / 1e9
    roofline = {"bound": "hbm", "kernel": "k_propagate", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": measured_traffic(H, W),
                "alg_bytes_per_launch": prop_bytes, "avg_launch_ms": round(prop_launch_ms, 4),
                "launches_per_step": launches}
    l1 = measured_l1_rate(H, W)
    if l1 is not None:   # the limit that binds: the L1 gathers of the exact SSDs, not HBM
        roofline["l1_tag_lookups_per_cu_cycle"] = {"achieved": l1, "peak": 1.0, "frac": l1,
                                                   "source": "rocprofv3 TCP_TOTAL_CACHE_ACCESSES, "
                                                             "profiles/lookup_counters_2048x2048_speed1.md"}
    lk_bytes = lookup_alg_bytes(H, W, H, W, info["kcoherence"])
    lookup_ms = statistics.mean(total_ms)

    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Mpatch/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"plain PatchTable speed 1, {H}x{W} query vs {H}x{W} database (synthetic seed {args.seed})",
                   "query": f"{H}x{W}", "database": f"{H}x{W}", "params": "PatchTableParams()+set_speed(1)",
                   "nslices": info["nslices"], "ncells": info["ncells"], "jf_rounds": p.jf_rounds},
        "table_build_ms": round(statistics.median(build_ms), 3),
        "table_build_stage_ms": {k: round(v, 3) for k, v in stages.items() if k in
                                 ("gck", "quantiser", "populate", "dt", "kcoherence", "build_total")},
        "lookup_stage_ms": {"query_gck": round(st["query_gck"], 4), "candidates": round(statistics.mean(cand_ms), 4),
                            "propagation": round(statistics.mean(prop_ms), 4), "total": round(lookup_ms, 4)},
        "lookup_alg_GBs": round(lk_bytes / (lookup_ms / 1e3) / 1e9, 2),
        "mean_ssd": round(mean_ssd, 6),
        "mean_dist": round(mean_dist, 6),
        "roofline": roofline,
    }

    if args.copy_probe:
        copy_probe(torch)

    if not args.no_cpu_baseline:
        log("cpu baseline: oracle (1 thread) on a 1024x768 pair, seed 2")
        cb = cpu_baseline()
        out["cpu_baseline"] = {"value": round(cb["mpatch_s"], 4), "unit": "Mpatch/s", "cores": 1, "kind": "port",
                               "sample": "oracle raster-order lookup at speed 1 on a 1024x768 pair (seed 2) of the "
                                         "same generator, 1 thread; build %.0f ms" % cb["build_ms"],
                               "build_ms": round(cb["build_ms"], 1)}
        out["matched_ssd_sample"] = {"pair": "1024x768 seed 2", "gpu_mean_ssd": round(cb["gpu_mean_ssd"], 6),
                                     "cpu_mean_ssd": round(cb["cpu_mean_ssd"], 6),
                                     "ratio": round(cb["gpu_mean_ssd"] / cb["cpu_mean_ssd"], 5),
                                     "gpu_mean_dist": round(cb["gpu_mean_dist"], 6),
                                     "cpu_mean_dist": round(cb["cpu_mean_dist"], 6)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
