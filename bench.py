#!/usr/bin/env python
"""Construction benchmark: RecSplit MPHF of n = 5,000,000 keys, l = 16, b = 2000.

One step = one complete GPURecSplit build (bucket partition, every seed search,
Golomb-Rice + Elias-Fano encoding, copy of the structure to the host) of the
5M splitmix64(seed 1) keys, which are resident in HBM before timing starts.
Prints one JSON line (metric ns/object, lower is better) with:
  roofline      the dominant search kernel (and every other search kernel) against
                MI355X's int32 vector issue peak, in algorithmic operations
  cpu_baseline  the CPU oracle (SimdRecSplit algorithm, vectorised C) on a bounded
                sample of the same workload, all host threads and 1 thread
and, as extra keys, the other workloads measured on the same GPU:
  plain_l18_b50_n5M   5M keys (seed 2), l = 18, b = 50
  plain_l24_scaled    40,000 keys (seed 24), l = 24, b = 2000 (20 buckets)
  batched_4x5M        four 5M-key l = 16 sets in one batch vs the plain build
  batched_4096        4096 sets, log-uniform sizes in [1e3, 1e6] (~600M keys), l = 12,
                      b = 100, one batch, vs the same 4096 sets built one by one
  served              a scaled served stream through the construction server
  served_half_rate    the same stream at half the arrival rate (latency without
                      queueing at the server's capacity)
  level_tails         busy share and exit tail of every search launch (one more,
                      untimed build with RS_WAVE_TIMES=1)
  dropin_host_keys    the headline build from host keys (rs_build: PCIe upload included)
  plan_graph          small builds: ordinary build vs build plan vs captured-graph replay

    python bench.py [--gpus 1] [--steps 3] [--warmup 1] [--quick]
"""
import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

# Roofline of the seed search: int32 vector issue.  Peak = MI355X's published FP32
# vector peak (157.3 TFLOP/s) counted as lane operations (one per FMA).
VALU_PEAK_TOPS = 157.3 / 2
# Algorithmic int32 operations of one key-seed evaluation, from the formula the
# reference evaluates (AbstractParallelRecSplit.hpp:277-284, 327-337): 64-bit
# key + seed (2), fold of the halves (1), MurmurHash3 fmix32 to its low 16 bits
# (xor-shift 2, multiply 1, xor-shift 2, multiply 1, xor-shift 2, mask 1 = 9),
# remap (multiply, shift = 2) and the node predicate's update (2: part count
# increment, bit set, or compare-add).
OPS_PER_EVAL = 16
KINDS = ["fanout2", "upper_aggr", "lower_aggr", "leaf_rotation", "leaf_bruteforce"]

# HBM bytes per launch of the search kernels at the plain l = 16 workload, from
# two rocprofv3 --pmc runs (FETCH_SIZE scaled x2.0, WRITE_SIZE x1.0: the factors of
# a 256 MiB copy probe); the profile summaries under profiles/ give the commands.
# Most of it is register spills written around chunk claims, not key traffic.
TRAFFIC_BYTES = {"lower_aggr": 15.90e9, "upper_aggr": 10.40e9, "leaf_rotation": 12.20e9}


def log(msg: str):
    print(f"# {msg}", file=sys.stderr, flush=True)


def kernel_rooflines(stats: list) -> dict:
    """Per search kernel: event-timed ms (mean over steps), algorithmic
    evaluations, evaluations/s, achieved TOP/s and fraction of the peak."""
    out = {}
    last = stats[-1]
    for k in KINDS:
        ms = float(np.mean([s["search_ms"][k] for s in stats]))
        ev = last["evals"][k]
        if ms <= 0 or ev == 0:
            continue
        eps = ev / (ms / 1e3)
        out[k] = {"kernel_ms": ms, "evals": ev, "evals_per_s": eps, "achieved": eps * OPS_PER_EVAL / 1e12,
                  "frac": eps * OPS_PER_EVAL / 1e12 / VALU_PEAK_TOPS, "nodes": last["nodes"][k],
                  "exec_ratio": last["executed_evals"][k] / ev,
                  "seeds": last["seeds"][k], "seeds_per_s": last["seeds"][k] / (ms / 1e3)}
    return out


def plain(gra, ctx, d_keys, n, leaf, bucket, steps, warmup, label) -> dict:
    """Times `steps` complete builds of the resident keys (after `warmup`)."""
    import torch

    for i in range(warmup):
        rs = gra.GPURecSplit(d_keys, bucket, leaf_size=leaf, ctx=ctx)
        log(f"{label} warmup {i}: {rs.stats()['total_ms']:.1f} ms")
        del rs
    torch.cuda.synchronize()
    stats = []
    t0 = time.perf_counter()
    for i in range(steps):
        rs = gra.GPURecSplit(d_keys, bucket, leaf_size=leaf, ctx=ctx)
        stats.append(rs.stats())
        log(f"{label} step {i}: {stats[-1]['total_ms']:.1f} ms")
        if i + 1 < steps:
            del rs
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    idx = rs.query_device(d_keys)  # verification outside the timed region
    ok = bool(torch.equal(torch.sort(idx).values, torch.arange(n, device=idx.device)))
    last = stats[-1]
    return {"ms_per_step": ms, "value": ms * 1e6 / n, "unit": "ns/object", "n": n, "leaf": leaf, "bucket": bucket,
            "steps": steps, "bits_per_key": rs.get_bits() / n, "bijection_verified": ok,
            "kernels": kernel_rooflines(stats),
            "phase_ms": {"partition": last["partition_ms"], "search": last["search_ms"], "split": last["split_ms"],
                         "encode": last["encode_ms"], "total": last["total_ms"]},
            "evals_per_s": sum(last["evals"].values()) / (sum(last["search_ms"].values()) / 1e3),
            "seeds_tested_per_s": sum(last["seeds"].values()) / (sum(last["search_ms"].values()) / 1e3)}


def cpu_baseline(leaf: int, bucket: int, seed: int, threads: int, buckets_per_thread: int = 5) -> dict:
    """Oracle (vectorised C restatement of SimdRecSplit) on a bounded sample of the
    same key stream: `buckets_per_thread` buckets per thread, median of 3 runs."""
    sys.path.insert(0, str(ROOT / "tests"))
    from oracle_lib import OracleMPHF, splitmix_keys

    n_s = bucket * buckets_per_thread * threads
    keys = splitmix_keys(seed, n_s)
    walls = []
    for _ in range(3):
        t0 = time.perf_counter()
        rs = OracleMPHF.build(keys, leaf, bucket, rotate=True, threads=threads)
        walls.append(time.perf_counter() - t0)
    assert rs.bijection_errors(keys) == 0
    dt = float(np.median(walls))
    return {"value": dt / n_s * 1e9, "unit": "ns/object", "cores": threads, "kind": "port",
            "sample": f"first {n_s} keys of splitmix64(seed {seed}) ({n_s // bucket} buckets of {bucket}), "
                      f"l={leaf}, rotation on, {threads} thread(s), median of 3 runs "
                      f"({', '.join(f'{w:.2f}' for w in walls)} s)"}


def probe(device: int = 0) -> dict:
    """VALU issue rates (lane-ops/s) of the instructions of the seed-search inner
    loop, from bench_probe.hip, plus one calibration copy of 256 MiB."""
    import ctypes as C

    lib = C.CDLL(str(ROOT / "bench_probe.so"))
    rates = {}
    for kind, name in enumerate(["v_xor_b32", "v_mul_lo_u32", "v_lshl_add_u64", "v_mul_u32_u24", "v_xor_b32_sdwa",
                                 "v_mad_u32_u24", "v_lshlrev_b32_sdwa"]):
        r, ms = C.c_double(), C.c_double()
        assert lib.rs_probe_valu(device, kind, 256, C.byref(r), C.byref(ms)) == 0
        rates[name] = r.value
    ms = C.c_double()
    assert lib.rs_probe_copy(device, 256 << 20, C.byref(ms)) == 0
    base = rates["v_xor_b32"]
    return {"lane_ops_per_s": rates, "issue_slots": {k: base / v for k, v in rates.items()},
            "copy_256MiB_ms": ms.value, "copy_GBps": 2 * (256 << 20) / (ms.value / 1e3) / 1e9}


def batched_4x5m(gra, ctx, args, plain_ms: float) -> dict:
    """Four sets of n keys (splitmix64 seeds seed..seed+3) built in one batch."""
    import torch
    from gpurecsplit_amd.keys import splitmix64_keys_device

    B = 4
    d = torch.cat([splitmix64_keys_device(args.seed + i, args.n) for i in range(B)])
    off = np.arange(B + 1, dtype=np.uint64) * np.uint64(args.n)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    rss, st = gra.build_batch(d, off, args.bucket, leaf_size=args.leaf, ctx=ctx)
    ms = (time.perf_counter() - t0) * 1e3
    log(f"batched 4x{args.n}: {ms:.1f} ms")
    same = rss[0].serialize() == gra.GPURecSplit(d[: args.n], args.bucket, leaf_size=args.leaf, ctx=ctx).serialize()
    return {"sets": B, "keys_per_set": args.n, "ms": ms, "value": ms * 1e6 / (B * args.n), "unit": "ns/object",
            "objects_per_s_vs_plain": (B * args.n / ms) / (args.n / plain_ms), "first_set_equals_plain": same,
            "search_share": sum(st["search_ms"].values()) / st["total_ms"]}


def batched_4096(gra, ctx, quick: bool, sequential: bool = True) -> dict:
    """4096 sets, per-set sizes log-uniform in [1e3, 1e6] (numpy default_rng(3)),
    keys from one SplitMix64(3) stream generated in HBM, l = 12, b = 100: one
    untimed batch that sizes the context's buffers, one timed batch, then the same
    sets as 4096 plain builds one after another.  Every set of the timed batch is
    checked for bijectivity by the GPU query kernel, and 8 sets byte for byte
    against the oracle."""
    import torch
    from gpurecsplit_amd.keys import splitmix64_keys_device

    sys.path.insert(0, str(ROOT / "tests"))
    from oracle_lib import OracleMPHF

    rng = np.random.default_rng(3)
    hi = 100_000 if quick else 1_000_000
    sizes = np.exp(rng.uniform(np.log(1_000), np.log(hi), size=4096)).astype(np.int64)
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    n = int(off[-1])
    log(f"4096 sets: {n} keys")
    d = splitmix64_keys_device(3, n)
    torch.cuda.synchronize()
    # first build: grows the context's device buffers to this size (untimed
    # warm-up, reported); the timed build reuses them, like every plain step
    t0 = time.perf_counter()
    rss, st = gra.build_batch(d, off, 100, leaf_size=12, ctx=ctx)
    ms_first = (time.perf_counter() - t0) * 1e3
    del rss
    log(f"4096 sets batched, first build: {ms_first:.1f} ms")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    rss, st = gra.build_batch(d, off, 100, leaf_size=12, ctx=ctx)
    ms = (time.perf_counter() - t0) * 1e3
    log(f"4096 sets batched: {ms:.1f} ms (library {st['total_ms']:.1f} ms)")
    bij = True
    for i in range(4096):
        a, b = int(off[i]), int(off[i + 1])
        idx = rss[i].query_device(d[a:b])
        bij &= bool(torch.equal(torch.sort(idx).values, torch.arange(b - a, device=idx.device)))
    sample = list(range(0, 4096, 512))
    oracle_ok = all(rss[i].serialize() == OracleMPHF.build(d[int(off[i]):int(off[i + 1])].cpu().numpy().view(np.uint64),
                                                           12, 100, threads=16).serialize() for i in sample)
    del rss
    seq_ms = float("nan")
    if sequential:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for i in range(4096):
            rs = gra.GPURecSplit(d[int(off[i]):int(off[i + 1])], 100, leaf_size=12, ctx=ctx)
            del rs
        torch.cuda.synchronize()
        seq_ms = (time.perf_counter() - t0) * 1e3
        log(f"4096 sets one by one: {seq_ms:.1f} ms")
    return {"sets": 4096, "keys": n, "leaf": 12, "bucket": 100, "sizes": f"log-uniform [1e3, {hi:.0e}]",
            "ms": ms, "value": ms * 1e6 / n, "unit": "ns/object", "ms_first_build": ms_first,
            "library_total_ms": st["total_ms"], "search_ms": st["search_ms"],
            "kernels": kernel_rooflines([st]),
            "partition_ms": st["partition_ms"], "encode_ms": st["encode_ms"],
            "sequential_plain_ms": seq_ms, "objects_per_s_vs_sequential_plain": seq_ms / ms,
            "bijection_all_sets": bij, "oracle_bytes_sets": sample, "oracle_bytes_equal": oracle_ok}


def dropin_host_keys(gra, ctx, d_keys, leaf: int, bucket: int, device_ms: float, steps: int = 2) -> dict:
    """The drop-in constructor as a C++ caller uses it (rs_build): the keys start in
    host memory and the build copies them to HBM (PCIe-inclusive).  Not the headline
    value, which times device-resident keys."""
    keys = np.ascontiguousarray(d_keys.cpu().numpy().view(np.uint64))
    gra.GPURecSplit(keys, bucket, leaf_size=leaf, ctx=ctx)
    times = []
    for _ in range(steps):
        t0 = time.perf_counter()
        rs = gra.GPURecSplit(keys, bucket, leaf_size=leaf, ctx=ctx)
        times.append((time.perf_counter() - t0) * 1e3)
        rs.close()
    ms = float(np.mean(times))
    return {"ms_per_step": ms, "value": ms * 1e6 / keys.shape[0], "unit": "ns/object", "steps": steps,
            "upload_and_host_ms": ms - device_ms, "key_bytes": int(keys.nbytes)}


def plan_graph(gra, ctx, splitmix64_keys_device, reps: int = 20) -> dict:
    """Small builds three ways, median wall ms of `reps` each with the serialized
    bytes in host memory at the end: the ordinary build (the host reads sizes back
    between phases), a build plan (rs_plan_enqueue: one queued build, one
    synchronisation) and a replay of the plan captured in a graph.  All three
    give the same bytes.  Each entry also gives the ordinary build's search ms and
    launch waves per kind (levels of little work run on fewer waves)."""
    import torch
    out = {}
    for n, leaf, b in [(10_000, 8, 100), (10_000, 12, 100), (100_000, 12, 100), (10_000, 16, 2000),
                       (100_000, 16, 2000), (1_000_000, 8, 100)]:
        keys = splitmix64_keys_device(20 + leaf, n)
        torch.cuda.synchronize()

        def timed(fn):
            ts, data = [], None
            for _ in range(reps):
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                data = fn()
                ts.append((time.perf_counter() - t0) * 1e3)
            return float(np.median(ts)), data

        def eager():
            rs = gra.GPURecSplit(keys, b, leaf_size=leaf, ctx=ctx)
            d = rs.serialize()
            rs.close()
            return d

        eager()
        e_ms, ref = timed(eager)
        st = gra.GPURecSplit(keys, b, leaf_size=leaf, ctx=ctx).stats()
        plan = gra.Plan(n, b, leaf, ctx=ctx)
        plan.build(keys)
        p_ms, p_data = timed(lambda: plan.build(keys))
        static = keys.clone()
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            plan.enqueue(static)

        def replay():
            g.replay()
            torch.cuda.synchronize()
            return plan.result()

        replay()
        g_ms, g_data = timed(replay)
        out[f"n{n}_l{leaf}_b{b}"] = {"build_ms": e_ms, "plan_ms": p_ms, "graph_replay_ms": g_ms,
                                     "same_bytes": p_data == ref and g_data == ref,
                                     "plan_device_bytes": plan.device_bytes,
                                     "search_ms": {k: v for k, v in st["search_ms"].items() if v > 0},
                                     "kind_waves": {k: v for k, v in st["kind_waves"].items() if v > 0}}
        log(f"plan/graph n={n} l={leaf} b={b}: build {e_ms:.2f} ms, plan {p_ms:.2f} ms, graph {g_ms:.2f} ms")
        del g
        plan.close()
    return out


def level_tails(gra, d_keys, leaf: int, bucket: int) -> dict:
    """One more build of the headline keys (outside the timed steps) on a context
    created with RS_WAVE_TIMES=1: per kind, the share of the resident search
    waves' time spent in the search loop, and the time from the median wave's exit
    to the last exit.  It bounds what running tree levels without a launch
    boundary between them could recover."""
    os.environ["RS_WAVE_TIMES"] = "1"
    try:
        diag = gra.Context(0)
    finally:
        del os.environ["RS_WAVE_TIMES"]
    rs = gra.GPURecSplit(d_keys, bucket, leaf_size=leaf, ctx=diag)
    st = rs.stats()
    rs.close()
    diag.close()
    kinds = [k for k in KINDS if st["search_ms"][k] > 0]
    return {"busy": {k: st["search_busy"][k] for k in kinds}, "tail_ms": {k: st["search_tail_ms"][k] for k in kinds},
            "search_ms": {k: st["search_ms"][k] for k in kinds}}


def served(args) -> dict:
    """Scaled served stream through the construction server: Poisson arrivals
    (seed 4), sizes log-uniform in [1e4, 1e6], l in {8, 12, 16, 24} with weights
    {.3, .3, .3, .1}, b = 2000 for l >= 16 else 100, l = 24 jobs capped at 4000 keys
    (a 24-leaf key costs ~3e8 evaluations)."""
    from gpurecsplit_amd.serve import poisson_stream, run_stream

    stream = poisson_stream(args.serve_jobs, rate_per_s=args.serve_rate, seed=4, size_lo=10_000,
                            size_hi=1_000_000, leaf_weights={8: 0.3, 12: 0.3, 16: 0.3, 24: 0.1},
                            bucket_of_leaf=lambda l: 2000 if l >= 16 else 100,
                            size_cap_of_leaf=lambda l: 4000 if l == 24 else 10**9)
    log(f"served: {len(stream)} jobs at {args.serve_rate}/s")
    done = [0]

    def on_done(j):
        done[0] += 1
        if done[0] % 20 == 0:
            log(f"served: {done[0]}/{len(stream)} jobs collected")

    jobs, summ = run_stream(stream, slots=4, time_limit_s=args.serve_time_limit, on_done=on_done)
    log(f"served: done in {summ['wall_s']:.1f} s")
    errors = [j.error for j in jobs if j.error]
    ok = not errors
    for j in jobs[:: max(1, len(jobs) // 20)]:  # bijection of a sample of results
        if j.result is not None:
            q = j.result.query(j.keys)
            ok &= bool(np.array_equal(np.sort(q), np.arange(j.n, dtype=np.uint64)))
    by_leaf = {}
    for j in jobs:
        e = by_leaf.setdefault(j.leaf, {"jobs": 0, "keys": 0, "lat": []})
        e["jobs"] += 1
        e["keys"] += j.n
        e["lat"].append(j.latency)
    summ.update({"stream": f"scaled: {len(stream)} jobs, Poisson {args.serve_rate}/s, sizes log-uniform [1e4, 1e6], "
                           "l=24 capped at 4000 keys, 4 slots",
                 "verified_sample": ok, "errors": errors[:3],
                 "by_leaf": {l: {"jobs": e["jobs"], "keys": e["keys"], "p50_latency_s": float(np.percentile(e["lat"], 50)),
                                 "p99_latency_s": float(np.percentile(e["lat"], 99))} for l, e in sorted(by_leaf.items())}})
    return summ


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=5_000_000)
    ap.add_argument("--leaf", type=int, default=16)
    ap.add_argument("--bucket", type=int, default=2000)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--quick", action="store_true", help="headline line only (plus C3): no batched/served/CPU legs")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--probe", action="store_true", help="also run the VALU issue-rate and copy probes")
    ap.add_argument("--serve-jobs", type=int, default=160, help="jobs of the served stream (0: skip)")
    ap.add_argument("--serve-rate", type=float, default=16.0, help="job arrivals per second of the served stream")
    ap.add_argument("--serve-time-limit", type=float, default=120.0, help="per-batch build time limit (s)")
    ap.add_argument("--no-4096", action="store_true", help="skip the 4096-set batched workload")
    ap.add_argument("--no-sequential", action="store_true",
                    help="batched4096: skip the 4096 one-by-one builds (profiling runs)")
    ap.add_argument("--only", choices=["l18", "l24", "batched4x5M", "batched4096", "served", "plan_graph"],
                    help="run one extra workload alone and print its JSON (profiling runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(buckets would shard by contiguous ranges across GPUs)", file=sys.stderr)
        return 2

    import torch
    import gpurecsplit_amd as gra
    from gpurecsplit_amd.keys import splitmix64_keys_device

    if not torch.cuda.is_available():
        print("no GPU visible", file=sys.stderr)
        return 1
    ctx = gra.Context(0)
    if args.only:
        if args.only == "l18":
            d = splitmix64_keys_device(2, 5_000_000)
            out = plain(gra, ctx, d, 5_000_000, 18, 50, args.steps, args.warmup, "l18 b50")
        elif args.only == "l24":
            d = splitmix64_keys_device(24, 40_000)
            out = plain(gra, ctx, d, 40_000, 24, 2000, args.steps, args.warmup, "l24 b2000")
        elif args.only == "batched4x5M":
            out = batched_4x5m(gra, ctx, args, float("nan"))
        elif args.only == "batched4096":
            out = batched_4096(gra, ctx, quick=False, sequential=not args.no_sequential)
        elif args.only == "plan_graph":
            out = plan_graph(gra, ctx, splitmix64_keys_device)
        else:
            out = served(args)
        print(json.dumps({"only": args.only, **out}))
        return 0
    d_keys = splitmix64_keys_device(args.seed, args.n)
    torch.cuda.synchronize()
    log(f"keys resident: n={args.n} l={args.leaf} b={args.bucket}")
    head = plain(gra, ctx, d_keys, args.n, args.leaf, args.bucket, args.steps, args.warmup, "plain")
    kern = head["kernels"]
    dom = max(kern, key=lambda k: kern[k]["kernel_ms"])
    line = {
        "metric": "construction ns/object at l=16,b=2000,n=5M (batched vs plain); p99 job latency",
        "value": head["value"],
        "unit": "ns/object",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32/u64",
        "data": "synthetic (splitmix64 128-bit key hashes)",
        "config": {"workload": f"plain GPURecSplit<{args.leaf}> build, n={args.n}, b={args.bucket}, "
                               f"rotation on, splitmix64 seed {args.seed}", "mode": "plain",
                   "leaf": args.leaf, "bucket": args.bucket, "n": args.n},
        "roofline": {"bound": "valu", "kernel": f"k_search<{dom}>", "achieved": kern[dom]["achieved"],
                     "peak": VALU_PEAK_TOPS, "unit": "TOP/s", "frac": kern[dom]["frac"],
                     "traffic": TRAFFIC_BYTES.get(dom), "ops_per_eval": OPS_PER_EVAL,
                     "evals": kern[dom]["evals"], "kernel_ms": kern[dom]["kernel_ms"],
                     "evals_per_s": kern[dom]["evals_per_s"], "kernels": kern},
        "evals_per_s": head["evals_per_s"],
        "seeds_tested_per_s": head["seeds_tested_per_s"],
        "bits_per_key": head["bits_per_key"],
        "bijection_verified": head["bijection_verified"],
        "phase_ms": head["phase_ms"],
    }
    if not args.quick:
        line["level_tails"] = level_tails(gra, d_keys, args.leaf, args.bucket)
        line["dropin_host_keys"] = dropin_host_keys(gra, ctx, d_keys, args.leaf, args.bucket, head["ms_per_step"])
    del d_keys
    ok = head["bijection_verified"]
    d3 = splitmix64_keys_device(2, 5_000_000)
    c3 = plain(gra, ctx, d3, 5_000_000, 18, 50, max(1, min(args.steps, 3)), 1, "l18 b50")
    line["plain_l18_b50_n5M"] = c3
    ok &= c3["bijection_verified"]
    del d3
    if not args.quick:
        d24 = splitmix64_keys_device(24, 40_000)
        l24 = plain(gra, ctx, d24, 40_000, 24, 2000, 1, 0, "l24 b2000")
        l24["note"] = "scaled: 40,000 keys (20 buckets), seed 24"
        line["plain_l24_scaled"] = l24
        ok &= l24["bijection_verified"]
        del d24
        line["plan_graph"] = plan_graph(gra, ctx, splitmix64_keys_device)
        line["batched_4x5M"] = batched_4x5m(gra, ctx, args, head["ms_per_step"])
        if not args.no_4096:
            line["batched_4096"] = batched_4096(gra, ctx, quick=False)
        torch.cuda.empty_cache()
        if args.serve_jobs > 0:
            line["served"] = served(args)
            half = argparse.Namespace(**vars(args))
            half.serve_rate = args.serve_rate / 2
            line["served_half_rate"] = served(half)
    if args.probe:
        line["probe"] = probe(0)
    if not args.no_cpu_baseline and not args.quick:
        threads = max(1, min(16, len(os.sched_getaffinity(0))))
        line["cpu_baseline"] = cpu_baseline(args.leaf, args.bucket, args.seed, threads)
        line["cpu_baseline_1thread"] = cpu_baseline(args.leaf, args.bucket, args.seed, 1)
    print(json.dumps(line))
    return 0 if ok else 1


if __name__ == "__main__":
    sys.exit(main())
