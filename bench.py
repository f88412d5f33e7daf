#!/usr/bin/env python
"""Construction benchmark: RecSplit MPHF of n = 5,000,000 keys, l = 16, b = 2000.

One step = one complete GPURecSplit build (bucket partition, every seed search,
Golomb-Rice + Elias-Fano encoding, copy of the structure to the host) of the
5M splitmix64(seed 1) keys, which are resident in HBM before timing starts.
Prints one JSON line (metric ns/object, lower is better) with:
  roofline      dominant search kernel against the int32 VALU issue peak
  cpu_baseline  the CPU oracle (SimdRecSplit algorithm, vectorised C, all
                threads used) on a bounded sample of the same workload

    python bench.py [--gpus 1] [--steps 3] [--warmup 1]
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

# Roofline of the seed search: int32 VALU issue.  Peak = MI355X's published FP32
# vector peak (157.3 TFLOP/s) as lane operations: one lane-op per FMA.
VALU_PEAK_TOPS = 157.3 / 2
# VALU instructions per key-seed evaluation in the inner loop of each search
# kernel (default launch: 4 candidates per lane, 1 for fanout-2 nodes; LDS
# increment table), counted from the gfx950 assembly of librecsplit_amd.so, and
# the same weighted by the issue cost measured by rs_probe.hip in v_xor_b32 slots
# (v_mul_lo_u32 1.65, v_lshl_add_u64 1.72, v_mul_u32_u24 1.54, v_xor_b32_sdwa 1.56,
# v_mad_u32_u24 1.56, v_lshlrev_b32_sdwa 1.49, other ops 1).
INSTR_PER_EVAL = {"fanout2": 11.5, "upper_aggr": 10.88, "lower_aggr": 10.88, "leaf_rotation": 10.94,
                  "leaf_bruteforce": 10.88}
SLOTS_PER_EVAL = {"fanout2": 14.82, "upper_aggr": 15.11, "lower_aggr": 15.11, "leaf_rotation": 15.20,
                  "leaf_bruteforce": 15.13}
XOR_RATE = 5.50e13  # v_xor_b32 lane-ops/s measured by rs_probe.hip (5.38e13-5.75e13 over runs)

# HBM bytes per launch of each search kernel at this workload, from two rocprofv3
# --pmc runs (FETCH_SIZE x2.0 + WRITE_SIZE x1.0, the factors of the 256 MiB copy
# probe); profiles/k_search_plain_l16_b2000_n5M.md.
TRAFFIC_BYTES = {"lower_aggr": 8.095e9, "upper_aggr": 2.083e9, "leaf_rotation": 4.820e9}


def cpu_baseline(leaf: int, bucket: int, seed: int) -> dict:
    """Oracle (vectorised C restatement of SimdRecSplit) on a bounded sample:
    5 buckets' worth of keys per thread of the same generator."""
    sys.path.insert(0, str(ROOT / "tests"))
    from oracle_lib import OracleMPHF, splitmix_keys

    threads = max(1, min(16, len(os.sched_getaffinity(0))))
    n_s = bucket * 5 * threads
    keys = splitmix_keys(seed, n_s)
    t0 = time.perf_counter()
    rs = OracleMPHF.build(keys, leaf, bucket, rotate=True, threads=threads)
    dt = time.perf_counter() - t0
    assert rs.bijection_errors(keys) == 0
    return {"value": dt / n_s * 1e9, "unit": "ns/object", "cores": threads, "kind": "port",
            "sample": f"first {n_s} keys of splitmix64(seed {seed}) ({n_s // bucket} buckets of {bucket}), "
                      f"l={leaf}, rotation on, {threads} threads, wall {dt:.1f} s"}


def probe(device: int = 0) -> dict:
    """VALU issue rates (lane-ops/s) of the instructions of the seed-search inner
    loop, from rs_probe.hip, plus one calibration copy of 256 MiB."""
    import ctypes as C

    lib = C.CDLL(str(ROOT / "gpurecsplit_amd" / "librs_probe.so"))
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


def batched(gra, ctx, args, plain_ms: float) -> dict:
    """B sets of n keys each (splitmix64 seeds seed..seed+B-1) built in one batch."""
    import torch
    from gpurecsplit_amd.keys import splitmix64_keys

    B = args.batch_sets
    print(f"# batched: {B} sets of {args.n} keys", file=sys.stderr, flush=True)
    keys = np.concatenate([splitmix64_keys(args.seed + i, args.n) for i in range(B)])
    off = np.arange(B + 1, dtype=np.uint64) * np.uint64(args.n)
    d = torch.from_numpy(keys.view(np.int64)).to("cuda:0")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    rss, st = gra.build_batch(d, off, args.bucket, leaf_size=args.leaf, ctx=ctx)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3
    print(f"# batched: {ms:.1f} ms", file=sys.stderr, flush=True)
    # the first set is the plain workload's key set: identical output
    same = rss[0].serialize() == gra.GPURecSplit(d[: args.n], args.bucket, leaf_size=args.leaf, ctx=ctx).serialize()
    return {"sets": B, "keys_per_set": args.n, "ms": ms, "value": ms * 1e6 / (B * args.n), "unit": "ns/object",
            "objects_per_s_vs_plain": (B * args.n / ms) / (args.n / plain_ms), "first_set_equals_plain": same}


def served(args) -> dict:
    """Served stream (scaled): Poisson arrivals, log-uniform sizes in [1e4, 1e6],
    l in {8, 12, 16, 24} with weights {.3, .3, .3, .1}, b = 2000 for l >= 16 else 100;
    l = 24 jobs capped at 5000 keys (an l=24 key costs ~3e8 evaluations)."""
    from gpurecsplit_amd.serve import poisson_stream, run_stream

    stream = poisson_stream(args.serve_jobs, rate_per_s=6.0, seed=4, size_lo=10_000, size_hi=1_000_000,
                            leaf_weights={8: 0.3, 12: 0.3, 16: 0.3, 24: 0.1},
                            bucket_of_leaf=lambda l: 2000 if l >= 16 else 100,
                            size_cap_of_leaf=lambda l: 5000 if l == 24 else 10**9)
    print(f"# served: {len(stream)} jobs", file=sys.stderr, flush=True)
    done = [0]

    def on_batch(batch):
        done[0] += len(batch)
        print(f"# served: batch of {len(batch)} (l={batch[0].leaf}, {sum(j.n for j in batch)} keys) done, "
              f"{done[0]}/{len(stream)} jobs", file=sys.stderr, flush=True)

    jobs, summ = run_stream(stream, workers=2, time_limit_s=args.serve_time_limit, on_batch=on_batch)
    print(f"# served: done in {summ['wall_s']:.1f} s", file=sys.stderr, flush=True)
    errors = [j.error for j in jobs if j.error]
    ok = not errors
    for j in jobs[:: max(1, len(jobs) // 10)]:  # bijection of a sample of results
        if j.result is not None:
            q = j.result.query(j.keys)
            ok &= bool(np.array_equal(np.sort(q), np.arange(j.n, dtype=np.uint64)))
    summ.update({"stream": "scaled: rate 6/s, sizes log-uniform [1e4, 1e6], l=24 capped at 5000 keys, 2 workers",
                 "verified_sample": ok, "errors": errors[:3]})
    return summ


def batched_4096(gra, ctx) -> dict:
    """The 4096-set batched workload: per-set sizes log-uniform in [1e3, 1e6]
    (numpy default_rng(3)), keys from one SplitMix64(3) stream generated in HBM,
    l = 12, b = 100, all sets in one rs_build_batch call.  Verifies the bijection
    of 32 evenly spaced sets by host queries."""
    import torch
    from gpurecsplit_amd.keys import splitmix64_keys_device

    rng = np.random.default_rng(3)
    sizes = np.exp(rng.uniform(np.log(1_000), np.log(1_000_000), size=4096)).astype(np.int64)
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    n = int(off[-1])
    print(f"# 4096 sets: {n} keys", file=sys.stderr, flush=True)
    d = splitmix64_keys_device(3, n)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    rss, st = gra.build_batch(d, off, 100, leaf_size=12, ctx=ctx)
    ms = (time.perf_counter() - t0) * 1e3
    print(f"# 4096 sets: {ms:.1f} ms", file=sys.stderr, flush=True)
    ok = True
    for i in range(0, 4096, 128):
        k = d[int(off[i]):int(off[i + 1])].cpu().numpy().view(np.uint64)
        ok &= bool(np.array_equal(np.sort(rss[i].query(k)), np.arange(sizes[i], dtype=np.uint64)))
    return {"sets": 4096, "keys": n, "leaf": 12, "bucket": 100, "ms": ms, "value": ms * 1e6 / n,
            "unit": "ns/object", "search_ms": st["search_ms"], "bijection_sample_ok": ok}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=5_000_000)
    ap.add_argument("--leaf", type=int, default=16)
    ap.add_argument("--bucket", type=int, default=2000)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--probe", action="store_true", help="also run the VALU issue-rate and copy probes")
    ap.add_argument("--batch-sets", type=int, default=4, help="sets of n keys built together (0: skip)")
    ap.add_argument("--serve-jobs", type=int, default=120, help="jobs of the served stream (0: skip)")
    ap.add_argument("--sets4096", action="store_true", help="also run the 4096-set l=12 b=100 batched workload")
    ap.add_argument("--serve-time-limit", type=float, default=60.0, help="per-batch build time limit of the served stream (s)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(buckets shard by contiguous ranges across GPUs; see DESIGN.md)", file=sys.stderr)
        return 2

    import torch
    import gpurecsplit_amd as gra
    from gpurecsplit_amd.keys import splitmix64_keys

    if not torch.cuda.is_available():
        print("no GPU visible", file=sys.stderr)
        return 1
    keys = splitmix64_keys(args.seed, args.n)
    d_keys = torch.from_numpy(keys.view(np.int64)).to("cuda:0")
    ctx = gra.Context(0)
    print(f"# keys resident: n={args.n} l={args.leaf} b={args.bucket}", file=sys.stderr, flush=True)

    for i in range(args.warmup):
        rs = gra.GPURecSplit(d_keys, args.bucket, leaf_size=args.leaf, ctx=ctx)
        print(f"# warmup {i}: {rs.stats()['total_ms']:.1f} ms", file=sys.stderr, flush=True)
        del rs
    torch.cuda.synchronize()
    stats = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        rs = gra.GPURecSplit(d_keys, args.bucket, leaf_size=args.leaf, ctx=ctx)
        stats.append(rs.stats())
        print(f"# step {i}: {stats[-1]['total_ms']:.1f} ms", file=sys.stderr, flush=True)
        if i + 1 < args.steps:
            del rs
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps

    # verification outside the timed region: bijection through the GPU query kernel
    idx = rs.query_device(d_keys)
    ok = bool(torch.equal(torch.sort(idx).values, torch.arange(args.n, device=idx.device)))
    bits_per_key = rs.get_bits() / args.n

    # dominant search kernel: the kind with the most event-timed search time
    last = stats[-1]
    kind = max(last["search_ms"], key=lambda k: last["search_ms"][k])
    kms = np.mean([s["search_ms"][kind] for s in stats])
    evals = last["evals"][kind]
    evals_per_s = evals / (kms / 1e3)
    achieved = evals_per_s * INSTR_PER_EVAL[kind] / 1e12
    total_evals = sum(last["evals"].values())
    search_total_ms = sum(last["search_ms"].values())

    line = {
        "metric": "construction ns/object at l=16,b=2000,n=5M (batched vs plain); p99 job latency",
        "value": ms_per_step * 1e6 / args.n,
        "unit": "ns/object",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32/u64",
        "data": "synthetic (splitmix64 128-bit key hashes)",
        "config": {"workload": f"plain GPURecSplit<{args.leaf}> build, n={args.n}, b={args.bucket}, "
                               f"rotation on, splitmix64 seed {args.seed}", "mode": "plain",
                   "leaf": args.leaf, "bucket": args.bucket, "n": args.n},
        "roofline": {"bound": "valu", "kernel": f"k_search<{kind}>", "achieved": achieved,
                     "peak": VALU_PEAK_TOPS, "unit": "TOP/s", "frac": achieved / VALU_PEAK_TOPS,
                     "traffic": TRAFFIC_BYTES.get(kind), "evals": evals, "kernel_ms": kms,
                     "evals_per_s": evals_per_s, "instr_per_eval": INSTR_PER_EVAL[kind],
                     "issue_bound_frac": evals_per_s * SLOTS_PER_EVAL[kind] / XOR_RATE},
        "seeds_evals_per_s": total_evals / (search_total_ms / 1e3),
        "exec_ratio": {k: (last["executed_evals"][k] / last["evals"][k]) if last["evals"][k] else None
                       for k in last["evals"]},
        "bits_per_key": bits_per_key,
        "bijection_verified": ok,
        "phase_ms": {"partition": last["partition_ms"], "search": last["search_ms"], "split": last["split_ms"],
                     "encode": last["encode_ms"], "total": last["total_ms"]},
    }
    if args.batch_sets > 1:
        line["batched"] = batched(gra, ctx, args, ms_per_step)
    if args.sets4096:
        line["batched_4096"] = batched_4096(gra, ctx)
    if args.serve_jobs > 0:
        line["served"] = served(args)
    if args.probe:
        line["probe"] = probe(0)
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(args.leaf, args.bucket, args.seed)
    print(json.dumps(line))
    return 0 if ok else 1


if __name__ == "__main__":
    sys.exit(main())
