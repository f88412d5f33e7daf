"""Writes tests/golden/headline_mphf.json: the oracle's MPHF of each headline
workload at its stated size, as a SHA-256 of the serialized bytes plus sizes and
counters.

The GPU tests compare the HIP build's bytes with these digests, so the full-size
configurations are byte-checked without re-running minutes of CPU work on every
test run.  The digests come from the CPU oracle (oracle/liboracle.so), the C
restatement of SimdRecSplit; rerun this script after any change to the oracle.

Workloads (keys: SplitMix64(seed), key i = (next(), next())):
  plain_l16_b2000_n5M   n = 5,000,000, seed 1, l = 16, b = 2000, rotation
  plain_l18_b50_n5M     n = 5,000,000, seed 2, l = 18, b = 50, rotation
  plain_l24_b2000_n8000 n = 8,000, seed 24, l = 24, b = 2000, rotation (4 buckets:
                        6-part upper nodes, 9-part lower nodes, 24-key leaves)
  plain_l8_b100_n1M     n = 1,000,000, seed 42, l = 8, b = 100, rotation

    python tests/golden/make_headline_fixtures.py [name ...] [--threads T]
"""
import argparse
import hashlib
import json
import os
import sys
import time
from pathlib import Path

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))

from oracle_lib import OracleMPHF, splitmix_keys  # noqa: E402

WORKLOADS = {
    "plain_l16_b2000_n5M": dict(n=5_000_000, seed=1, leaf=16, bucket=2000),
    "plain_l18_b50_n5M": dict(n=5_000_000, seed=2, leaf=18, bucket=50),
    "plain_l24_b2000_n8000": dict(n=8_000, seed=24, leaf=24, bucket=2000),
    "plain_l8_b100_n1M": dict(n=1_000_000, seed=42, leaf=8, bucket=100),
}
OUT = HERE / "headline_mphf.json"


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("names", nargs="*")
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 1)
    args = ap.parse_args()
    data = json.loads(OUT.read_text()) if OUT.exists() else {}
    for name in args.names or list(WORKLOADS):
        w = WORKLOADS[name]
        keys = splitmix_keys(w["seed"], w["n"])
        t0 = time.perf_counter()
        rs = OracleMPHF.build(keys, w["leaf"], w["bucket"], rotate=True, threads=args.threads)
        dt = time.perf_counter() - t0
        blob = rs.serialize()
        errors = rs.bijection_errors(keys)
        assert errors == 0, f"{name}: oracle MPHF is not a bijection ({errors} errors)"
        data[name] = dict(w, rotate=True, sha256=hashlib.sha256(blob).hexdigest(), bytes=len(blob),
                          nseeds=rs.nseeds, evals=rs.evals, max_seed=int(rs.seeds().max()))
        print(f"{name}: {len(blob)} bytes, {rs.nseeds} seeds, {rs.evals} evals, {dt:.1f} s", flush=True)
        OUT.write_text(json.dumps(data, indent=1, sort_keys=True) + "\n")


if __name__ == "__main__":
    main()
