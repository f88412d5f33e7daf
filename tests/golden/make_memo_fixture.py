"""Writes tests/golden/golomb_memo.json from the reference's committed memo table.

The reference ships its Golomb-Rice memo (golomb parameter << 27 | subtree nodes
<< 16 | subtree fixed bits) as a C++ table in include/function/golombRiceMemo.hpp.
This script reads that file AS TEXT and stores, per leaf size 2..24, only an
FNV-1a 64-bit checksum of the 3000 little-endian u32 entries plus a handful of
sample entries (m = 0..8 and a few larger sizes).  The oracle and the product each
regenerate the table from the formula and are checked against this fixture.

Run (in a container that has the reference checkout):
    python tests/golden/make_memo_fixture.py /root/reference
"""
import json
import re
import struct
import sys
from pathlib import Path

SAMPLES = [0, 1, 2, 3, 4, 5, 6, 7, 8, 100, 500, 1000, 2000, 2999]


def fnv1a64(data: bytes) -> int:
    h = 0xCBF29CE484222325
    for b in data:
        h ^= b
        h = (h * 0x100000001B3) & 0xFFFFFFFFFFFFFFFF
    return h


def main(ref_root: str) -> None:
    text = (Path(ref_root) / "include/function/golombRiceMemo.hpp").read_text()
    body = text[text.index("golomb_rice_memo{{") :]
    rows = re.findall(r"\{([^{}]*)\}", body)
    tables = []
    for row in rows:
        vals = [int(v, 16) for v in re.findall(r"0x[0-9A-Fa-f]+", row)]
        if len(vals) == 3000:
            tables.append(vals)
    assert len(tables) == 23, len(tables)
    out = {}
    for i, t in enumerate(tables):
        leaf = i + 2
        raw = struct.pack("<3000I", *t)
        out[str(leaf)] = {
            "fnv1a64": f"{fnv1a64(raw):016x}",
            "samples": {str(m): t[m] for m in SAMPLES},
        }
    dst = Path(__file__).with_name("golomb_memo.json")
    dst.write_text(json.dumps(out, indent=1, sort_keys=True) + "\n")
    print(f"wrote {dst}")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
