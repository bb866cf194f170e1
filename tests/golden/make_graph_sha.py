"""Writes graph_sha256.json: SHA-256 of the CSR (int32 little-endian nindex then
nlist) of every synthetic graph the benchmark configurations use, so a change of the
generators (splitmix64 streams, R-MAT, Chung-Lu, ER sampling) shows up as a failing
test instead of silently moving the workloads."""
import hashlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "fringe-sgc_amd"))
import fsgc  # noqa: E402

SPECS = [
    "er:4096:32768:1",                       # C1
    "rmat:20:16:0.57:0.19:0.19:7",           # C2
    "rmat:22:16:0.57:0.19:0.19:11",          # C3
    "chunglu:4000000:2.1:20:100000:19",      # C4
    "rmat:23:16:0.57:0.19:0.19:23",          # C5
    "rmat:21:32:0.57:0.19:0.19:29",          # C5
]


def digest(spec):
    ni, nl = fsgc.generate_csr(spec)
    h = hashlib.sha256()
    h.update(ni.astype("<i4").tobytes())
    h.update(nl.astype("<i4").tobytes())
    return {"nodes": int(len(ni) - 1), "directed_edges": int(len(nl)), "sha256": h.hexdigest()}


if __name__ == "__main__":
    out = {s: digest(s) for s in SPECS}
    json.dump(out, open(os.path.join(HERE, "graph_sha256.json"), "w"), indent=1)
    print(json.dumps(out, indent=1))
