"""The fringeCount CLI end to end (the reference's process boundary,
fringeCount.cu:2265-2331): an .egr written by fsgc_gen and a motif.mo written by
fringePreprocess, the "%lld occurrences" line against the oracle's reference-compatible
u64, the "exact:" line against its exact count, and the same output with and without
the degree relabelling the CLI applies to wedge cores. Marked gpu."""
import os
import re
import subprocess

import pytest

import oracle

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "fringe-sgc_amd", "bin")

CASES = [
    # (graph spec, motif: nodes then edge pairs as fringePreprocess takes them)
    ("er:4096:32768:1", ["5", "0", "1", "1", "2", "0", "2", "0", "3", "0", "4"]),                 # edge core (C1)
    ("rmat:10:8:0.57:0.19:0.19:7", ["8", "0", "1", "1", "2", "1", "3", "1", "4", "1", "5", "0", "6", "2", "7"]),  # wedge
    ("rmat:10:8:0.57:0.19:0.19:7", ["9", "0", "1", "0", "2", "0", "3", "1", "2", "1", "3", "2", "3", "0", "4",
                                    "0", "5", "0", "6", "1", "7", "1", "8"]),                     # triangle (C4 motif)
]


def _run(args, cwd):
    p = subprocess.run(args, cwd=cwd, capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stdout + p.stderr
    return p.stdout


@pytest.mark.parametrize("spec,motif", CASES)
def test_fringecount_cli(gpu, tmp_path, spec, motif):
    print("cli", spec, motif[0], flush=True)
    egr = str(tmp_path / "g.egr")
    _run([os.path.join(BIN, "fsgc_gen"), "graph", spec, egr], tmp_path)
    _run([os.path.join(BIN, "fringePreprocess")] + motif, tmp_path)
    mo_path = str(tmp_path / "motif.mo")
    mo = open(mo_path, "rb").read()
    ni, nl = gpu.generate_csr(spec)
    o = oracle.count(mo, ni, nl)
    outs = []
    for extra in ([], ["--relabel", "none"], ["--relabel", "desc"]):
        out = _run([os.path.join(BIN, "fringeCount"), egr, mo_path] + extra, tmp_path)
        occ = int(re.search(r"^(-?\d+) occurrences$", out, re.M).group(1))
        exact = int(re.search(r"^exact: (\d+) occurrences", out, re.M).group(1))
        assert re.search(r"^runtime: \d+\.\d{6} s$", out, re.M)
        assert occ == (o["ref_compat"] if o["ref_compat"] < 2**63 else o["ref_compat"] - 2**64)  # printed signed
        assert exact == o["exact"]
        outs.append((occ, exact))
    assert outs[0] == outs[1] == outs[2]
