"""Work items run in several consecutive launches (FSGC_SEGMENT_ITEMS) give the same
results as one launch (marked gpu). The library reads the variable once per process,
so the segmented counts come from a child process."""
import json
import os
import subprocess
import sys

import pytest

from test_gpu_parity import C2_MOTIF

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SPEC = "rmat:11:8:0.57:0.19:0.19:29"

CHILD = r"""
import json, sys
sys.path.insert(0, sys.argv[1])
import fsgc
fsgc.init(0)
g = fsgc.Graph.generate(sys.argv[2])
mos = [m for m in fsgc.random_motifs(3, 32) if m[0] <= 3]
mos.append(fsgc.preprocess(8, [(0, 1), (1, 2), (1, 3), (1, 4), (1, 5), (0, 6), (2, 7)])[0])
print(json.dumps([[str(r.exact), str(r.sg), r.core_matches] for r in (g.count(m) for m in mos)]))
"""


def test_segmented_launches_equal_single(gpu):
    g = gpu.Graph.generate(SPEC)
    mos = [m for m in gpu.random_motifs(3, 32) if m[0] <= 3]
    mos.append(gpu.preprocess(*C2_MOTIF)[0])  # linear wedge: the per-centre pass runs after the last segment
    want = [[str(r.exact), str(r.sg), r.core_matches] for r in (g.count(m) for m in mos)]
    env = dict(os.environ, FSGC_SEGMENT_ITEMS="5")
    out = subprocess.run([sys.executable, "-c", CHILD, os.path.join(ROOT, "fringe-sgc_amd"), SPEC], env=env,
                         capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr[-2000:]
    got = json.loads(out.stdout.strip().splitlines()[-1])
    assert got == want
