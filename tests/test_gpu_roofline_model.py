"""The pivot kernel's algorithmic bytes (fsgc_result.alg_bytes) equal the host model
of its schedule (oracle.pivot_units / pivot_bytes) exactly: the roofline figure can be
recomputed from unit counts. Marked gpu."""
import os

import pytest

import oracle
from test_gpu_parity import C2_MOTIF

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("spec", ["rmat:12:16:0.57:0.19:0.19:7", "chunglu:20000:2.1:12:3000:9", "er:3000:20000:4"])
def test_pivot_alg_bytes_equal_model(gpu, spec):
    ni, nl = gpu.generate_csr(spec)
    g = gpu.Graph.from_csr(ni, nl)
    c2 = gpu.preprocess(*C2_MOTIF)[0]       # c3 launch (16 waves, 2048-position windows)
    p13 = gpu.preprocess_pattern(13)[0]     # closed-form wedge: no c3, LDS hub counters
    maxdeg = int(max(ni[1:] - ni[:-1]))
    s = 2 if maxdeg <= 65536 else 4
    n, m = len(ni) - 1, len(nl)
    # C2's motif is linear in the triple region: by default a launch without c3 plus the
    # per-centre pass (8 bytes per vertex and per position); with its per-wedge c3, a c3
    # launch (16 waves, 2048-position windows)
    for mo, c3, cap, l32, lin in ((c2, False, 0, False, True), (c2, True, 0, False, False),
                                  (c2, True, 64, False, False), (c2, True, 0, True, False),
                                  (p13, False, 0, False, False)):
        print(spec, "c3", c3, "lds_cap", cap, "lists32", l32, "linear", lin, flush=True)
        if l32:
            os.environ["FSGC_PIVOT_LISTS32"] = "1"
        if c3:
            os.environ["FSGC_PIVOT_C3_PER_WEDGE"] = "1"
        try:
            r = g.count(mo, gpu.Tune(lds_cap=cap))
        finally:
            os.environ.pop("FSGC_PIVOT_LISTS32", None)
            os.environ.pop("FSGC_PIVOT_C3_PER_WEDGE", None)
        u = oracle.pivot_units(ni, nl, ecap=(cap or 2048) if c3 else 2048, hub_slots=0 if c3 else 39552, c3=c3)
        want = oracle.pivot_bytes(u, 4 if l32 else s, hub_launch=not c3) + (8 * (n + m) if lin else 0)
        assert int(r.alg_bytes) == want, (spec, c3, cap, l32, lin, u)
    g.free()
