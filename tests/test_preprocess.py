"""The motif compiler against the reference's own fringePreprocess outputs (golden
fixtures; live against oracle/_ref/fringePreprocess when it is built). CPU only."""
import json
import os
import subprocess
import tempfile

import pytest

import fsgc
import oracle

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "mo_golden.json")))


@pytest.mark.parametrize("row", GOLDEN["motifs"], ids=lambda r: r["source"])
def test_matches_reference_preprocessor(row):
    if "number" in row:
        mo, log = fsgc.preprocess_pattern(row["number"])
    else:
        mo, log = fsgc.preprocess(row["nodes"], [tuple(e) for e in row["edges"]])
    assert fsgc.live_fields(mo) == row["live"]
    assert [l for l in log.splitlines() if l.startswith("including")] == row["including"]
    # unused slots are zero in our output
    n, a = mo[0], mo[1]
    assert all(b == 0 for b in mo[3 + max(0, n - 1):8]) and all(b == 0 for b in mo[8 + max(0, n - 1):13])
    assert all(b == 0 for b in mo[13 + a:76]) and all(b == 0 for b in mo[76 + a:139])


def test_rejects_like_reference():
    for r in GOLDEN["rejects"]:
        args = r["args"]
        n = args[0]
        e = [(args[i], args[i + 1]) for i in range(1, len(args), 2)]
        with pytest.raises(fsgc.FsgcError):
            fsgc.preprocess(n, e)
        assert r["returncode"] != 0
    with pytest.raises(fsgc.FsgcError):
        fsgc.preprocess(1, [])
    with pytest.raises(fsgc.FsgcError):
        fsgc.preprocess(3, [(0, 5)])
    with pytest.raises(fsgc.FsgcError):
        fsgc.preprocess_pattern(31)


def test_print_matching_order_text():
    mo, log = fsgc.preprocess(5, [(0, 1), (1, 2), (0, 2), (0, 3), (0, 4)])
    txt = fsgc.print_matching_order(mo)
    assert txt.startswith("starting from A >= 4\n\nfind B >= 2\n connecting to A\n\n")
    assert "anchor sets\n A: 2\n A B: 1\n" in txt
    assert txt in log


@pytest.mark.skipif(not os.path.exists(oracle.REF_PREPROCESS), reason="reference preprocessor not built")
def test_live_against_reference_binary():
    """Random patterns beyond the fixtures, against the reference binary itself."""
    for i, (n, e) in enumerate(fsgc.random_patterns(99, 80)):
        with tempfile.TemporaryDirectory() as d:
            r = subprocess.run([oracle.REF_PREPROCESS, str(n)] + [str(x) for p in e for x in p], cwd=d,
                               capture_output=True, text=True, timeout=30)
            assert r.returncode == 0
            ref = open(os.path.join(d, "motif.mo"), "rb").read()
        mine, _ = fsgc.preprocess(n, e)
        assert fsgc.live_fields(ref) == fsgc.live_fields(mine), (n, e)


def test_cli_writes_motif_files(tmp_path):
    exe = os.path.join(os.path.dirname(fsgc.LIB_PATH), "bin", "fringePreprocess")
    if not os.path.exists(exe):
        pytest.skip("CLI not built")
    r = subprocess.run([exe, "5", "0", "1", "1", "2", "0", "2", "0", "3", "0", "4"], cwd=tmp_path,
                       capture_output=True, text=True, timeout=30)
    assert r.returncode == 0, r.stderr
    mo = open(tmp_path / "motif.mo", "rb").read()
    assert len(mo) == 139
    assert fsgc.live_fields(mo) == {"nodes": 2, "anchorsets": 2, "startdeg": 4, "destdeg": [2], "list": [1],
                                    "anch": [1, 3], "fcnt": [2, 1]}
    assert (tmp_path / "motif.dot").read_text().startswith("graph {")
    r = subprocess.run([exe, "4", "0", "1", "2", "3"], cwd=tmp_path, capture_output=True, text=True, timeout=30)
    assert r.returncode != 0 and "disconnected" in r.stderr
