"""A batch captured once as a HIP graph and replayed (marked gpu): every replay,
with eager counting in between, equals the eager batch and fsgc_count alone."""
import pytest

pytestmark = pytest.mark.gpu

KEYS = ("exact", "sg", "aut", "ref_compat", "core_matches", "flags", "motif_class")


def key(r):
    return tuple(getattr(r, k) for k in KEYS)


# general cores (4-6 core vertices) only on graphs where their enumeration is small
@pytest.mark.parametrize("spec,max_core", [("er:2000:8000:4", 6), ("rmat:12:8:0.57:0.19:0.19:3", 3),
                                           ("rmat:7:3:0.57:0.19:0.19:5", 6)])
def test_captured_batch_replays_equal_eager(gpu, spec, max_core):
    g = gpu.Graph.generate(spec)
    mos = [m for m in gpu.random_motifs(3, 64) if m[0] <= max_core]
    eager, _ = g.count_batch(mos)
    want = [key(r) for r in eager]
    print(spec, len(mos), "motifs", flush=True)
    cb = g.capture_batch(mos)
    for rep in range(3):
        got, ms = cb.run()
        print(f"  replay {rep}: {ms:.2f} ms", flush=True)
        assert [key(r) for r in got] == want, rep
        # eager work between replays
        one = g.count(mos[rep])
        assert key(one) == want[rep]
    cb.free()


def test_captured_batch_refuses_slices(gpu):
    g = gpu.Graph.generate("er:200:800:1")
    mo = gpu.preprocess(3, [(0, 1), (1, 2), (0, 2)])[0]
    with pytest.raises(gpu.FsgcError):
        g.capture_batch([mo], gpu.Tune(slice_mod=3, slice_rem=1))
