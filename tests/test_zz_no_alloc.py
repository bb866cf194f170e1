"""fsgc_count_noalloc (the count for callers that cannot allocate) against fsgc_count,
through the C ABI: every result field identical, on two calls in a row, the second on
a smaller graph, and the refusals. Marked gpu."""
import pytest

pytestmark = pytest.mark.gpu

KEYS = ("exact", "sg", "aut", "ref_compat", "core_matches", "flags", "motif_class")


def motifs(fsgc):
    mos = [fsgc.preprocess(k + 1, [(0, i) for i in range(1, k + 1)])[0] for k in (2, 3, 5)]  # stars
    mos += [fsgc.preprocess(4, [(0, 1), (0, 2), (1, 3)])[0],                                 # edge cores
            fsgc.preprocess(6, [(0, 1), (0, 2), (0, 3), (1, 4), (1, 5)])[0],
            fsgc.preprocess(5, [(0, 1), (0, 2), (0, 3), (0, 4)])[0]]
    mos += [fsgc.preprocess_pattern(1)[0]]                                                  # host count
    for mo in mos:
        assert fsgc.plan_schedule(mo, 1000) in ("edge-pass", "host"), fsgc.live_fields(mo)
    return mos


def test_noalloc_equals_count_two_calls(gpu):
    mos = motifs(gpu)
    for spec in ("rmat:12:16:0.57:0.19:0.19:3", "er:500:2000:7"):  # the second input is smaller
        g = gpu.Graph.generate(spec)
        g.count(mos[3])  # builds the graph's supports (allocating path)
        for mo in mos:
            print(spec, gpu.live_fields(mo), flush=True)
            a = g.count(mo)
            b = g.count_noalloc(mo)
            assert tuple(getattr(b, k) for k in KEYS) == tuple(getattr(a, k) for k in KEYS), (spec, gpu.live_fields(mo))
            c = g.count_noalloc(mo)  # the array keeps the last call's contents: a repeat is identical
            assert tuple(getattr(c, k) for k in KEYS) == tuple(getattr(a, k) for k in KEYS)
        g.free()


def test_noalloc_refusals(gpu):
    g = gpu.Graph.generate("er:300:900:5")
    star = gpu.preprocess(3, [(0, 1), (0, 2)])[0]
    with pytest.raises(gpu.FsgcError) as e:
        g.count_noalloc(star)  # supports not built
    assert e.value.code == -11
    g.count(star)
    wedge = gpu.preprocess(8, [(0, 1), (1, 2), (1, 3), (1, 4), (1, 5), (0, 6), (2, 7)])[0]
    with pytest.raises(gpu.FsgcError) as e:
        g.count_noalloc(wedge)  # not an edge-pass motif
    assert e.value.code == -95
    assert g.count_noalloc(star).exact == g.count(star).exact
    g.free()
