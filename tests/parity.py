"""Bit-pattern comparison of GPU results against the oracle's."""
import numpy as np


def bits(a):
    a = np.ascontiguousarray(a)
    return a.view({1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}[a.dtype.itemsize])


def assert_same(a, b, what):
    ba, bb = bits(a), bits(b)
    bad = np.nonzero(ba != bb)[0]
    assert bad.size == 0, f"{what}: {bad.size} differ, first {bad[:5].tolist()}"


STAT_KEYS = ("iterations", "equivalent", "blocks", "skipped", "fallback_samples",
             "closed_form_samples")


def compare_frames(gpu, cpu, raw=True, stat_keys=STAT_KEYS):
    gh, gs, gst = gpu
    for k in ("class", "nu", "de", "normal"):
        if k in cpu["columns"]:
            assert_same(gs[k], cpu["columns"][k], k)
    if raw:
        o = cpu["outcomes"]
        r = gs["raw"]
        assert_same(r["kind"], o["kind"], "kind")
        esc = o["kind"] == 0
        inter = o["kind"] == 1
        assert_same(r["n"][esc | inter], o["n"][esc | inter], "n")
        for k in ("zr", "zi", "dr", "di", "dexp"):
            assert_same(r[k][esc], o[k][esc], k)
    st = cpu["stats"]
    for k in stat_keys:
        assert gst[k] == st[k], (k, gst[k], st[k])
    assert gst["reference_len"] == st["reference_len"]
