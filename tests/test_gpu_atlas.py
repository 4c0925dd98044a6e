"""Reusing stored orbits (docs/spec/ATLAS.md:127-170, REF-01/REF-02): an orbit written as
canonical slab chunks and read back opens a frame (fdg_frame_open_orbit) that renders
exactly as one whose orbit was computed, and a deeper frame's orbit serves a shallower
frame at the same centre exactly as the oracle's render with that supplied orbit.
"""
import numpy as np
import pytest

import fractodactyl_amd as fa
import oracle
from fdsfile import kern05, parse_fds
from parity import assert_same

pytestmark = pytest.mark.gpu

COLS = fa.COL_NU | fa.COL_DE | fa.COL_NORMAL


def store_and_load(re, im, bits, centre, size=1024):
    slabs = fa.orbit_slabs(re, im, bits, size)
    ids = [fa.chunk_id(c) for c in slabs]
    manifest = fa.orbit_manifest(centre[0], centre[1], bits, ids)
    cre, cim, b, got_ids = fa.orbit_manifest_read(manifest)
    assert (cre, cim, b) == (centre[0], centre[1], bits) and got_ids == ids
    by_id = {fa.chunk_id(c): c for c in slabs}
    parts = [fa.orbit_slab_read(by_id[i]) for i in got_ids]
    return (np.concatenate([p[3] for p in parts]), np.concatenate([p[4] for p in parts]), b)


@pytest.mark.parametrize("name,route", [("valley-1e-28", fa.ROUTE_BLA), ("seahorse", fa.ROUTE_PLAIN),
                                        ("i-1e-300", fa.ROUTE_BLA)])
def test_a_stored_orbit_renders_like_a_computed_one(name, route):
    h, _ = parse_fds(dict(kern05())[name])
    v = fa.View(h["re"], h["im"], h["width"], h["rotation"])
    p = fa.Params(h["nx"], h["ny"], 1, h["max_iter"], h["escape_radius"], COLS)
    re, im, bits = fa.reference(v, p)
    orbit = store_and_load(re, im, bits, (h["re"], h["im"]))
    assert np.array_equal(orbit[0], re) and np.array_equal(orbit[1], im)
    with fa.ResidentFrame(v, p, route) as f:
        f.run()
        h1, c1, _ = f.read(raw=True)
    with fa.ResidentFrame(v, p, route, orbit=orbit) as f:
        f.run()
        h2, c2, _ = f.read(raw=True)
    assert h1["kernel"] == h2["kernel"]
    for k in ("class", "nu", "de", "normal"):
        assert_same(c2[k], c1[k], k)


def test_a_deeper_orbit_serves_a_shallower_frame():
    # grid.rs:95-109: one deep orbit serves every shallower frame at its centre; the header
    # records the orbit's precision
    h, _ = parse_fds(dict(kern05())["valley-1e-28"])
    deep_v = fa.View(h["re"], h["im"], "1e-28", 0.0)
    p = fa.Params(128, 72, 1, 30_000, 1e10, COLS)
    re, im, bits = fa.reference(deep_v, p)
    orbit = store_and_load(re, im, bits, (h["re"], h["im"]))
    v = fa.View(h["re"], h["im"], "1e-22", 0.0)
    with fa.ResidentFrame(v, p, fa.ROUTE_BLA, orbit=orbit) as f:
        f.run()
        hd, cols, _ = f.read(raw=True)
    want = oracle.render(h["re"], h["im"], "1e-22", 0.0, 128, 72, 30_000, columns=COLS,
                         bla=True, orbit=(re, im, bits))
    assert hd["kernel"].startswith(want["kernel"]), (hd["kernel"], want["kernel"])
    assert f"bits={bits}" in hd["kernel"]
    for k in ("class", "nu", "de", "normal"):
        assert_same(cols[k], want["columns"][k], k)


def test_supplied_orbit_refusals():
    h, _ = parse_fds(dict(kern05())["valley-1e-28"])
    p = fa.Params(64, 36, 1, 5000, 1e10, COLS)
    deep = fa.View(h["re"], h["im"], "1e-28", 0.0)
    re, im, bits = fa.reference(deep, p)
    shallow = fa.View(h["re"], h["im"], "1e-5", 0.0)  # the f64 tier: exactly 53 bits
    with pytest.raises(fa.FdgError, match="53"):
        fa.ResidentFrame(shallow, p, fa.ROUTE_PLAIN, orbit=(re, im, bits))
    deeper = fa.View(h["re"], h["im"], "1e-40", 0.0)
    with pytest.raises(fa.FdgError, match="at least"):
        fa.ResidentFrame(deeper, p, fa.ROUTE_PLAIN, orbit=(re, im, bits))
    with pytest.raises(fa.FdgError, match="Z_0 = 0"):
        fa.ResidentFrame(deep, p, fa.ROUTE_PLAIN, orbit=(re[1:], im[1:], bits))


def test_orbits_shorter_than_the_loop_reads_are_refused():
    # Z_1 is read from the first step, and the scaled tier's rebase from a delta below
    # 1e-270 continues to Z_2: 1 point (2 on the scaled tier) is refused before any launch,
    # where the reference would index past its orbit (sample.rs:367)
    for name, short in (("valley-1e-28", 1), ("i-1e-300", 2)):
        h, _ = parse_fds(dict(kern05())[name])
        v = fa.View(h["re"], h["im"], h["width"], h["rotation"])
        p = fa.Params(h["nx"], h["ny"], 1, h["max_iter"], h["escape_radius"], COLS)
        re, im, bits = fa.reference(v, p)
        with pytest.raises(fa.FdgError, match="at least 2|too short") as e:
            fa.ResidentFrame(v, p, fa.ROUTE_PLAIN, orbit=(re[:short], im[:short], bits))
        assert e.value.code == -1
