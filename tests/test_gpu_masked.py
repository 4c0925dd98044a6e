"""Masked renders (fd_kernel::render_with_masked / render_bla_masked,
crates/fd-kernel/src/grid.rs:173-209) and supersampled grids on the GPU against the oracle.

A masked render computes only the samples whose mask entry is set and leaves the others
unset; the jet-AA film passes (crates/fd-cli/src/jetaa.rs:1-45) render each pixel's
top-left sample first and a rejected subset second.  Columns are compared bit for bit
over the whole grid (unset samples are zero in both), raw outcomes and work counters over
the computed samples.
"""
import numpy as np
import pytest

import fractodactyl_amd as fa
import oracle
from fdsfile import kern05, parse_fds
from parity import assert_same

pytestmark = pytest.mark.gpu

COLS = fa.COL_NU | fa.COL_DE | fa.COL_NORMAL
STATS = ("iterations", "equivalent", "blocks", "skipped", "fallback_samples",
         "closed_form_samples")


def frame(name):
    h, _ = parse_fds(dict(kern05())[name])
    return h, fa.View(h["re"], h["im"], h["width"], h["rotation"])


def first_samples(nx, ny):
    """jetaa.rs `first`: each 2x2 pixel's top-left sample of an ss 2 grid."""
    i = np.arange(nx * ny)
    return (((i % nx) % 2 == 0) & ((i // nx) % 2 == 0)).astype(np.uint8)


def check(g, c, mask):
    _, gs, gst = g
    for k in ("class", "nu", "de", "normal"):
        assert_same(gs[k], c["columns"][k], k)
    on = mask != 0
    r, o = gs["raw"], c["outcomes"]
    assert np.all(r["kind"][~on] == 255)
    assert_same(r["kind"][on], o["kind"][on], "kind")
    esc = on & (o["kind"] == 0)
    for k in ("zr", "zi", "dr", "di", "dexp"):
        assert_same(r[k][esc], o[k][esc], k)
    for k in STATS:
        assert gst[k] == c["stats"][k], (k, gst[k], c["stats"][k])


@pytest.mark.parametrize("name,route", [("seahorse-fx", fa.ROUTE_PLAIN),
                                        ("valley-1e-28", fa.ROUTE_BLA),
                                        ("i-1e-1000", fa.ROUTE_BLA)])
def test_masked_render_equals_oracle(name, route):
    h, v = frame(name)
    nx, ny = h["nx"], h["ny"]
    rng = np.random.default_rng(7)
    for mask in (first_samples(nx, ny), (rng.random(nx * ny) < 0.3).astype(np.uint8)):
        p = fa.Params(nx, ny, 1, h["max_iter"], h["escape_radius"], COLS)
        g = fa.render_masked(v, p, mask, route, raw=True)
        c = oracle.render(h["re"], h["im"], h["width"], h["rotation"], nx, ny, h["max_iter"],
                          h["escape_radius"], columns=COLS, bla=route == fa.ROUTE_BLA,
                          mask=mask)
        check(g, c, mask)


def test_two_passes_on_a_resident_frame_make_the_full_render():
    h, v = frame("seahorse")
    nx, ny = h["nx"], h["ny"]
    p = fa.Params(nx, ny, 1, h["max_iter"], h["escape_radius"], COLS)
    full = fa.render(v, p, fa.ROUTE_BLA, raw=True)
    first = first_samples(nx, ny)
    with fa.ResidentFrame(v, p, fa.ROUTE_BLA) as f:
        f.mask(first)
        s1 = f.run()
        _, c1, _ = f.read(raw=True)
        f.mask(1 - first)
        s2 = f.run()
        _, c2, _ = f.read(raw=True)
        f.mask(None)
        f.run()
        _, c3, _ = f.read(raw=True)
    on = first != 0
    for k in ("class", "nu", "de", "normal"):
        merged = np.where(on, c1[k], c2[k])
        assert_same(merged, full[1][k], k)
        assert_same(c3[k], full[1][k], k + " (mask cleared)")
    assert s1["equivalent"] + s2["equivalent"] == full[2]["equivalent"]
    # later passes reuse the build flags the first pass's set-up cleared
    assert s1["bla_levels"] == s2["bla_levels"] == full[2]["bla_levels"]


def test_mixed_form_with_a_mask():
    h, v = frame("rung-7676")
    nx, ny = h["nx"], h["ny"]
    p = fa.Params(nx, ny, 1, h["max_iter"], h["escape_radius"], COLS)
    mask = first_samples(nx, ny)
    a = fa.render_masked(v, p, mask, fa.ROUTE_BLA, fa.FORM_F64, raw=True)
    m = fa.render_masked(v, p, mask, fa.ROUTE_BLA, fa.FORM_MIXED, raw=True)
    assert np.all(m[1]["raw"]["kind"][mask == 0] == 255)
    esc = (mask != 0) & (m[1]["raw"]["kind"] == 0) & (a[1]["raw"]["kind"] == 0)
    assert esc.sum() > 0
    for k in ("n", "zr", "zi", "dr", "di", "dexp"):
        assert_same(m[1]["raw"][k][esc], a[1]["raw"][k][esc], k)


@pytest.mark.parametrize("name,route", [("seahorse", fa.ROUTE_BLA),
                                        ("i-1e-300", fa.ROUTE_BLA)])
def test_supersampled_grid_equals_oracle(name, route):
    # ss 2: the grid holds 2x2 samples per output pixel; de is in output pixels
    h, v = frame(name)
    nx, ny = 2 * (h["nx"] // 2), 2 * (h["ny"] // 2)
    p = fa.Params(nx, ny, 2, h["max_iter"], h["escape_radius"], COLS)
    g = fa.render(v, p, route, raw=True)
    c = oracle.render(h["re"], h["im"], h["width"], h["rotation"], nx, ny, h["max_iter"],
                      h["escape_radius"], columns=COLS, ss=2, bla=route == fa.ROUTE_BLA)
    check(g, c, np.ones(nx * ny, np.uint8))
