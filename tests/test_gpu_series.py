"""Series starts on the GPU (fdg_render_series / fdg_frame_series) against the oracle's
render_series (crates/fd-kernel/src/grid.rs:218-226, series.rs:342-374), bit for bit:
columns, raw outcomes and the work counters, including the retry from zero of samples
whose series start ends Unresolved (grid.rs:623-649).
"""
import numpy as np
import pytest

import fractodactyl_amd as fa
import oracle
from parity import assert_same

pytestmark = pytest.mark.gpu

V0 = ("-0.743643887037158", "0.131825904205312")
NUCLEUS = ("-0.8296874999999996981089858669738646984522311568651036619209",
           "0.1859374999999989417838556112859064221737945709381783514212")
STATS = ("iterations", "equivalent", "blocks", "skipped", "fallback_samples",
         "closed_form_samples")


def series_for(v, p):
    re, im, _ = fa.reference(v, p)
    rho = fa.bla_dc_max(v, p)
    return fa.series_build(re, im, rho, float(v.width) / p.nx * p.ss, p.max_iter)


def to_oracle(sr):
    return dict(n=sr.n, rho=sr.rho, points=sr.points, b=sr.b, state_bound=sr.state_bound,
                derivative_bound=sr.derivative_bound, shift_px=sr.shift_px)


def check(g, c, mask=None):
    gh, gs, gst = g
    for k in ("class", "nu", "de", "normal"):
        if k in gs:
            assert_same(gs[k], c["columns"][k], k)
    r, o = gs["raw"], c["outcomes"]
    on = np.ones(len(r["kind"]), bool) if mask is None else mask != 0
    assert_same(r["kind"][on], o["kind"][on], "kind")
    done = on & (o["kind"] != oracle.UNRESOLVED)
    assert_same(r["n"][done], o["n"][done], "n")
    esc = on & (o["kind"] == oracle.ESCAPED)
    for k in ("zr", "zi", "dr", "di", "dexp"):
        assert_same(r[k][esc], o[k][esc], k)
    for k in STATS:
        assert gst[k] == c["stats"][k], (k, gst[k], c["stats"][k])
    assert gh["kernel"].startswith(c["kernel"]), (gh["kernel"], c["kernel"])
    assert " series/0" in gh["kernel"]


CASES = [
    # (centre, width, nx, ny, max_iter, route, columns)
    (V0, "1e-10", 96, 54, 20_000, fa.ROUTE_PLAIN, fa.COL_NU | fa.COL_DE | fa.COL_NORMAL),
    (V0, "1e-10", 96, 54, 20_000, fa.ROUTE_PLAIN, fa.COL_NU),
    (V0, "1e-16", 128, 72, 20_000, fa.ROUTE_BLA, fa.COL_NU | fa.COL_DE | fa.COL_NORMAL),
    (V0, "1e-14", 128, 72, 20_000, fa.ROUTE_BLA, fa.COL_NU),
    (NUCLEUS, "2.3713737e-29", 33, 19, 10_000, fa.ROUTE_PLAIN, fa.COL_NU | fa.COL_DE),
    (NUCLEUS, "2.3713737e-29", 65, 37, 10_000, fa.ROUTE_BLA, fa.COL_NU | fa.COL_DE),
    (NUCLEUS, "5.0118723e-15", 192, 108, 20_000, fa.ROUTE_PLAIN,
     fa.COL_NU | fa.COL_DE | fa.COL_NORMAL),
]


@pytest.mark.parametrize("centre,width,nx,ny,mi,route,cols", CASES)
def test_series_render_equals_oracle(centre, width, nx, ny, mi, route, cols):
    v = fa.View(centre[0], centre[1], width, 0.0)
    p = fa.Params(nx, ny, 1, mi, 1e10, cols)
    sr = series_for(v, p)
    assert sr.n > 0, sr.stop
    print(f"series n={sr.n} orders={len(sr.b)} stop={sr.stop}", flush=True)
    g = fa.render_series(v, p, sr, route, raw=True)
    c = oracle.render(centre[0], centre[1], width, 0.0, nx, ny, mi, columns=cols,
                      bla=route == fa.ROUTE_BLA, series=to_oracle(sr))
    check(g, c)


def test_series_retry_restores_the_nucleus_interior():
    # the exact centre of an odd grid: its series start can only end Unresolved, and the
    # retry from zero proves it Interior through the nucleus heuristic (series.rs:645-693)
    v = fa.View(NUCLEUS[0], NUCLEUS[1], "2.3713737e-29", 0.0)
    p = fa.Params(33, 19, 1, 10_000, 1e10, fa.COL_NU | fa.COL_DE)
    sr = series_for(v, p)
    _, plain, _ = fa.render(v, p, raw=True)
    _, acc, st = fa.render_series(v, p, sr, raw=True)
    centre = 19 // 2 * 33 + 33 // 2
    assert plain["raw"]["kind"][centre] == fa.INTERIOR
    assert_same(acc["class"], plain["class"], "class")
    # the retried samples spent their unresolved attempt too
    assert st["equivalent"] > 0


def test_masked_series_passes_on_a_resident_frame():
    v = fa.View(V0[0], V0[1], "1e-16", 0.0)
    p = fa.Params(128, 72, 1, 20_000, 1e10, fa.COL_NU | fa.COL_DE | fa.COL_NORMAL)
    sr = series_for(v, p)
    mask = (np.arange(p.nx * p.ny) % 3 == 0).astype(np.uint8)
    with fa.ResidentFrame(v, p, fa.ROUTE_BLA) as f:
        f.series(sr)
        f.mask(mask)
        f.run()
        g = f.read(raw=True)
        c = oracle.render(V0[0], V0[1], "1e-16", 0.0, p.nx, p.ny, p.max_iter,
                          columns=p.columns, bla=True, series=to_oracle(sr), mask=mask)
        check(g, c, mask)
        # clearing the series gives the plain BLA render
        f.mask(None)
        f.series(None)
        f.run()
        h, cols, _ = f.read(raw=True)
        assert " series/0" not in h["kernel"]
        _, full, _ = fa.render(v, p, fa.ROUTE_BLA, raw=True)
        assert_same(cols["nu"], full["nu"], "nu")


def test_series_without_prefix_is_the_plain_render():
    v = fa.View(V0[0], V0[1], "1e-10", 0.0)
    p = fa.Params(96, 54, 1, 20_000, 1e10, fa.COL_NU | fa.COL_DE)
    sr = series_for(v, p).replace(n=0)
    h, s, st = fa.render_series(v, p, sr, raw=True)
    _, plain, pst = fa.render(v, p, raw=True)
    assert h["kernel"].endswith(" series/0 gpu-hip-f64/1"), h["kernel"]
    for k in ("class", "nu", "de"):
        assert_same(s[k], plain[k], k)
    assert st["iterations"] == pst["iterations"]


def test_series_refusals():
    v = fa.View(V0[0], V0[1], "1e-10", 0.0)
    p = fa.Params(96, 54, 1, 20_000, 1e10, fa.COL_NU)
    sr = series_for(v, p)
    with pytest.raises(fa.FdgError, match="another orbit"):
        fa.render_series(v, p, sr.replace(points=sr.points + 1))
    with pytest.raises(fa.FdgError, match="series radius"):
        fa.render_series(v, p, sr.replace(rho=sr.rho / 2))
    deep = fa.View("-1.2", "0.1", "1e-300", 0.0)
    with pytest.raises(fa.FdgError, match="f64 deltas"):
        fa.render_series(deep, fa.Params(96, 54, 1, 2000, 1e10, fa.COL_NU), sr)
    with fa.ResidentFrame(v, p, fa.ROUTE_PLAIN, fa.FORM_FE32) as f:
        with pytest.raises(fa.FdgError, match="FP64 form"):
            f.series(sr)


SERIES_CENTRE = ("-0.8296874999999996981089858669738646984522311568651036619209",
                 "0.1859374999999989417838556112859064221737945709381783514212")


@pytest.mark.parametrize("width,route", [("5.0118723e-15", fa.ROUTE_PLAIN),
                                         ("7.9432823e-26", fa.ROUTE_BLA),
                                         ("1.9952623e-20", fa.ROUTE_BLA)])
def test_prepare_renders_what_it_chose(width, route):
    # Series::prepare's counterpart: whatever it decides, the frame then renders exactly
    # the oracle's render_series (admitted) or plain/BLA render (declined)
    v = fa.View(SERIES_CENTRE[0], SERIES_CENTRE[1], width, 0.0)
    p = fa.Params(192, 108, 1, 20_000, 1e10, fa.COL_NU | fa.COL_DE | fa.COL_NORMAL)
    with fa.ResidentFrame(v, p, route) as f:
        sr, ok = f.series_prepare()
        print(f"prepare {width}: n={sr.n} admitted={ok}", flush=True)
        f.run()
        g = f.read(raw=True)
    c = oracle.render(v.center_re, v.center_im, width, 0.0, p.nx, p.ny, p.max_iter,
                      columns=p.columns, bla=route == fa.ROUTE_BLA,
                      series=to_oracle(sr) if ok else None)
    if ok:
        check(g, c)
    else:
        assert " series/0" not in g[0]["kernel"]
        for k in ("class", "nu", "de", "normal"):
            assert_same(g[1][k], c["columns"][k], k)


def test_prepare_declines_small_frames():
    v = fa.View(SERIES_CENTRE[0], SERIES_CENTRE[1], "5.0118723e-15", 0.0)
    p = fa.Params(96, 54, 1, 20_000, 1e10, fa.COL_NU)  # 5,184 < 16,384 samples
    with fa.ResidentFrame(v, p, fa.ROUTE_PLAIN) as f:
        sr, ok = f.series_prepare()
        assert not ok and sr.n == 0
