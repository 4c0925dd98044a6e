"""A whole jet-AA film frame on the GPU (fdg_frame_jet: pass 1, host rebuild, pass 2 on the
rejected samples, merge; crates/fd-cli/src/film.rs:1359-1457) against the same pipeline
on the oracle, bit for bit.
"""
import numpy as np
import pytest

import fractodactyl_amd as fa
import oracle
from fdsfile import kern05, parse_fds
from parity import assert_same

pytestmark = pytest.mark.gpu

COLS = fa.COL_NU | fa.COL_DE | fa.COL_NORMAL


def oracle_jet(re_, im_, width, rotation, nx, ny, max_iter, bla):
    w, h = nx // 2, ny // 2
    first = oracle.jetaa_first(w, h)
    kw = dict(columns=COLS, ss=2, bla=bla, outcomes=False)
    s = oracle.render(re_, im_, width, rotation, nx, ny, max_iter, 1e10, mask=first, **kw)
    cols = s["columns"]
    todo = oracle.jetaa_rebuild(cols, w, h)
    if todo.any():
        s2 = oracle.render(re_, im_, width, rotation, nx, ny, max_iter, 1e10, mask=todo, **kw)
        oracle.jetaa_merge(cols, s2["columns"], todo)
    return cols, int(w * h + todo.sum())


CASES = [
    ("-0.7453", "0.1127", "0.0065", 0.0, 192, 108, 20_000, fa.ROUTE_PLAIN),
    ("-0.7453", "0.1127", "0.0065", 0.7, 192, 108, 20_000, fa.ROUTE_BLA),
]
_h, _ = parse_fds(dict(kern05())["valley-1e-28"])
CASES.append((_h["re"], _h["im"], _h["width"], 0.0, 128, 72, _h["max_iter"], fa.ROUTE_BLA))
_h, _ = parse_fds(dict(kern05())["i-1e-300"])
CASES.append((_h["re"], _h["im"], _h["width"], 0.0, 192, 108, _h["max_iter"], fa.ROUTE_BLA))


@pytest.mark.parametrize("re_,im_,width,rot,nx,ny,mi,route", CASES)
def test_jet_frame_equals_the_oracle_pipeline(re_, im_, width, rot, nx, ny, mi, route):
    v = fa.View(re_, im_, width, rot)
    p = fa.Params(nx, ny, 2, mi, 1e10, COLS)
    with fa.ResidentFrame(v, p, route) as f:
        h, cols, st, done = f.jet()
        # the frame's mask is cleared afterwards: a plain run is the exact ss 2 render
        f.run()
        _, full, _ = f.read()
    want, want_done = oracle_jet(re_, im_, width, rot, nx, ny, mi, route == fa.ROUTE_BLA)
    print(f"jet {width}: computed {done} of {nx * ny} samples", flush=True)
    assert done == want_done
    for k in ("class", "nu", "de", "normal"):
        assert_same(cols[k], want[k], k)
    exact = oracle.render(re_, im_, width, rot, nx, ny, mi, 1e10, columns=COLS, ss=2,
                          bla=route == fa.ROUTE_BLA, outcomes=False)["columns"]
    for k in ("class", "nu", "de", "normal"):
        assert_same(full[k], exact[k], k + " (exact)")
    assert done < nx * ny


def test_jet_frame_refusals():
    v = fa.View("-0.7453", "0.1127", "0.0065", 0.0)
    with fa.ResidentFrame(v, fa.Params(96, 54, 1, 2000, 1e10, COLS), fa.ROUTE_PLAIN) as f:
        with pytest.raises(fa.FdgError, match="ss 2"):
            f.jet()
    with fa.ResidentFrame(v, fa.Params(96, 54, 2, 2000, 1e10, fa.COL_NU), fa.ROUTE_PLAIN) as f:
        with pytest.raises(fa.FdgError, match="nu, de and normal"):
            f.jet()


def test_a_sample_mask_set_before_a_jet_frame_is_kept():
    # fdg_frame_jet sets its own pass masks; a mask the caller had set is set again after,
    # so the next plain run computes exactly the caller's samples
    v = fa.View("-0.7453", "0.1127", "0.0065", 0.0)
    p = fa.Params(96, 54, 2, 2000, 1e10, COLS)
    rng = np.random.default_rng(5)
    mask = (rng.random(96 * 54) < 0.3).astype(np.uint8)
    with fa.ResidentFrame(v, p, fa.ROUTE_PLAIN) as f:
        f.mask(mask)
        f.jet()
        f.run()
        _, after, _ = f.read()
    want = oracle.render("-0.7453", "0.1127", "0.0065", 0.0, 96, 54, 2000, 1e10, columns=COLS,
                         ss=2, bla=False, outcomes=False, mask=mask)["columns"]
    for k in ("class", "nu", "de", "normal"):
        assert_same(after[k], want[k], k)
