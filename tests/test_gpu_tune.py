"""The launch tuning (fdg_frame_launch) and its autotuner (fdg_frame_tune): every candidate
launch gives the same outputs and sample statistics, bit for bit, in every delta form; the
tuner keeps a launch no slower than the default; bad launches are refused.
"""
import numpy as np
import pytest

import fractodactyl_amd as fa
from fdsfile import kern05, parse_fds
from parity import assert_same

pytestmark = pytest.mark.gpu

COLS = fa.COL_NU | fa.COL_DE | fa.COL_NORMAL
SAMPLE_STATS = ("iterations", "equivalent", "blocks", "skipped", "plain_steps",
                "fallback_samples", "closed_form_samples", "max_sample_ops")
LAUNCHES = [(0, 0, 0), (4, 2, 256), (16, 8, 4096), (1, 1, 64), (32, 64, 16384), (8, 4, 1024)]


def frame(name="i-1e-300"):
    h, _ = parse_fds(dict(kern05())[name])
    v = fa.View(h["re"], h["im"], h["width"], h["rotation"])
    return v, fa.Params(h["nx"], h["ny"], 1, h["max_iter"], h["escape_radius"], COLS)


@pytest.mark.parametrize("form", [fa.FORM_F64, fa.FORM_FE32, fa.FORM_MIXED])
def test_every_launch_gives_the_same_frame(form):
    v, p = frame()
    ref = None
    with fa.ResidentFrame(v, p, fa.ROUTE_BLA, form) as f:
        for launch in LAUNCHES:
            f.launch(*launch)
            st = f.run()
            _, s, _ = f.read(raw=True)
            if ref is None:
                ref = (s, st)
                continue
            for k in ("class", "nu", "de", "normal"):
                assert_same(s[k], ref[0][k], f"{k} {launch}")
            for k in ("kind", "n", "zr", "zi", "dr", "di", "dexp", "ops"):
                assert_same(s["raw"][k], ref[0]["raw"][k], f"{k} {launch}")
            for k in SAMPLE_STATS:
                assert st[k] == ref[1][k], (k, launch)


def test_jet_frame_and_masked_pass_under_a_tuned_launch():
    # a sparse masked pass (the jet frame's second pass) under a non-default launch
    v, p = frame("valley-1e-28")
    p = fa.Params(128, 72, 2, p.max_iter, p.escape_radius, COLS)
    with fa.ResidentFrame(v, p, fa.ROUTE_BLA) as f:
        _, a, _, done_a = f.jet()
        f.launch(16, 2, 256)
        _, b, _, done_b = f.jet()
    assert done_a == done_b
    for k in ("class", "nu", "de", "normal"):
        assert_same(a[k], b[k], k)


def test_tuner_keeps_a_launch_no_slower_than_the_default():
    v, p = frame()
    with fa.ResidentFrame(v, p, fa.ROUTE_BLA) as f:
        f.run()
        _, want, _ = f.read(raw=True)
        best, best_ms, default_ms = f.tune(reps=2)
        print("tuned", best, best_ms, default_ms, flush=True)
        assert best_ms <= default_ms
        assert best["groups_per_cu"] in (0, 4, 8, 16)
        f.run()
        _, got, _ = f.read(raw=True)
    for k in ("class", "nu", "de", "normal"):
        assert_same(got[k], want[k], k)


@pytest.mark.parametrize("launch,what", [((33, 0, 0), "groups_per_cu"),
                                         ((0, 65, 0), "claim_div"),
                                         ((0, 0, 100), "claim_cap"),
                                         ((0, 0, 16448), "claim_cap"),
                                         ((0, 0, 0, 65), "refill")])
def test_bad_launches_are_refused(launch, what):
    v, p = frame()
    with fa.ResidentFrame(v, fa.Params(96, 54, 1, 1000, 1e10, COLS), fa.ROUTE_PLAIN) as f:
        with pytest.raises(fa.FdgError, match=what):
            f.launch(*launch)
        with pytest.raises(fa.FdgError, match="reps"):
            f.tune(reps=0)
