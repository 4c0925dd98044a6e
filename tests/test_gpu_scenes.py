"""GPU parity on the benchmark scenes (the seeded Newton-nucleus frames the benchmark
renders), bit for bit against the oracle: the plain route at 480x270 and 1e-20, the
BLA route at 1920x1080 and 1e-100 (every sample Interior), and the 1e-1000 scaled-tier
frame with max_iter 1e6 on a reduced grid of the same width.
"""
import numpy as np
import pytest

import fractodactyl_amd as fa
import oracle
from fractodactyl_amd.scenes import scene
from parity import compare_frames

pytestmark = pytest.mark.gpu

COLS = fa.COL_NU | fa.COL_DE | fa.COL_NORMAL


@pytest.fixture(scope="module")
def ctx():
    c = fa.Context(0)
    yield c
    c.close()


def both(ctx, k, nx=None, ny=None):
    sc = scene(k)
    cfg = sc["config"]
    nx, ny = nx or cfg["nx"], ny or cfg["ny"]
    v = fa.View(sc["center_re"], sc["center_im"], cfg["width"], 0.0)
    p = fa.Params(nx, ny, 1, cfg["max_iter"], cfg["escape_radius"], COLS)
    route = fa.ROUTE_BLA if cfg["bla"] else fa.ROUTE_PLAIN
    g = fa.render(v, p, route, ctx=ctx, raw=True)
    c = oracle.render(sc["center_re"], sc["center_im"], cfg["width"], 0.0, nx, ny,
                      cfg["max_iter"], cfg["escape_radius"], columns=COLS, bla=cfg["bla"],
                      bla_eps=fa.EPS)
    return g, c


def test_plain_scene_480x270(ctx):
    g, c = both(ctx, 1)
    compare_frames(g, c)
    assert "bla/1" not in g[0]["kernel"]
    # the nucleus centre's fixed-point orbit repeated exactly after one period
    assert g[2]["reference_period"] == scene(1)["nucleus"]["period"]
    assert g[2]["reference_repeat_at"] == g[2]["reference_period"] + 1
    assert (g[1]["class"] & 3 == 0).any() and (g[1]["class"] & 3 == 2).any()


def test_bla_scene_1080p_interior(ctx):
    g, c = both(ctx, 2)
    compare_frames(g, c)
    assert (g[1]["class"] & 3 == 1).all()
    assert g[2]["blocks"] > 0


def test_scaled_scene_1e1000_reduced_grid(ctx):
    g, c = both(ctx, 3, 96, 54)
    compare_frames(g, c)
    assert g[0]["kernel"].startswith("pert-fx-scaled/1 bits=")
    assert (g[1]["class"] & 3 == 0).any()
    assert g[2]["reference_period"] == scene(3)["nucleus"]["period"]
    assert g[2]["reference_repeat_at"] == g[2]["reference_period"] + 1
