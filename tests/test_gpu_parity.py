"""GPU parity: the HIP path through the C ABI against the oracle, bit for bit.

The FP64 form restates the reference CPU kernel's exact operation order, so every
comparison here is exact equality of the IEEE bit patterns: BLA tables, raw outcomes,
encoded columns and work counters.  Cases are the reference's own golden frames
(KERN-05), its mpmath truth pack and its BLA unit-test orbits.
"""
import math

import numpy as np
import pytest

import fractodactyl_amd as fa
import oracle
import truthpack
from fdsfile import kern05, parse_fds
from parity import assert_same, bits, compare_frames

pytestmark = pytest.mark.gpu

EPS_MAX = 1.0 / (1 << 41)
TIER_OF = {"seahorse-fx": fa.TIER_FIXED, "seahorse-scaled": fa.TIER_SCALED}


@pytest.fixture(scope="module")
def ctx():
    c = fa.Context(0)
    yield c
    c.close()


def render_both(ctx, re_, im_, width, rot, nx, ny, iters, tier=-1, bla=False, eps=fa.EPS,
                columns=fa.COL_NU | fa.COL_DE | fa.COL_NORMAL, er=1e10):
    v = fa.View(re_, im_, width, rot)
    p = fa.Params(nx, ny, 1, iters, er, columns, tier)
    g = fa.render(v, p, fa.ROUTE_BLA if bla else fa.ROUTE_PLAIN, bla_eps=eps, ctx=ctx, raw=True)
    c = oracle.render(re_, im_, width, rot, nx, ny, iters, er, columns=columns, tier=tier,
                      bla=bla, bla_eps=eps)
    return g, c


@pytest.mark.parametrize("name", [n for n, _ in kern05()])
def test_kern05_plain_route_is_bit_exact(ctx, name):
    data = dict(kern05())[name]
    h, _ = parse_fds(data)
    print(f"case {name}", flush=True)
    g, c = render_both(ctx, h["re"], h["im"], h["width"], h["rotation"], h["nx"], h["ny"],
                       h["max_iter"], TIER_OF.get(name, -1))
    compare_frames(g, c)
    assert g[0]["kernel"] == c["kernel"] + " gpu-hip-f64/1"


@pytest.mark.parametrize("name", ["seahorse", "seahorse-fx", "seahorse-scaled", "cusp-1e-8",
                                  "valley-1e-28", "i-1e-50-rot", "tip-1e-200", "i-1e-300",
                                  "i-1e-1000", "rung-7676"])
def test_kern05_bla_route_is_bit_exact(ctx, name):
    data = dict(kern05())[name]
    h, _ = parse_fds(data)
    print(f"case {name} bla", flush=True)
    g, c = render_both(ctx, h["re"], h["im"], h["width"], h["rotation"], h["nx"], h["ny"],
                       h["max_iter"], TIER_OF.get(name, -1), bla=True)
    compare_frames(g, c)
    assert g[0]["kernel"] == c["kernel"] + " gpu-hip-f64/1"
    # the level count the device derives from the build's flags (read back beside the
    # statistics) is the reference's Bla::build level count
    assert g[2]["bla_levels"] == c["stats"]["bla_levels"]


@pytest.mark.parametrize("name", ["tip-1e-200", "i-1e-1000"])
def test_repeated_passes_on_a_resident_frame(ctx, name):
    """Passes after the first reuse the table's level flags that the previous pass's set-up
    cleared: every pass reports the oracle's level count and counters and the same
    outcomes."""
    h, _ = parse_fds(dict(kern05())[name])
    print(f"case {name} repeated", flush=True)
    v = fa.View(h["re"], h["im"], h["width"], h["rotation"])
    cols = fa.COL_NU | fa.COL_DE | fa.COL_NORMAL
    p = fa.Params(h["nx"], h["ny"], 1, h["max_iter"], h["escape_radius"], cols)
    c = oracle.render(h["re"], h["im"], h["width"], h["rotation"], h["nx"], h["ny"],
                      h["max_iter"], h["escape_radius"], columns=cols, bla=True)
    assert c["stats"]["bla_levels"] > 0
    with fa.ResidentFrame(v, p, fa.ROUTE_BLA, ctx=ctx) as f:
        for _ in range(3):
            st = f.run()
            gh, gs, _ = f.read(raw=True)
            compare_frames((gh, gs, st), c)
            assert st["bla_levels"] == c["stats"]["bla_levels"]


@pytest.mark.parametrize("cr,ci,iters,eps,dc,dc_exp", [
    (-0.743643887037158, 0.131825904205312, 5000, EPS_MAX, 1e-17, 0),
    (-0.743643887037158, 0.131825904205312, 20000, 2.0 ** -50, 1e-25, 0),
    (0.2501, 0.0, 100000, EPS_MAX, 2e-17, 0),
    (-1.7548776662466927, 0.0, 100000, 2.0 ** -50, 2e-16, 0),
    (0.0, 1.0, 100000, 2.0 ** -50, 1.3, -3330),
])
def test_gpu_bla_table_is_bit_exact(ctx, cr, ci, iters, eps, dc, dc_exp):
    zr, zi, _ = oracle.reference(repr(cr), repr(ci), "4", 0.0, 1, 1, iters, 1e10,
                                 tier=oracle.TIER_F64)
    want = oracle.bla_build(zr, zi, eps, dc, dc_exp)
    got = fa.build_bla(zr, zi, eps, dc, dc_exp, ctx=ctx)
    assert len(got) == len(want)
    for j, (a, b) in enumerate(zip(got, want)):
        assert_same(a, b, f"level {j + 1}")


def test_gpu_bla_table_on_a_deep_fixed_orbit(ctx):
    h, _ = parse_fds(dict(kern05())["rung-7676"])
    zr, zi = oracle.reference_fixed(h["re"], h["im"], 3445, 150_000)
    pl = oracle.plane(h["re"], h["im"], h["width"], 0.0, 96, 54)
    dc = pl["h_m"] * math.hypot(48.0, 27.0)
    want = oracle.bla_build(zr, zi, 2.0 ** -50, dc, pl["h_e"])
    got = fa.build_bla(zr, zi, 2.0 ** -50, dc, pl["h_e"], ctx=ctx)
    assert len(got) == len(want) > 5
    for j, (a, b) in enumerate(zip(got, want)):
        assert_same(a, b, f"level {j + 1}")


@pytest.mark.parametrize("name", truthpack.CASES)
def test_truth_pack_on_gpu(ctx, name):
    case = truthpack.load(name)
    edges = truthpack.edges()
    expected = sum(1 for e in edges if e[0] == name)
    v = fa.View(case["re"], case["im"], case["width"], case["rotation"])
    p = fa.Params(case["nx"], case["ny"], case["ss"], case["max_iter"], case["escape_radius"])
    for route in (fa.ROUTE_PLAIN, fa.ROUTE_BLA):
        _, s, _ = fa.render(v, p, route, bla_eps=EPS_MAX, ctx=ctx)
        sc = truthpack.compare(case, s["class"], s["nu"], s["de"], s["normal"], edges)
        assert sc["class_mismatches"] == sc["known_edges"] == expected, (route, sc["mismatches"])
        assert sc["scalar_failures"] == 0, (route, sc)


def test_larger_frame_with_rotation_and_bla(ctx):
    g, c = render_both(ctx, "-0.743643887037158704752191506114774",
                       "0.131825904205311970493132056385139", "3e-25", 0.7, 320, 180, 60000,
                       bla=True)
    compare_frames(g, c)
    assert g[2]["blocks"] > 0


def test_class_only_columns(ctx):
    g, c = render_both(ctx, "-0.7453", "0.1127", "0.0065", 0.0, 96, 54, 5000, columns=0)
    assert set(g[1]) == {"class", "raw"}
    assert_same(g[1]["class"], c["columns"]["class"], "class")


def test_requests_the_reference_refuses(ctx):
    v = fa.View("-0.7453", "0.1127", "0.0065")
    with pytest.raises(fa.FdgError) as e:
        fa.render(v, fa.Params(16, 8, columns=fa.COL_NU | fa.COL_BOUND), ctx=ctx)
    assert e.value.code == -2
    with pytest.raises(fa.FdgError) as e:
        fa.render_bla(v, fa.Params(16, 8), bla_eps=2.0 ** -23, ctx=ctx)
    assert e.value.code == -1
    with pytest.raises(fa.FdgError) as e:
        fa.render_bla(v, fa.Params(16, 8, escape_radius=3.0), ctx=ctx)
    assert e.value.code == -1
    with pytest.raises(fa.FdgError):
        fa.render(fa.View("0", "1", "1e-300"), fa.Params(16, 8, tier=fa.TIER_F64), ctx=ctx)


@pytest.mark.parametrize("nx,ny", [(96, 64), (100, 60), (64, 8)])
def test_tiled_and_row_ordered_grids_are_bit_exact(ctx, nx, ny):
    # A wave's 64 samples are an 8 x 8 tile of the grid when both sides are multiples of 8
    # (96 x 64, 64 x 8), else 64 consecutive samples of the row-major scan (100 x 60): both
    # orders must compute every sample once, each equal to the oracle's.
    print(f"case tiles {nx}x{ny}", flush=True)
    g, c = render_both(ctx, "-0.7453", "0.1127", "0.0065", 0.3, nx, ny, 5000, bla=True)
    compare_frames(g, c)


@pytest.mark.parametrize("refill", [0, 1, 16])
def test_masked_tiled_frame_equals_the_full_frame(ctx, refill):
    # a sparse mask on a tiled grid: the masked pass computes exactly the masked samples,
    # equal to the full frame's, whatever the refill threshold
    print(f"case masked tiles refill {refill}", flush=True)
    v = fa.View("-0.7453", "0.1127", "0.0065", 0.0)
    p = fa.Params(64, 64, 1, 5000, 1e10, fa.COL_NU | fa.COL_DE | fa.COL_NORMAL)
    _, full, _ = fa.render(v, p, fa.ROUTE_BLA, ctx=ctx)
    mask = (np.arange(64 * 64) * 7919 % 13 == 0).astype(np.uint8)
    with fa.ResidentFrame(v, p, fa.ROUTE_BLA, ctx=ctx) as f:
        f.launch(refill=refill)
        f.mask(mask)
        f.run()
        _, got, _ = f.read()
    sel = mask.astype(bool)
    for k in ("class", "nu", "de", "normal"):
        a = np.asarray(got[k]).reshape(64 * 64, -1)[sel]
        b = np.asarray(full[k]).reshape(64 * 64, -1)[sel]
        assert_same(a, b, k)
