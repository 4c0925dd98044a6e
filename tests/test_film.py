"""Films: keyframes at one centre rendered against one shared reference orbit.

The reference renders a frame from a supplied, deeper orbit with
fd_kernel::render_with / render_bla (crates/fd-kernel/src/grid.rs:95-109,190-210); `fd
reuse` (crates/fd-cli/src/reuse.rs:1-60) shares one orbit, computed at the deepest frame's
precision, across the frames at one centre and requires that "same precision must mean
bit-identical output".  CPU tests pin the oracle's supplied-orbit path to that contract
and to the reference's refusals; GPU tests check the film entry points (one batched
launch in a HIP graph, or one launch per frame; per-frame or shared BLA tables) bit for
bit against the oracle rendering each frame from the same orbit and table bound.
"""
import numpy as np
import pytest

import fractodactyl_amd as fa
import oracle
from fdsfile import kern05, parse_fds
from parity import assert_same, compare_frames

NX, NY, ITERS = 96, 54, 20_000
COLS = oracle.COL_NU | oracle.COL_DE | oracle.COL_NORMAL


@pytest.fixture(scope="module")
def centre():
    h, _ = parse_fds(dict(kern05())["rung-7676"])
    return h["re"], h["im"]


def widths(k=6):
    """1e-10 (an f64-tier frame at this grid) down to 1e-300 (scaled tier)."""
    return [f"{10.0 ** (-10 - 290 * i / (k - 1)):.3e}" for i in range(k)]


def deep_orbit(c, ws):
    """The orbit of the deepest frame, at its precision (reuse.rs:33-44)."""
    zr, zi, bits = oracle.reference(c[0], c[1], ws[-1], 0.0, NX, NY, ITERS, 1e10)
    return zr, zi, bits


def dc_bound(c, w):
    pl = oracle.plane(c[0], c[1], w, 0.0, NX, NY)
    half = np.hypot(NX / 2, NY / 2)
    if pl["tier"] == oracle.TIER_SCALED:
        return pl["h_m"] * half, pl["h_e"]
    return pl["h_m"] * 2.0 ** pl["h_e"] * half, 0


def test_supplied_orbit_at_own_precision_is_identical(centre):
    # reuse.rs:67-70: same precision must mean bit-identical output (and kernel id)
    for w, bla in (("1e-40", False), ("1e-120", True), ("1e-300", True)):
        zr, zi, bits = oracle.reference(centre[0], centre[1], w, 0.0, NX, NY, ITERS, 1e10)
        own = oracle.render(centre[0], centre[1], w, 0.0, NX, NY, ITERS, columns=COLS, bla=bla)
        sup = oracle.render(centre[0], centre[1], w, 0.0, NX, NY, ITERS, columns=COLS, bla=bla,
                            orbit=(zr, zi, bits))
        assert sup["kernel"] == own["kernel"]
        for k in own["columns"]:
            assert_same(sup["columns"][k], own["columns"][k], k)
        assert sup["stats"]["iterations"] == own["stats"]["iterations"]


def test_deeper_orbit_records_its_precision(centre):
    ws = widths()
    zr, zi, bits = deep_orbit(centre, ws)
    own = oracle.render(centre[0], centre[1], ws[1], 0.0, NX, NY, ITERS, columns=COLS, bla=True)
    sup = oracle.render(centre[0], centre[1], ws[1], 0.0, NX, NY, ITERS, columns=COLS, bla=True,
                        orbit=(zr, zi, bits))
    assert f"bits={bits} " in sup["kernel"] and f"bits={bits} " not in own["kernel"]
    # More precision may change bits (reuse.rs:67-70) but not the picture.
    assert (sup["columns"]["class"] == own["columns"]["class"]).mean() > 0.99


def test_supplied_orbit_refusals(centre):
    zr, zi, bits = oracle.reference(centre[0], centre[1], "1e-120", 0.0, NX, NY, ITERS, 1e10)
    with pytest.raises(ValueError, match="needs at least"):
        oracle.render(centre[0], centre[1], "1e-200", 0.0, NX, NY, ITERS, orbit=(zr, zi, bits))
    with pytest.raises(ValueError, match="needs 53"):
        oracle.render(centre[0], centre[1], "1e-3", 0.0, NX, NY, ITERS, orbit=(zr, zi, bits))
    with pytest.raises(ValueError, match="Z_0 = 0"):
        oracle.render(centre[0], centre[1], "1e-120", 0.0, NX, NY, ITERS,
                      orbit=(zr[1:], zi[1:], bits))
    dc, de = dc_bound(centre, "1e-120")
    with pytest.raises(ValueError, match="beyond the BLA table"):
        oracle.render(centre[0], centre[1], "1e-120", 0.0, NX, NY, ITERS, bla=True,
                      orbit=(zr, zi, bits), table_dc=(dc / 2, de))
    # a table for a larger bound is accepted (grid.rs:399-411)
    oracle.render(centre[0], centre[1], "1e-120", 0.0, NX, NY, ITERS, bla=True,
                  orbit=(zr, zi, bits), table_dc=(dc * 4, de))


def oracle_frames(c, ws, orbit, shared):
    """Each frame from the orbit its group shares (reuse.rs:124-142): f64-tier frames
    the 53-bit orbit at the rounded centre (the one the oracle computes for them), the
    others the deep orbit; shared tables at the largest dc bound of each group."""
    def f64(w):
        return oracle.plane(c[0], c[1], w, 0.0, NX, NY)["tier"] == oracle.TIER_F64

    def big(group):
        return max((dc_bound(c, w) for w in group), key=lambda t: np.log2(t[0]) + t[1])
    groups = {True: [w for w in ws if f64(w)], False: [w for w in ws if not f64(w)]}
    return [oracle.render(c[0], c[1], w, 0.0, NX, NY, ITERS, columns=COLS, bla=True,
                          orbit=None if f64(w) else orbit,
                          table_dc=big(groups[f64(w)]) if shared else None) for w in ws]


@pytest.fixture(scope="module")
def film_case(centre):
    ws = widths()
    orbit = deep_orbit(centre, ws)
    return ws, orbit, {s: oracle_frames(centre, ws, orbit, s) for s in (False, True)}


def check_film(film, cpu):
    for k, c in enumerate(cpu):
        g = film.read(k, raw=True)
        assert g[0]["kernel"] == c["kernel"] + " gpu-hip-f64/1"
        compare_frames(g, c, stat_keys=("iterations", "equivalent"))


@pytest.mark.gpu
@pytest.mark.parametrize("shared", [False, True])
def test_film_matches_oracle(centre, film_case, shared):
    ws, orbit, cpu = film_case
    views = [fa.View(centre[0], centre[1], w) for w in ws]
    tables = fa.TABLES_SHARED if shared else fa.TABLES_PER_FRAME
    with fa.Film(views, fa.Params(NX, NY, 1, ITERS, 1e10, COLS), tables=tables) as film:
        st = film.run(fa.LAUNCH_BATCH)
        assert st["reference_bits"] == orbit[2] and st["reference_len"] == len(orbit[0])
        assert st["equivalent"] == sum(c["stats"]["equivalent"] for c in cpu[shared])
        check_film(film, cpu[shared])
        # one launch per frame, then the graph replayed twice more, eager work between
        film.run(fa.LAUNCH_PER_FRAME)
        check_film(film, cpu[shared])
        for _ in range(2):
            film.run(fa.LAUNCH_BATCH)
            check_film(film, cpu[shared])


@pytest.mark.gpu
def test_film_refusals_and_centres(centre):
    p = fa.Params(16, 8, 1, 1000)
    with pytest.raises(fa.FdgError, match="share one centre"):
        fa.Film([fa.View(centre[0], centre[1], "1e-20"), fa.View("-0.75", "0.1", "1e-20")], p)
    with pytest.raises(fa.FdgError) as e:
        fa.Film([fa.View(centre[0], centre[1], "1e-20")], fa.Params(16, 8, columns=fa.COL_BOUND))
    assert e.value.code == -2
    # the same decimal written another way is the same centre (orbit.rs:228-247)
    with fa.Film([fa.View("-0.75000", "0.1", "1e-20"), fa.View("-75e-2", "0.10", "1e-30")],
                 p) as f:
        f.run()


def oracle_jet_frames(c, ws, orbit):
    """Each frame's jet-AA pipeline (film.rs:1359-1457) on the oracle, from the orbit its
    group shares, per-frame tables."""
    out, total = [], 0
    w, h = NX // 2, NY // 2
    for wd in ws:
        f64 = oracle.plane(c[0], c[1], wd, 0.0, NX, NY)["tier"] == oracle.TIER_F64
        kw = dict(columns=COLS, ss=2, bla=True, outcomes=False, orbit=None if f64 else orbit)
        first = oracle.jetaa_first(w, h)
        cols = oracle.render(c[0], c[1], wd, 0.0, NX, NY, ITERS, mask=first, **kw)["columns"]
        todo = oracle.jetaa_rebuild(cols, w, h)
        if todo.any():
            s2 = oracle.render(c[0], c[1], wd, 0.0, NX, NY, ITERS, mask=todo, **kw)["columns"]
            oracle.jetaa_merge(cols, s2, todo)
        out.append(cols)
        total += w * h + int(todo.sum())
    return out, total


@pytest.mark.gpu
def test_film_jet_matches_oracle(centre):
    # both jet-AA passes of every keyframe in one batched launch each, against each
    # frame's jet pipeline on the oracle from the same shared orbit
    ws = widths()
    orbit = deep_orbit(centre, ws)
    want, total = oracle_jet_frames(centre, ws, orbit)
    views = [fa.View(centre[0], centre[1], w) for w in ws]
    with fa.Film(views, fa.Params(NX, NY, 2, ITERS, 1e10, COLS)) as film:
        cols, done = film.jet()
        assert done == total < len(ws) * NX * NY
        for k in range(len(ws)):
            for key in ("class", "nu", "de", "normal"):
                assert_same(cols[k][key], want[k][key], f"frame {k} {key}")
        # the mask is cleared: a plain run is each frame's exact ss 2 render again
        film.run(fa.LAUNCH_BATCH)
        g = film.read(0)
        exact = oracle.render(centre[0], centre[1], ws[0], 0.0, NX, NY, ITERS, columns=COLS,
                              ss=2, bla=True, outcomes=False)["columns"]
        assert_same(g[1]["nu"], exact["nu"], "nu (mask cleared)")


@pytest.mark.gpu
def test_config4_film_all_keyframes_row_sampled():
    # the benchmark film itself: 64 keyframes of 960x540, widths 1e-10 .. 1e-300 toward the
    # period-16116 nucleus, max_iter 2e5, one shared orbit, per-frame tables, one batched
    # launch; four rows of every keyframe (different rows per frame) against the oracle
    # rendering the same rows from the same orbit, bit for bit
    from fractodactyl_amd.scenes import film_widths, scene
    sc = scene(4)
    cfg = sc["config"]
    ws = film_widths(sc)
    nx, ny, it = cfg["nx"], cfg["ny"], cfg["max_iter"]
    c = (sc["center_re"], sc["center_im"])
    zr, zi, bits = oracle.reference(c[0], c[1], ws[-1], 0.0, nx, ny, it, cfg["escape_radius"])
    views = [fa.View(c[0], c[1], w) for w in ws]
    escaped = 0
    with fa.Film(views, fa.Params(nx, ny, 1, it, cfg["escape_radius"], COLS)) as film:
        film.run(fa.LAUNCH_BATCH)
        for k, w in enumerate(ws):
            rows = [(k * 7 + r * 131) % ny for r in range(4)]
            mask = np.zeros(nx * ny, np.uint8)
            for j in rows:
                mask[j * nx:(j + 1) * nx] = 1
            f64 = oracle.plane(c[0], c[1], w, 0.0, nx, ny)["tier"] == oracle.TIER_F64
            want = oracle.render(c[0], c[1], w, 0.0, nx, ny, it, cfg["escape_radius"],
                                 columns=COLS, bla=True, mask=mask,
                                 orbit=None if f64 else (zr, zi, bits))
            g = film.read(k, raw=True)
            assert g[0]["kernel"] == want["kernel"] + " gpu-hip-f64/1"
            sel = mask.astype(bool)
            for key in ("class", "nu", "de", "normal"):
                assert_same(g[1][key][sel], want["columns"][key][sel], f"frame {k} {key}")
            o, r = want["outcomes"], g[1]["raw"]
            assert_same(r["kind"][sel], o["kind"][sel], f"frame {k} kind")
            esc = sel & (o["kind"] == 0)
            done = sel & (o["kind"] != 2)
            assert_same(r["n"][done], o["n"][done], f"frame {k} n")
            for key in ("zr", "zi", "dr", "di", "dexp"):
                assert_same(r[key][esc], o[key][esc], f"frame {k} {key}")
            escaped += int(esc.sum())
    print(f"config 4: 64 keyframes x 4 rows, {escaped} escaped samples compared", flush=True)
    assert escaped > 0
