"""The 1e-3000 jet-AA scene (3840x2160 pixels at 2x2 samples, a 10,500-bit satellite
nucleus of period 140742, max_iter 1e7) on a reduced grid at the same width and max_iter,
bit for bit against the oracle, through the jet-AA frame (crates/fd-cli/src/film.rs:
1365-1456) and the exact ss 2 pass, with the orbit at its 2^20-point cap
(crates/fd-kernel/src/reference.rs:16).

Every sample lies deep inside the nucleus's component.  The nucleus heuristic declines the
orbit (reference.rs:102-120 wants 1e-140 < |Z_P|, and Z_P is 0 in f64), so each sample
rebases onto that exact zero once a period (sample.rs:418-449): the rebase's two-term sum
is NaN (0 x 2^k with k > 1023), the step from Z_0 drops the NaN square far below dc (Rust's
saturating `as i64` gives NaN the exponent 0), and the sample restarts from dc, BLA blocks
and all.  No sample escapes and the scaled tier's periodicity never judges the returns
(the delta is below f64's range), so every sample is Unresolved at max_iter: there is no
escaped sample whose Jacobian a finite difference could check on this scene.
"""
import time

import numpy as np
import pytest

import fractodactyl_amd as fa
import oracle
from fractodactyl_amd.scenes import scene
from parity import assert_same, compare_frames

pytestmark = pytest.mark.gpu

COLS = fa.COL_NU | fa.COL_DE | fa.COL_NORMAL
NX, NY, MAX_ITER = 192, 108, 10_000_000


@pytest.fixture(scope="module")
def case():
    sc = scene(5)
    cfg = sc["config"]
    print(f"config 5 oracle orbit: {NX}x{NY} ss2, max_iter {MAX_ITER}", flush=True)
    t = time.time()
    zr, zi, bits = oracle.reference(sc["center_re"], sc["center_im"], cfg["width"], 0.0, NX,
                                    NY, MAX_ITER, cfg["escape_radius"])
    print(f"oracle orbit {len(zr)} points at {bits} bits in {time.time() - t:.1f} s",
          flush=True)
    return sc, cfg, (zr, zi, bits)


def oracle_frame(sc, cfg, orbit, **kw):
    return oracle.render(sc["center_re"], sc["center_im"], cfg["width"], 0.0, NX, NY,
                         MAX_ITER, cfg["escape_radius"], columns=COLS, ss=2, bla=True,
                         bla_eps=fa.EPS, orbit=orbit, **kw)


def test_config5_jet_and_exact_frames_equal_the_oracle(case):
    sc, cfg, orbit = case
    assert len(orbit[0]) == 1 << 20
    v = fa.View(sc["center_re"], sc["center_im"], cfg["width"], 0.0)
    p = fa.Params(NX, NY, 2, MAX_ITER, cfg["escape_radius"], COLS)
    t = time.time()
    with fa.ResidentFrame(v, p, fa.ROUTE_BLA) as f:
        _, jet, _, done = f.jet()
        f.run()
        g = f.read(raw=True)
    print(f"gpu frame (orbit, jet, exact pass) in {time.time() - t:.1f} s; jet computed "
          f"{done} of {NX * NY}", flush=True)
    assert g[0]["kernel"].startswith("pert-fx-scaled/1 bits=")
    assert g[2]["reference_len"] == 1 << 20

    c = oracle_frame(sc, cfg, orbit)
    compare_frames(g, c)
    assert (g[1]["raw"]["kind"] == fa.UNRESOLVED).all()
    assert g[2]["equivalent"] == NX * NY * MAX_ITER
    assert g[2]["blocks"] > 0 and g[2]["skipped"] > 0

    w, h = NX // 2, NY // 2
    kw = dict(outcomes=False)
    first = oracle_frame(sc, cfg, orbit, mask=oracle.jetaa_first(w, h), **kw)["columns"]
    todo = oracle.jetaa_rebuild(first, w, h)
    if todo.any():
        oracle.jetaa_merge(first, oracle_frame(sc, cfg, orbit, mask=todo, **kw)["columns"],
                           todo)
    assert done == w * h + int(todo.sum())
    for k in ("class", "nu", "de", "normal"):
        assert_same(jet[k], first[k], k)
