"""Writes direct_hp.json: direct high-precision iteration of z <- z^2 + c at 256 seeded
pixels of each benchmark frame, the ground truth the perturbation forms are compared with
(the reference's own check of this kind is tools/oracle.py:70-80 in fractodactyl).

Each pixel's c is the exact value the renderer iterates: the frame's centre (the f64 centre
on the f64 tier, the decimal centre otherwise) plus the pixel's f64 offset (view.rs:117-122
unit_offset times h, or h_m 2^h_e on the scaled tier; grid.rs:652-725), written out as an
exact decimal.  z <- z^2 + c is iterated in binary fixed point with the oracle's
restatement of fd_fixed::orbit's arithmetic (crates/fd-fixed/src/orbit.rs; oracle.direct)
at two precisions, the frame's reference precision + 64 and + 192 bits, with dz/dc carried
alongside in scaled form; past |z|^2 > 2^32 the few steps to the escape radius run in f64
(|z| >= 2^16 makes them well conditioned).  A pixel whose two precisions disagree (escape
count, or z at escape beyond 1e-9 relative) is kept with "converged": false and left out
of the comparisons.

The pixels: 256 seeded uniformly over the frame, plus, for frames whose uniform pixels
rarely escape (the 1e-1000 frame: 9 of 256), 128 seeded among the pixels the oracle's
render finds escaped.  For each pixel: n = the first iterate with fl(re^2 + im^2) > R^2 (-1: none through
max_iter), z_n as f64, a tie flag when |z_n|^2 or |z_{n-1}|^2 lies within 1e-12 of R^2
(there the escape count may legitimately differ by one), and for escaped pixels log2 of
the distance estimate in pixels, de = 2 |z| ln|z| / |dz/dc| / h (tools/oracle.py:119).  The
comparison gates the smooth count as a displacement, |d nu| de ln2 / 2 pixels
(tools/oracle.py:8-12): near the set nu is steep, and a sub-pixel shift of c within a
perturbation form's precision moves it by far more than 1e-3.

Usage: python tests/golden/make_direct_hp.py [config ...]   (default 1 2 3; about 5
minutes on 8 cores, nearly all of it config 3's bounded pixels at 1e6 iterations).
"""
import json
import math
import multiprocessing as mp
import os
import sys
from decimal import Decimal, getcontext

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from fractodactyl_amd.scenes import scene  # noqa: E402

OUT = os.path.join(HERE, "direct_hp.json")
PIXELS = 256
GUARD = (64, 192)


def exact_c(sc, pl, i, j):
    """Decimal strings of the pixel's c, as the renderer's tier represents it."""
    x, y = i + 0.5 - pl["half_x"], j + 0.5 - pl["half_y"]
    ux = pl["cos_r"] * x + pl["sin_r"] * y
    uy = pl["sin_r"] * x - pl["cos_r"] * y
    getcontext().prec = 6000
    if pl["tier"] == oracle.TIER_SCALED:
        scale = Decimal(2) ** pl["h_e"]
        dr, di = Decimal(ux * pl["h_m"]) * scale, Decimal(uy * pl["h_m"]) * scale
    else:
        h = pl["h_m"] * oracle.exp2i(pl["h_e"])
        dr, di = Decimal(ux * h), Decimal(uy * h)
    if pl["tier"] == oracle.TIER_F64:
        cr, ci = Decimal(pl["c_re"]), Decimal(pl["c_im"])
    else:
        cr, ci = Decimal(sc["center_re"]), Decimal(sc["center_im"])
    return format(cr + dr, "f"), format(ci + di, "f")


def log2_de_px(zr, zi, d, pl):
    """log2 of 2 |z| ln|z| / |dz/dc| / h, with dz/dc = (d[0] + i d[1]) 2^d[2] and
    h = h_m 2^h_e."""
    az = math.hypot(zr, zi)
    return (math.log2(2.0 * az * math.log(az)) - math.log2(math.hypot(d[0], d[1])) - d[2]
            - math.log2(pl["h_m"]) - pl["h_e"])


def one(args):
    k, i, j = args
    sc = scene(k)
    c = sc["config"]
    pl = oracle.plane(sc["center_re"], sc["center_im"], c["width"], 0.0, c["nx"], c["ny"])
    cre, cim = exact_c(sc, pl, i, j)
    r2 = c["escape_radius"] ** 2
    runs = [oracle.direct(cre, cim, int(pl["bits"]) + g, c["max_iter"], r2) for g in GUARD]
    (n1, z1, _, p1, q1), (n2, z2, d2, p2, q2) = runs
    conv = n1 == n2 and (n1 < 0 or abs(z1 - z2) <= 1e-9 * abs(z2))
    tie = n2 >= 0 and (abs(q2 / r2 - 1) < 1e-12 or abs(p2 / r2 - 1) < 1e-12)
    lde = log2_de_px(z2.real, z2.imag, d2, pl) if n2 >= 0 else None
    return dict(i=i, j=j, n=n2, zr=z2.real, zi=z2.imag, tie=bool(tie), converged=bool(conv),
                log2_de_px=lde)


# frames whose uniform pixels rarely escape get this many more pixels, drawn (seeded) from
# those the oracle's perturbation render finds escaped, so that escape counts and smooth
# counts are checked on enough pixels (selection by the render does not bias a per-pixel
# comparison with direct iteration)
ESCAPED_EXTRA = {3: 128}


def pixels(k):
    sc = scene(k)
    c = sc["config"]
    rng = np.random.default_rng(7000 + k)
    ii = rng.integers(0, c["nx"], PIXELS)
    jj = rng.integers(0, c["ny"], PIXELS)
    out = [(k, int(a), int(b)) for a, b in zip(ii, jj)]
    extra = ESCAPED_EXTRA.get(k, 0)
    if extra:
        cand = rng.choice(c["nx"] * c["ny"], 40 * extra, replace=False)
        mask = np.zeros(c["nx"] * c["ny"], np.uint8)
        mask[cand] = 1
        r = oracle.render(sc["center_re"], sc["center_im"], c["width"], 0.0, c["nx"], c["ny"],
                          c["max_iter"], c["escape_radius"], columns=0, bla=c["bla"],
                          mask=mask)
        esc = [int(q) for q in cand if r["outcomes"]["kind"][q] == oracle.ESCAPED][:extra]
        out += [(k, q % c["nx"], q // c["nx"]) for q in esc]
    return out


def main():
    ks = [int(a) for a in sys.argv[1:]] or [1, 2, 3]
    out = json.load(open(OUT)) if os.path.exists(OUT) else {}
    with mp.Pool(min(8, os.cpu_count() or 1)) as pool:
        for k in ks:
            sc = scene(k)
            c = sc["config"]
            pl = oracle.plane(sc["center_re"], sc["center_im"], c["width"], 0.0, c["nx"],
                              c["ny"])
            rows = pool.map(one, pixels(k), chunksize=1)
            out[str(k)] = dict(
                centre=[sc["center_re"], sc["center_im"]], width=c["width"], nx=c["nx"],
                ny=c["ny"], max_iter=c["max_iter"], escape_radius=c["escape_radius"],
                tier=int(pl["tier"]), bits=[int(pl["bits"]) + g for g in GUARD],
                **{key: [r[key] for r in rows]
                   for key in ("i", "j", "n", "zr", "zi", "tie", "converged", "log2_de_px")})
            esc = sum(r["n"] >= 0 for r in rows)
            print(f"config {k}: {esc} escaped, {len(rows) - esc} bounded, "
                  f"{sum(not r['converged'] for r in rows)} unconverged", flush=True)
            json.dump(out, open(OUT, "w"), indent=None, separators=(",", ":"))


if __name__ == "__main__":
    main()
