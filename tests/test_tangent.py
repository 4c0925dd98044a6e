"""dz/dc and the screen-space Jacobian against central finite differences.

The reference carries dz/dc through plain steps (sample.rs:352-360) and BLA blocks
(sample.rs:304-312); its own known answer pins log2|dz/dc| at one 1e-1000 pixel
(sample.rs:1256-1286).  Here the derivative is checked against central differences of z
(Richardson-extrapolated, at the largest of steps 1e-4..1e-7 sample whose estimate has
converged) computed by the oracle's perturbation kernel at the displaced
offsets, on seeded escaped samples of the rung-7676 frame at 5.05e-997 (scaled tier,
BLA on) -- first for the oracle (CPU), then for the GPU through the C ABI
(fdg_jacobian), whose dz/dc is bit-identical to the oracle's.
"""
import math

import numpy as np
import pytest

import fractodactyl_amd as fa
import oracle
from fdsfile import kern05, parse_fds

STEPS = (1e-4, 1e-5, 1e-6, 1e-7)   # finite-difference steps tried, in samples
TOL = 1e-6          # relative tolerance of the 2x2 Jacobian (and of FD convergence)
NX, NY, ITERS = 96, 54, 150_000


@pytest.fixture(scope="module")
def frame():
    h, _ = parse_fds(dict(kern05())["rung-7676"])
    re_, im_, w = h["re"], h["im"], h["width"]
    pl = oracle.plane(re_, im_, w, 0.0, NX, NY)
    zr, zi, bits = oracle.reference(re_, im_, w, 0.0, NX, NY, ITERS, 1e10)
    levels = oracle.bla_build(zr, zi, oracle.EPS_EXACT, pl["h_m"] * math.hypot(NX / 2, NY / 2),
                              pl["h_e"])
    return dict(re=re_, im=im_, w=w, pl=pl, zr=zr, zi=zi, levels=levels,
                nucleus=oracle.nucleus_period(zr, zi))


def z_at(f, ux, uy):
    pl = f["pl"]
    r = oracle.sample(f["zr"], f["zi"], ux * pl["h_m"], uy * pl["h_m"], ITERS, 1e20,
                      extended=True, offset_exp=pl["h_e"], derivative=True,
                      bla_levels=f["levels"], nucleus=f["nucleus"])
    return r


def unit(pl, i, j):
    x, y = i + 0.5 - pl["half_x"], j + 0.5 - pl["half_y"]
    return pl["cos_r"] * x + pl["sin_r"] * y, pl["sin_r"] * x - pl["cos_r"] * y


def central(f, ux, uy, dx, dy, eps, n0):
    p = z_at(f, ux + eps * dx, uy + eps * dy)
    m = z_at(f, ux - eps * dx, uy - eps * dy)
    if not (p["kind"] == m["kind"] == oracle.ESCAPED and p["n"] == m["n"] == n0):
        return None
    return np.array([(p["zr"] - m["zr"]) / (2 * eps), (p["zi"] - m["zi"]) / (2 * eps)])


def richardson(f, i, j, n0, eps):
    """Central differences of z over the sample grid at steps eps and eps/2, extrapolated
    ((4 D(e/2) - D(e)) / 3, error O(e^4)): rows Re/Im z, columns x/y."""
    pl = f["pl"]
    ux, uy = unit(pl, i, j)
    c, s = pl["cos_r"], pl["sin_r"]
    cols = []
    for dx, dy in ((c, s), (s, -c)):  # d(ux, uy)/dx and d(ux, uy)/dy
        a = central(f, ux, uy, dx, dy, eps, n0)
        b = central(f, ux, uy, dx, dy, eps / 2, n0)
        if a is None or b is None:
            return None
        cols.append((4 * b - a) / 3)
    return np.array([[cols[0][0], cols[1][0]], [cols[0][1], cols[1][1]]])


def fd_jacobian(f, i, j, n0):
    """The first Richardson estimate, from the largest step down, that agrees with the
    estimate at the next smaller step to TOL -- a converged finite difference, found
    without looking at the analytic derivative -- or None.  Samples near the filaments
    bend z on sub-sample scales (z_n at n ~ 1.3e5 is a polynomial of degree 2^n in c), so
    the step that converges differs from sample to sample."""
    prev = None
    for eps in STEPS:
        cur = richardson(f, i, j, n0, eps)
        if cur is not None and prev is not None:
            if np.abs(cur - prev).max() <= TOL * np.abs(cur).max():
                return prev
        prev = cur
    return None


def jac_from_d(pl, dr, di, dexp):
    sc = 2.0 ** (dexp + pl["h_e"])
    ar, ai = dr * pl["h_m"] * sc, di * pl["h_m"] * sc
    c, s = pl["cos_r"], pl["sin_r"]
    return np.array([[ar * c - ai * s, ar * s + ai * c], [ar * s + ai * c, ai * s - ar * c]])


def seeded_pixels(k=12):
    rng = np.random.default_rng(1234)
    return [(int(rng.integers(0, NX)), int(rng.integers(0, NY))) for _ in range(k)]


def test_oracle_derivative_matches_finite_differences(frame):
    checked = 0
    for i, j in seeded_pixels():
        ux, uy = unit(frame["pl"], i, j)
        r = z_at(frame, ux, uy)
        if r["kind"] != oracle.ESCAPED:
            continue
        fd = fd_jacobian(frame, i, j, r["n"])
        if fd is None:
            continue
        jd = jac_from_d(frame["pl"], r["dr"], r["di"], r["dexp"])
        err = np.abs(fd - jd).max() / np.abs(jd).max()
        assert err < TOL, (i, j, err)
        checked += 1
    assert checked >= 6


@pytest.mark.gpu
def test_gpu_jacobian_matches_finite_differences(frame):
    view = fa.View(frame["re"], frame["im"], frame["w"], 0.0)
    params = fa.Params(NX, NY, 1, ITERS, 1e10)
    h, s, st = fa.render_bla(view, params, raw=True)
    j, e = fa.jacobian(h, s)
    checked = 0
    for i, jj in seeded_pixels():
        idx = jj * NX + i
        if s["raw"]["kind"][idx] != fa.ESCAPED:
            continue
        fd = fd_jacobian(frame, i, jj, int(s["raw"]["n"][idx]))
        if fd is None:
            continue
        jg = j[idx] * 2.0 ** int(e[idx])
        err = np.abs(fd - jg).max() / np.abs(jg).max()
        assert err < TOL, (i, jj, err)
        checked += 1
    assert checked >= 6


def test_host_jacobian_formula():
    # fdg_jacobian is host-only: a synthetic escaped outcome, rotation 0.4
    view = fa.View("0", "1", "1e-1000", 0.4)
    params = fa.Params(4, 2)
    n = 8
    raw = {"kind": np.zeros(n, np.uint8), "n": np.full(n, 10, np.uint64),
           "zr": np.ones(n), "zi": np.zeros(n), "dr": np.linspace(1, 2, n),
           "di": np.linspace(-1, 0.5, n), "dexp": np.full(n, 3330, np.int64),
           "ops": np.zeros(n, np.uint32)}
    raw["kind"][3] = fa.INTERIOR
    j, e = fa.jacobian({"view": view, "params": params}, {"raw": raw})
    pl = oracle.plane("0", "1", "1e-1000", 0.4, 4, 2)
    for i in range(n):
        if i == 3:
            assert not j[i].any()
            continue
        want = jac_from_d(pl, raw["dr"][i], raw["di"][i], 3330)
        got = j[i] * 2.0 ** int(e[i])
        assert np.allclose(got, want, rtol=1e-15, atol=0)
