"""The reference GPU route's certified error control, restated (test infrastructure: the
oracle; the library's own versions are checked against it at the end):

* Reference::error_radius (crates/fd-kernel/src/reference.rs:122-148): q[m] bounds the
  stored orbit's distance from the exact orbit.  Pinned by the mathematics: against a
  512-bit fixed-point orbit of the same (dyadic) centre, at the reference's own test centres
  (sample.rs:1042-1090), q[m] >= |A_m - Z_m| at every point, and not by more than the
  rounding a few steps accumulate (a bound a dropped term would break either way).
* GuardedBla::build (crates/fd-kernel/src/bla/guarded.rs:103-134): the reference's own
  test (guarded.rs:136-212) restated -- direct binary64 windows inside each sampled valid
  block stay within the derivative envelope, over more than 500 blocks, while the state
  envelope alone demonstrably misses some -- and its refusals.
* external_escape_ok (crates/fd-kernel/src/store.rs:119-198): closed cases of its five
  conditions (escape proven, derivative radius, displacement, DE ratio, angle).
"""
import math

import numpy as np
import pytest

import oracle

U = np.finfo(np.float64).eps / 2
EPS_MAX = 1.0 / (1 << 41)
CENTRES = [(0.375, 0.28125), (-0.7421875, 0.0546875), (0.25, 0.0)]


def exact_orbit(cr, ci, n):
    return oracle.reference_fixed(repr(cr), repr(ci), 512, n)


def f64_orbit(cr, ci, n):
    zr, zi, bits = oracle.reference(repr(cr), repr(ci), "1e-3", 0.0, 8, 8, n, 65536.0,
                                    tier=oracle.TIER_F64)
    assert bits == 53
    return zr, zi


@pytest.mark.parametrize("cr,ci", CENTRES)
def test_error_radius_covers_the_exact_orbit(cr, ci):
    ar, ai = exact_orbit(cr, ci, 10_000)
    for kind in ("f64", "fixed"):
        if kind == "f64":
            zr, zi = f64_orbit(cr, ci, 10_000)
            q = oracle.error_radius(zr, zi, None, cr, ci)
        else:
            zr, zi = oracle.reference_fixed(repr(cr), repr(ci), 256, 10_000)
            q = oracle.error_radius(zr, zi, 256, cr, ci)
        n = min(len(zr), len(ar))
        assert n > 30
        d = np.hypot(zr[:n] - ar[:n], zi[:n] - ai[:n])
        az = np.hypot(zr[:n], zi[:n])
        # |A_m - Z_m| <= q[m], A rounded to f64 (the reference's own allowance 4u|Z|)
        assert np.all(d <= q[:n] + 4 * U * az), (kind, np.max(d - q[:n]))
        assert np.all(q >= 0) and q[1] > 0
        # and tight: where the orbit has already drifted, q is not orders of magnitude above it
        live = d > 1e3 * np.finfo(float).tiny
        if live.any():
            assert np.median(q[:n][live] / d[live]) < 1e6, kind


def test_error_radius_fixed_defect_is_the_fraction_bits():
    # three truncated squares and the centre's truncation: 2^(4 - bits) per step
    zr = np.zeros(4)
    zi = np.zeros(4)
    q = oracle.error_radius(zr, zi, 100, 0.0, 0.0)
    tiny = np.finfo(float).tiny
    assert q[0] == tiny
    assert q[1] == pytest.approx(2.0 ** -96, rel=1e-12)
    assert np.all(np.diff(q) >= 0)
    # an f64 orbit's defect instead: 8u(|Z|^2 + |c|) + 2u|c| per step
    q64 = oracle.error_radius(zr, zi, None, 0.5, 0.0)
    assert q64[1] == pytest.approx(8 * U * 0.5 + 2 * U * 0.5, rel=1e-12)


def affine(x, a, b):
    return (a[0] * x[0] - a[1] * x[1] + b[0] + 0.0, a[0] * x[1] + a[1] * x[0] + b[1] + 0.0)


def norm(z):
    return abs(z[0]) + abs(z[1])


def test_guarded_bla_direct_derivative_windows():
    checked = naive_misses = 0
    worst = 0.0
    for cr, ci, c in ((-0.743643887037158, 0.131825904205312, 1e-28),
                      (-0.7453, 0.1127, 1e-24), (0.25, 0.0, 1e-24)):
        zr, zi, _ = oracle.reference(repr(cr), repr(ci), "1e-3", 0.0, 8, 8, 10_000, 1e10,
                                     tier=oracle.TIER_F64)
        q = np.zeros(len(zr))
        for eps in (EPS_MAX, 2.0 ** -50):
            levels = oracle.guarded_build(zr, zi, q, eps, c)
            for j, level in enumerate(levels):
                for i in range(0, len(level), 17):
                    g = level[i]
                    if not g[4] > 0.0:
                        continue
                    a, b, r, alpha, beta, da, db = (g[0], g[1]), (g[2], g[3]), *g[4:]
                    m, ln = 1 + (i << (j + 1)), 2 << j
                    delta = (0.9 * r, 0.05 * r)
                    dc = (0.3 * c, -0.2 * c)
                    d0 = (0.7, -0.3)
                    d, der = delta, d0
                    for k in range(m, m + ln):
                        z = (zr[k] + d[0], zi[k] + d[1])
                        der = affine(der, (2 * z[0], 2 * z[1]), (1.0, 0.0))
                        d = affine(d, (2 * zr[k] + d[0], 2 * zi[k] + d[1]), dc)
                    approx = affine(d0, a, b)
                    err = math.hypot(der[0] - approx[0], der[1] - approx[1])
                    bound = da * norm(d0) + db
                    rnd = 64 * np.finfo(float).eps * ln * (norm(der) + norm(approx))
                    assert err <= bound + rnd, (m, ln, err, bound)
                    state = affine(delta, a, (b[0] * dc[0] - b[1] * dc[1],
                                              b[0] * dc[1] + b[1] * dc[0]))
                    serr = math.hypot(d[0] - state[0], d[1] - state[1])
                    sb = alpha * norm(delta) + beta * norm(dc)
                    assert serr <= sb + 64 * np.finfo(float).eps * ln * (norm(d) + norm(state))
                    naive_misses += err > alpha * norm(d0) + beta
                    worst = max(worst, err / (bound + rnd))
                    checked += 1
    assert checked > 500, checked
    # the state envelope alone would not cover the derivative
    assert naive_misses > 0
    print(f"guarded windows checked {checked}, naive misses {naive_misses}, worst {worst:.3g}")


def test_guarded_bla_refusals():
    zr, zi, _ = oracle.reference("0.25", "0", "1e-3", 0.0, 8, 8, 20, 1e10, tier=oracle.TIER_F64)
    with pytest.raises(ValueError):
        oracle.guarded_build(zr, zi, np.full(len(zr), np.nan), 2.0 ** -50, 1e-20)
    with pytest.raises(ValueError):
        oracle.guarded_build(zr, zi, np.zeros(len(zr)), 2.0 ** -20, 1e-20)  # eps > 2^-24


def test_guarded_bla_envelopes_contain_the_plain_table():
    # same A, B as Bla::build_scaled where both are valid; radii only shrink, alpha/beta grow
    zr, zi, _ = oracle.reference("-0.7453", "0.1127", "1e-3", 0.0, 8, 8, 5000, 1e10,
                                 tier=oracle.TIER_F64)
    g = oracle.guarded_build(zr, zi, np.zeros(len(zr)), 2.0 ** -50, 1e-24)
    p = oracle.bla_build(zr, zi, 2.0 ** -50, 1e-24)
    assert len(g) >= 3
    for lg, lp in zip(g, p):
        both = (lg[:, 4] > 0) & (lp[:, 4] > 0)
        assert both.any()
        assert np.array_equal(lg[both, :4], lp[both, :4])
        assert np.all(lg[both, 4] <= lp[both, 4])
        assert np.all(lg[both, 5] >= lp[both, 5]) and np.all(lg[both, 6] >= lp[both, 6])
        assert not np.any((lg[:, 4] > 0) & ~(lp[:, 4] > 0))


def test_external_escape_ok_cases():
    n, z, d, dexp = 1234, (1e5, 2e4), (3.0, -1.0), 10
    px = (1.0, 0)  # de = 2|z| ln|z| / |dz/dc| ~ 700 output pixels
    assert oracle.external_escape_ok(n, z, d, dexp, 0.0, 0.0, *px)
    assert oracle.external_escape_ok(n, z, d, dexp, 1e-6, 1e-9, *px)
    # radii that do not prove escape (|z| - ez <= 2) or a derivative radius >= |D| / 2
    assert not oracle.external_escape_ok(n, z, d, dexp, math.hypot(*z) - 2.0, 0.0, *px)
    assert not oracle.external_escape_ok(n, z, d, dexp, 0.0, 0.5 * math.hypot(*d), *px)
    # a displacement above 5e-4 px, a DE ratio beyond 1e-3, an angle beyond 0.1 degree
    assert not oracle.external_escape_ok(n, z, d, dexp, 1e-2 * math.hypot(*z), 0.0, *px)
    assert not oracle.external_escape_ok(n, z, d, dexp, 0.0, 1e-3 * math.hypot(*d), *px)
    # non-finite or negative radii
    for ez, ed in ((math.nan, 0.0), (0.0, math.inf), (-1.0, 0.0)):
        assert not oracle.external_escape_ok(n, z, d, dexp, ez, ed, *px)


# ------------------------------------------------------------ the library against the oracle

def _product():
    fa = pytest.importorskip("fractodactyl_amd")
    try:
        fa.lib()
    except Exception as e:
        pytest.skip(str(e))
    return fa


def test_library_error_radius_and_guarded_table_equal_the_oracle():
    fa = _product()
    cases = [("-0.743643887037158704752191506114774", "0.131825904205311970493132056385139",
              "1e-28", 30_000)]
    for re_, im_, w, it in cases:
        zr, zi, bits = oracle.reference(re_, im_, w, 0.0, 64, 36, it, 1e10)
        pl = oracle.plane(re_, im_, w, 0.0, 64, 36)
        for b in (None, bits):
            qa = fa.error_radius(zr, zi, b, pl["c_re"], pl["c_im"])
            qo = oracle.error_radius(zr, zi, b, pl["c_re"], pl["c_im"])
            assert np.array_equal(qa.view(np.uint64), qo.view(np.uint64))
        q = oracle.error_radius(zr, zi, bits, pl["c_re"], pl["c_im"])
        dc = pl["h_m"] * 2.0 ** pl["h_e"] * math.hypot(32, 18)
        ga = fa.guarded_bla(zr, zi, q, 2.0 ** -50, dc)
        go = oracle.guarded_build(zr, zi, q, 2.0 ** -50, dc)
        assert len(ga) == len(go) >= 5
        for a, b in zip(ga, go):
            assert np.array_equal(a.view(np.uint64), b.view(np.uint64))
    with pytest.raises(fa.FdgError):
        fa.guarded_bla(zr, zi, np.full(len(zr), -1.0), 2.0 ** -50, 1e-20)


def test_library_admission_equals_the_oracle():
    fa = _product()
    rng = np.random.default_rng(7)
    agree = admitted = 0
    for _ in range(3000):
        n = int(rng.integers(10, 10**6))
        az = 10.0 ** rng.uniform(0.5, 6)
        th = rng.uniform(0, 2 * math.pi)
        z = (az * math.cos(th), az * math.sin(th))
        d = tuple(rng.normal(size=2))
        dexp = int(rng.integers(-50, 3000))
        ez = az * 10.0 ** rng.uniform(-16, -1)
        ed = math.hypot(*d) * 10.0 ** rng.uniform(-16, -1)
        px = (10.0 ** rng.uniform(-3, 3), -dexp + int(rng.integers(-20, 20)))
        a = fa.external_escape_ok(n, z, d, dexp, ez, ed, *px)
        b = oracle.external_escape_ok(n, z, d, dexp, ez, ed, *px)
        agree += a == b
        admitted += a
    assert agree == 3000
    assert 100 < admitted < 2900


def test_overlay_writes_admitted_states_through_the_encoder():
    fa = _product()
    re_, im_, w = "-0.7453", "0.1127", "0.0065"
    cols = fa.COL_NU | fa.COL_DE | fa.COL_NORMAL
    full = oracle.render(re_, im_, w, 0.0, 48, 27, 5000, 1e10, columns=cols, bla=False)
    o = full["outcomes"]
    present = o["kind"] == oracle.ESCAPED
    present[::3] = False  # overlay a subset; the others keep what the columns held
    out = {"class": np.full(48 * 27, 9, np.uint8), "nu": np.full(48 * 27, -1.0),
           "de": np.full(48 * 27, -1.0, np.float32), "normal": np.full(48 * 27, 7, np.uint16)}
    states = {k: o[k] for k in ("n", "zr", "zi", "dr", "di", "dexp")}
    states["present"] = present
    fa.overlay_escaped(fa.View(re_, im_, w, 0.0), fa.Params(48, 27, 1, 5000, 1e10, cols), out,
                       states)
    for k in ("class", "nu", "de", "normal"):
        assert np.array_equal(out[k][present].view(np.uint8),
                              full["columns"][k][present].view(np.uint8)), k
    assert np.all(out["class"][~present] == 9) and np.all(out["nu"][~present] == -1.0)
    bad = dict(states)
    bad["zr"] = np.where(present, 0.0, bad["zr"])  # |z| within the escape radius
    with pytest.raises(fa.FdgError):
        fa.overlay_escaped(fa.View(re_, im_, w, 0.0), fa.Params(48, 27, 1, 5000, 1e10, cols),
                           out, bad)
