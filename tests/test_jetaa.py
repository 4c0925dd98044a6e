"""Jet-AA's rebuild (crates/fd-cli/src/jetaa.rs): the library's host rebuild against the
oracle's restatement, the reference's own jet-AA tests (jetaa.rs:262-340), and a closed
form: columns sampled from a cubic F(s), for which the stencil's F'' and F''' are exact,
must be predicted at the other sample positions to within the columns' quantisation.
"""
import numpy as np
import pytest

import oracle

SEAHORSE = ("-0.7453", "0.1127", "0.0065")


def _product():
    fa = pytest.importorskip("fractodactyl_amd")
    try:
        fa.lib()
    except Exception as e:  # built by __graft_entry__.build()
        pytest.skip(str(e))
    return fa


def test_first_marks_each_pixels_top_left_sample():
    # jetaa.rs:318-323
    for first in [oracle.jetaa_first] + ([_product().jetaa_first]):
        m = first(3, 2).astype(bool)
        assert m.sum() == 6
        assert m[0] and not m[1] and m[2] and not m[6] and m[12]


def exact(rotation, w=96, h=54):
    """The exact ss 2 render of jetaa.rs:268-290's seahorse view (oracle)."""
    f = oracle.render(*SEAHORSE, rotation, 2 * w, 2 * h, 20_000, 1e10,
                      columns=oracle.COL_NU | oracle.COL_DE | oracle.COL_NORMAL, ss=2)
    return f["columns"], w, h


def keep_first(cols, w, h):
    keep = oracle.jetaa_first(w, h).astype(bool)
    s = {k: v.copy() for k, v in cols.items()}
    s["class"][~keep] = 0  # Class(0), as the reference's check() leaves them
    return s, keep


@pytest.mark.parametrize("rotation", [0.0, 0.7, 2.0])
def test_library_rebuild_equals_the_oracle(rotation):
    fa = _product()
    ex, w, h = exact(rotation)
    a, _ = keep_first(ex, w, h)
    b = {k: v.copy() for k, v in a.items()}
    ta = oracle.jetaa_rebuild(a, w, h)
    tb = fa.jetaa_rebuild(b, w, h)
    assert np.array_equal(ta, tb)
    for k in ("class", "nu", "de", "normal"):
        assert np.array_equal(a[k].view(np.uint8), b[k].view(np.uint8)), k
    # merge copies the masked samples only
    c = {k: v.copy() for k, v in a.items()}
    d = {k: v.copy() for k, v in a.items()}
    oracle.jetaa_merge(c, ex, ta)
    fa.jetaa_merge(d, ex, ta)
    for k in ("class", "nu", "de", "normal"):
        assert np.array_equal(c[k].view(np.uint8), d[k].view(np.uint8)), k


@pytest.mark.parametrize("rotation", [0.0, 0.7])
def test_accepted_predictions_match_the_exact_samples(rotation):
    # jetaa.rs:292-340: accepted share > 0.25, worst |dnu| < 1e-2, worst relative dDE < 5e-2,
    # and every accepted sample has the exact sample's kind
    ex, w, h = exact(rotation)
    s, keep = keep_first(ex, w, h)
    todo = oracle.jetaa_rebuild(s, w, h).astype(bool)
    acc = ~keep & ~todo
    share = acc.sum() / (3 * w * h)
    assert share > 0.25, share
    assert np.array_equal(s["class"][acc] & 3, ex["class"][acc] & 3)
    esc = acc & ((ex["class"] & 3) == 0)
    dnu = np.abs(s["nu"][esc] - ex["nu"][esc]).max()
    dde = (np.abs(s["de"][esc] - ex["de"][esc]) / ex["de"][esc]).max()
    assert dnu < 1e-2 and dde < 5e-2, (dnu, dde)


def cubic_columns(w, h, c):
    """Columns of the samples of a w x h pixel ss 2 grid for F(s) = sum c[k] s^k, s = x - i y
    in output pixels: nu = -log2 Re F, de = 2 Re F / |F'|, normal = arg F' (quantised as
    Samples::angle).  Returns the columns and the exact (nu, de, angle) per sample."""
    nx, ny = 2 * w, 2 * h
    j, i = np.divmod(np.arange(nx * ny), nx)
    s = (i * 0.5) - 1j * (j * 0.5)
    F = sum(ck * s ** k for k, ck in enumerate(c))
    dF = sum(k * ck * s ** (k - 1) for k, ck in enumerate(c) if k)
    nu = -np.log2(F.real)
    de = 2 * F.real / np.abs(dF)
    ang = np.angle(dF)
    t = np.round(ang * (65536.0 / (2 * np.pi))).astype(np.int64).astype(np.uint16)
    cols = {"class": np.zeros(nx * ny, np.uint8), "nu": nu.copy(),
            "de": de.astype(np.float32), "normal": t}
    return cols, (nu, de, ang)


@pytest.mark.parametrize("coef", [
    (60.0, 0.8 - 0.3j, 0.02 + 0.015j, -0.0004 + 0.0002j),
    (200.0, -1.5 + 0.4j, -0.01 + 0.03j, 0.0002 - 0.0007j),
])
def test_cubic_jets_are_predicted_exactly(coef):
    # F' is quadratic, so the neighbour stencil recovers F'' and F''' exactly and every
    # interior pixel's three predictions are consistent: each must be accepted and equal
    # the closed form to the precision of the stored columns (f32 de, 2^-16-turn normal)
    w, h = 24, 16
    cols, (nu, de, ang) = cubic_columns(w, h, coef)
    keep = oracle.jetaa_first(w, h).astype(bool)
    s = {k: v.copy() for k, v in cols.items()}
    for k in ("nu", "de", "normal"):
        s[k][~keep] = 0
    todo = oracle.jetaa_rebuild(s, w, h).astype(bool)
    nx = 2 * w
    px, py = (np.arange(4 * w * h) % nx) // 2, (np.arange(4 * w * h) // nx) // 2
    inner = (px >= 1) & (py >= 1) & (px + 2 < w) & (py + 2 < h) & ~keep
    assert not todo[inner].any()
    assert (s["class"][inner] == 0).all()
    assert np.abs(s["nu"][inner] - nu[inner]).max() < 1e-5
    assert (np.abs(s["de"][inner] - de[inner]) / de[inner]).max() < 1e-4
    got = s["normal"][inner].astype(np.float64) * (2 * np.pi / 65536.0)
    d = np.angle(np.exp(1j * (got - ang[inner])))
    assert np.abs(d).max() < 4 * (2 * np.pi / 65536.0)
    # the border pixels are never predicted
    assert todo[~inner & ~keep].all()
    # a wrong stencil sign breaks it: mirror the columns vertically (F(conj-ish) is no
    # longer the function whose jets are stored) and most predictions are rejected
    bad = {k: v.copy() for k, v in cols.items()}
    bad["normal"] = (65536 - bad["normal"].astype(np.int64)).astype(np.uint16)
    for k in ("nu", "de", "normal"):
        bad[k][~keep] = 0
    tb = oracle.jetaa_rebuild(bad, w, h).astype(bool)
    assert tb[inner].mean() > 0.5


def test_interior_pixels_predict_their_own_kind():
    # a pixel whose sample and five neighbours are all Interior predicts Interior
    # (heuristic evidence) for its three other samples; one Unresolved neighbour stops it
    w, h = 8, 6
    n = 4 * w * h
    cols = {"class": np.full(n, 1 | (2 << 2), np.uint8), "nu": np.zeros(n),
            "de": np.zeros(n, np.float32), "normal": np.zeros(n, np.uint16)}
    keep = oracle.jetaa_first(w, h).astype(bool)
    cols["class"][~keep] = 0
    nx = 2 * w
    cols["class"][(2 * 3) * nx + 2 * 4] = 2  # pixel (4, 3) Unresolved
    s = {k: v.copy() for k, v in cols.items()}
    todo = oracle.jetaa_rebuild(s, w, h).astype(bool)

    def sample(px, py, k):
        return (2 * py + k // 2) * nx + 2 * px + k % 2

    assert all(s["class"][sample(2, 2, k)] == 1 for k in (1, 2, 3))   # certified -> heuristic
    assert all(todo[sample(3, 3, k)] for k in (1, 2, 3))              # its right neighbour differs
    assert all(todo[sample(4, 3, k)] for k in (1, 2, 3))
    fa = _product()
    s2 = {k: v.copy() for k, v in cols.items()}
    assert np.array_equal(fa.jetaa_rebuild(s2, w, h), todo.astype(np.uint8))
    assert np.array_equal(s2["class"], s["class"])
