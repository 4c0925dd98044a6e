"""The FP32 floatexp delta form (FDG_FORM_FE32) against the FP64 form.

FE32 iterates FP32 mantissas with carried exponents, so it is close to, not equal to, the
reference's FP64 results; these tests pin its measured agreement on frames of the
reference's own golden set (KERN-05) and on the 1080p interior scene, as regression
guards: classes, escape counts and smooth iteration counts against the FP64 form, which
itself equals the oracle bit for bit (test_gpu_parity.py).
"""
import numpy as np
import pytest

import fractodactyl_amd as fa
from fdsfile import kern05, parse_fds
from fractodactyl_amd.scenes import scene

pytestmark = pytest.mark.gpu

COLS = fa.COL_NU | fa.COL_DE | fa.COL_NORMAL


def agreement(a, b):
    ka, kb = a[1]["raw"]["kind"], b[1]["raw"]["kind"]
    esc = (ka == 0) & (kb == 0)
    dn = b[1]["raw"]["n"][esc].astype(np.int64) - a[1]["raw"]["n"][esc].astype(np.int64)
    dnu = np.abs(b[1]["nu"][esc] - a[1]["nu"][esc])
    if not esc.any():  # (no sample escapes in both: nothing to compare but the classes)
        return dict(classes=float((ka == kb).mean()), n_within1=1.0, nu_1e3=1.0)
    return dict(classes=float((ka == kb).mean()), n_within1=float((np.abs(dn) <= 1).mean()),
                nu_1e3=float((dnu <= 1e-3).mean()))


def both(name, route, nx=None, ny=None):
    h, _ = parse_fds(dict(kern05())[name])
    v = fa.View(h["re"], h["im"], h["width"], h["rotation"])
    p = fa.Params(nx or h["nx"], ny or h["ny"], 1, h["max_iter"], h["escape_radius"], COLS)
    a = fa.render(v, p, route, fa.FORM_F64, raw=True)
    b = fa.render(v, p, route, fa.FORM_FE32, raw=True)
    assert b[0]["kernel"].endswith(" gpu-hip-fe32/1")
    return agreement(a, b)


# (frame, route, classes, n within 1, smooth count within 1e-3), measured on MI355X with
# the FE32 kernel's Brent/Newton periodicity (classes 1.0 on every KERN-05 frame; n within 1
# 0.94 / 1.0 / 0.86 / 1.0 / 1.0 / 1.0 / 0.92 / 0.64; smooth count 0.81 / 1.0 / 0.40 / 0.9988 /
# 0.9986 / 0.9989 / 0.63 / 0.0002), with margin below
CASES = [("seahorse", fa.ROUTE_PLAIN, 1.0, 0.93, 0.80),
         ("cusp-1e-8", fa.ROUTE_BLA, 1.0, 0.999, 0.999),
         ("valley-1e-28", fa.ROUTE_PLAIN, 1.0, 0.85, 0.39),
         ("i-1e-300", fa.ROUTE_BLA, 1.0, 0.999, 0.99),
         ("i-1e-1000", fa.ROUTE_BLA, 1.0, 0.999, 0.99),
         ("i-1e-50-rot", fa.ROUTE_PLAIN, 1.0, 0.999, 0.99),
         ("rung-7676", fa.ROUTE_BLA, 1.0, 0.90, 0.6),
         ("rung-7676", fa.ROUTE_PLAIN, 1.0, 0.6, 0.0)]


@pytest.mark.parametrize("name,route,cls,n1,nu", CASES)
def test_fe32_agreement_with_fp64(name, route, cls, n1, nu):
    g = both(name, route)
    print(name, g, flush=True)
    assert g["classes"] >= cls and g["n_within1"] >= n1 and g["nu_1e3"] >= nu, g


def test_fe32_finds_the_nucleus_interior():
    # every sample of the 1e-100 scene is Interior by the nucleus contraction heuristic
    sc = scene(2)
    c = sc["config"]
    v = fa.View(sc["center_re"], sc["center_im"], c["width"])
    p = fa.Params(192, 108, 1, c["max_iter"], c["escape_radius"], COLS)
    b = fa.render(v, p, fa.ROUTE_BLA, fa.FORM_FE32, raw=True)
    assert (b[1]["raw"]["kind"] == fa.INTERIOR).all()
    assert (b[1]["raw"]["n"] == sc["nucleus"]["period"] + 1).all()


@pytest.mark.parametrize("name,route", [("rung-7676", fa.ROUTE_BLA), ("rung-7676", fa.ROUTE_PLAIN),
                                        ("i-1e-300", fa.ROUTE_BLA), ("valley-1e-28", fa.ROUTE_BLA),
                                        ("seahorse", fa.ROUTE_PLAIN)])
def test_mixed_form_is_exact_where_it_recomputes_and_certified_elsewhere(name, route):
    # the mixed form recomputes in FP64 every sample its FP32 pass found escaped or could not
    # certify (error bound reaching the escape radius): its exterior classes, escape counts
    # and escaped columns are the FP64 form's (= the oracle's) exactly; every other sample
    # keeps the FP32 pass's outcome, whose class the FP64 form shares on these frames
    h, _ = parse_fds(dict(kern05())[name])
    v = fa.View(h["re"], h["im"], h["width"], h["rotation"])
    p = fa.Params(h["nx"], h["ny"], 1, h["max_iter"], h["escape_radius"], COLS)
    a = fa.render(v, p, route, fa.FORM_F64, raw=True)
    b = fa.render(v, p, route, fa.FORM_FE32, raw=True)
    m = fa.render(v, p, route, fa.FORM_MIXED, raw=True)
    assert m[0]["kernel"].endswith(" gpu-hip-mixed/1")
    ka, kb, km = a[1]["raw"]["kind"], b[1]["raw"]["kind"], m[1]["raw"]["kind"]
    assert np.array_equal(ka == fa.ESCAPED, km == fa.ESCAPED)
    assert np.array_equal(ka, km)
    esc = ka == fa.ESCAPED
    for k in ("n", "zr", "zi", "dr", "di", "dexp"):
        assert np.array_equal(m[1]["raw"][k][esc], a[1]["raw"][k][esc]), k
    for k in ("class", "nu", "de", "normal"):
        assert np.array_equal(m[1][k].view(np.uint8), a[1][k].view(np.uint8)), k
    # a sample is either the FP64 form's (recomputed) or the FP32 pass's (certified)
    same64 = np.ones(ka.size, bool)
    same32 = np.ones(ka.size, bool)
    for k in ("kind", "n", "zr", "zi", "dr", "di", "dexp"):
        same64 &= m[1]["raw"][k] == a[1]["raw"][k]
        same32 &= m[1]["raw"][k] == b[1]["raw"][k]
    assert (same64 | same32).all()
    # work is counted once per sample, by the pass whose outcome it keeps
    assert m[2]["equivalent"] == (
        int(np.where(km == fa.UNRESOLVED, h["max_iter"],
                     m[1]["raw"]["n"]).astype(np.int64).sum()))
