"""CPU checks of the C ABI (include/fdg.h) and the host logic of libfdg.so: the library
loads, exports every declared entry point, and its host-only pieces (argument parsing,
the arbitrary-precision reference orbit, the .fds writer) agree with the oracle."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import fractodactyl_amd as fa
import oracle
from fdsfile import kern05, parse_fds

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, "include", "fdg.h")).read()
    return sorted(set(re.findall(r"\b(fdg_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = fa.lib()
    names = declared()
    assert "fdg_render" in names and "fdg_build_bla" in names
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(fa.SIGNATURES), set(names) ^ set(fa.SIGNATURES)
    assert L.fdg_abi_version() == 1


def test_enabled_parses_the_route_flag():
    # frame.rs:62-68
    assert fa.enabled(None) == (False, 0)
    assert fa.enabled("off") == (False, 0)
    assert fa.enabled("on") == (True, fa.ROUTE_PLAIN)
    assert fa.enabled("bla") == (True, fa.ROUTE_BLA)
    with pytest.raises(ValueError):
        fa.enabled("yes")


def test_create_without_a_device_fails_cleanly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    with pytest.raises(fa.FdgError) as e:
        fa.Context(0)
    assert e.value.code == -3


def _host_reference(view, params):
    ln, bits = C.c_uint64(), C.c_uint32()
    v, p = fa._view(view), fa._params(params)
    assert fa.lib().fdg_reference(None, C.byref(v), C.byref(p), None, None, 0, C.byref(ln),
                                  C.byref(bits)) == 0
    re_ = np.zeros(ln.value)
    im_ = np.zeros(ln.value)
    assert fa.lib().fdg_reference(None, C.byref(v), C.byref(p), fa._p(re_), fa._p(im_),
                                  ln.value, C.byref(ln), C.byref(bits)) == 0
    return re_, im_, bits.value


@pytest.mark.parametrize("re_,im_,width,iters,tier", [
    ("-0.7453", "0.1127", "0.0065", 20000, -1),
    ("-0.7453", "0.1127", "0.0065", 20000, 2),
    ("-0.743643887037158704752191506114774", "0.131825904205311970493132056385139", "1e-28",
     50000, -1),
    ("0", "1", "1e-1000", 20000, -1),
    ("-2", "0", "1e-200", 20000, -1),
])
def test_host_reference_orbit_matches_oracle(re_, im_, width, iters, tier):
    v = fa.View(re_, im_, width, 0.0)
    p = fa.Params(192, 108, max_iter=iters, tier=tier)
    a_re, a_im, a_bits = _host_reference(v, p)
    b_re, b_im, b_bits = oracle.reference(re_, im_, width, 0.0, 192, 108, iters, 1e10, tier)
    assert a_bits == b_bits
    assert np.array_equal(a_re.view(np.int64), b_re.view(np.int64))
    assert np.array_equal(a_im.view(np.int64), b_im.view(np.int64))


@pytest.mark.parametrize("width,iters", [("1e-300", 30000), ("1e-700", 20000),
                                         ("1e-1000", 20000), ("1e-2000", 6000)])
def test_threaded_karatsuba_orbit_matches_oracle(width, iters):
    # the fixed-point orbit from 16 limbs up squares on three threads, and from 40 limbs
    # up by Karatsuba (one or two levels at these widths: 18 to 108 limbs); both are
    # exact, so the orbit is the oracle's single-threaded schoolbook one bit for bit
    from fractodactyl_amd.scenes import scene
    s3 = scene(3)
    re_, im_ = s3["center_re"], s3["center_im"]
    v = fa.View(re_, im_, width, 0.0)
    p = fa.Params(64, 36, max_iter=iters)
    a_re, a_im, a_bits = _host_reference(v, p)
    b_re, b_im, b_bits = oracle.reference(re_, im_, width, 0.0, 64, 36, iters, 1e10)
    assert a_bits == b_bits and len(a_re) == len(b_re) > 1000
    assert np.array_equal(a_re.view(np.int64), b_re.view(np.int64))
    assert np.array_equal(a_im.view(np.int64), b_im.view(np.int64))


@pytest.mark.parametrize("k", [1, 3])
def test_nucleus_orbit_repeats_exactly_and_equals_the_oracle(k):
    # At a nucleus centre known to the working precision, Z_P squares to zero in the
    # truncated fixed point, so Z_{P+1} = c = Z_1 exactly and the host copies the cycle
    # instead of squaring (the 1e-1000 frame: 16,117 of 1,000,001 points computed); the
    # orbit must still be the oracle's directly iterated one, bit for bit, to the last point
    from fractodactyl_amd.scenes import scene
    sc = scene(k)
    c = sc["config"]
    v = fa.View(sc["center_re"], sc["center_im"], c["width"], 0.0)
    p = fa.Params(64, 36, max_iter=c["max_iter"])
    a_re, a_im, a_bits = _host_reference(v, p)
    b_re, b_im, b_bits = oracle.reference(sc["center_re"], sc["center_im"], c["width"], 0.0,
                                          64, 36, c["max_iter"], 1e10)
    assert a_bits == b_bits and len(a_re) == len(b_re) == c["max_iter"] + 1
    assert np.array_equal(a_re.view(np.int64), b_re.view(np.int64))
    assert np.array_equal(a_im.view(np.int64), b_im.view(np.int64))
    P = sc["nucleus"]["period"]
    assert a_re[P + 1] == a_re[1] and a_im[P + 1] == a_im[1]


@pytest.mark.parametrize("k,digits", [(1, 25), (3, 1010)])
def test_newton_nucleus_reproduces_the_scene_nuclei(k, digits):
    # fdg_nucleus from a truncated start converges to the nucleus the scene holds: config 1's
    # seeded Newton nucleus (period 1611, 256 bits, found by Python-integer Newton at 288
    # bits) and config 3's ladder nucleus from the reference's own rung list (period 16116,
    # 3461 bits, the reference's decimals carry ~3720 bits)
    from decimal import Decimal, getcontext
    from fractodactyl_amd.scenes import scene
    getcontext().prec = 1500
    sc = scene(k)
    nuc = sc["nucleus"]
    re0, im0 = (nuc["re"], nuc["im"]) if k == 1 else (sc["center_re"], sc["center_im"])
    bits = {1: 256, 3: 3461}[k]
    cut = lambda s: s[:s.index(".") + 1 + digits]
    r, i, its, step = fa.nucleus(cut(re0), cut(im0), nuc["period"], bits)
    assert its <= 10 and step < -(bits - 8)
    err = max(abs(Decimal(r) - Decimal(re0)), abs(Decimal(i) - Decimal(im0)))
    assert err < Decimal(2) ** -(bits - 8), float(err)


def test_newton_nucleus_refuses_what_it_cannot_converge():
    # a start outside the period-16116 minibrot's basin escapes before the period
    from fractodactyl_amd.scenes import scene
    sc = scene(3)
    with pytest.raises(fa.FdgError, match="escaped"):
        fa.nucleus(sc["center_re"][:40], sc["center_im"][:40], 16116, 3461)
    with pytest.raises(fa.FdgError):
        fa.nucleus("0.25", "0", 0, 256)  # no period


def test_host_reference_rejects_bad_requests():
    v, p = fa._view(fa.View("abc", "0", "1e-3")), fa._params(fa.Params(8, 8))
    ln, bits = C.c_uint64(), C.c_uint32()
    assert fa.lib().fdg_reference(None, C.byref(v), C.byref(p), None, None, 0, C.byref(ln),
                                  C.byref(bits)) == -1
    assert b"decimal" in fa.lib().fdg_last_error(None)
    v = fa._view(fa.View("0", "0", "1e-3"))
    p = fa._params(fa.Params(8, 8, ss=3))
    assert fa.lib().fdg_reference(None, C.byref(v), C.byref(p), None, None, 0, C.byref(ln),
                                  C.byref(bits)) == -1
    p = fa._params(fa.Params(8, 8, tier=fa.TIER_F64))
    v = fa._view(fa.View("0", "1", "1e-300"))
    assert fa.lib().fdg_reference(None, C.byref(v), C.byref(p), None, None, 0, C.byref(ln),
                                  C.byref(bits)) == -1


def test_writer_matches_golden_bytes():
    # write.rs / header.rs through the product's writer, on the golden frames' own columns
    for name, data in kern05()[:4]:
        h, cols = parse_fds(data)
        header = dict(view=fa.View(h["re"], h["im"], h["width"], h["rotation"]),
                      params=fa.Params(h["nx"], h["ny"], h["ss"], h["max_iter"],
                                       h["escape_radius"], h["columns"] & ~1),
                      kernel=h["kernel"])
        assert fa.write(header, cols, minor=0) == data, name


def test_a_failed_host_allocation_returns_enomem():
    # In a child process whose address space is capped a little above what it uses, a
    # host-only call that needs far more memory (the jet-AA rebuild of a 2^20 x 2^20 pixel
    # grid: its first per-pixel array is 1 TiB) must return FDG_ENOMEM with a message, not
    # throw into the caller or abort.  No column memory is touched before that array.
    import subprocess
    import sys
    code = f"""
import ctypes as C, resource, sys
import numpy as np
sys.path.insert(0, {ROOT!r})
import fractodactyl_amd as fa
L = fa.lib()
def vm():
    for l in open("/proc/self/status"):
        if l.startswith("VmSize:"):
            return int(l.split()[1]) * 1024
lim = vm() + (512 << 20)
resource.setrlimit(resource.RLIMIT_AS, (lim, lim))
cols = dict(**{{"class": np.zeros(4, np.uint8), "nu": np.zeros(4), "de": np.zeros(4, np.float32),
            "normal": np.zeros(4, np.uint16)}})
todo = np.zeros(4, np.uint8)
rc = L.fdg_jetaa_rebuild(C.byref(fa._cols_struct(cols)), 1 << 20, 1 << 20, fa._p(todo))
print(rc, L.fdg_last_error(None).decode())
"""
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    rc, msg = r.stdout.strip().split(" ", 1)
    assert int(rc) == -4, r.stdout  # FDG_ENOMEM
    assert "memory" in msg
