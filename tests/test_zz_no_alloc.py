"""fdg_render_orbit_noalloc (the entry point that allocates nothing: its buffers sit in a
static device array per GPU) against fdg_frame_open_orbit + run + read, its allocating
counterpart: every raw outcome bit for bit and the work counters equal, on two calls in a
row — the second smaller, on another tier — so the second runs over the first's leftovers
in the shared array."""
import ctypes as C

import numpy as np
import pytest

import fractodactyl_amd as fa
from fdsfile import kern05, parse_fds

pytestmark = pytest.mark.gpu

COLS = fa.COL_NU | fa.COL_DE | fa.COL_NORMAL
FIELDS = (("kind", np.uint8), ("n", np.uint64), ("zr", np.float64), ("zi", np.float64),
          ("dr", np.float64), ("di", np.float64), ("dexp", np.int64), ("ops", np.uint32))
COUNTERS = ("iterations", "equivalent", "blocks", "skipped", "plain_steps", "fallback_samples",
            "closed_form_samples", "max_sample_ops", "bla_levels", "nucleus_period")


def frame(name, nx=None, ny=None):
    h, _ = parse_fds(dict(kern05())[name])
    v = fa.View(h["re"], h["im"], h["width"], h["rotation"])
    p = fa.Params(nx or h["nx"], ny or h["ny"], 1, h["max_iter"], h["escape_radius"], COLS)
    return v, p


class DeviceArrays:
    """Caller-owned device arrays from the HIP runtime libfdg.so itself links (the same
    loaded libamdhip64.so.7: a second HIP runtime in the process, such as torch's bundled
    one, does not see the device once this one holds it)."""

    def __init__(self, sizes):
        fa.lib()
        self.hip = C.CDLL("libamdhip64.so.7")
        self.hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.hip.hipFree.argtypes = [C.c_void_p]
        self.hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.ptr = {}
        for k, nbytes in sizes.items():
            p = C.c_void_p()
            assert self.hip.hipMalloc(C.byref(p), max(nbytes, 8)) == 0, k
            self.ptr[k] = p.value

    def get(self, k, n, t):
        a = np.empty(n, t)
        assert self.hip.hipMemcpy(a.ctypes.data, self.ptr[k], a.nbytes, 2) == 0  # to host
        return a

    def free(self):
        for p in self.ptr.values():
            self.hip.hipFree(p)
        self.ptr = {}


def noalloc(ctx, v, p, route, orbit):
    re, im, bits = orbit
    n = p.nx * p.ny
    dev = DeviceArrays({k: n * np.dtype(t).itemsize for k, t in FIELDS})
    try:
        out = fa._Outcomes(*(dev.ptr[k] for k, _ in FIELDS))
        st = fa._Stats()
        vv, pp = fa._view(v), fa._params(p)
        print(f"noalloc call: {v.width} {p.nx}x{p.ny} route {route} orbit {len(re)}",
              flush=True)
        rc = fa.lib().fdg_render_orbit_noalloc(ctx._h, C.byref(vv), C.byref(pp), route,
                                               fa.EPS, fa._p(re), fa._p(im), len(re), bits,
                                               C.byref(out), C.byref(st))
        ctx._check(rc)
        got = {k: dev.get(k, n, t) for k, t in FIELDS}
    finally:
        dev.free()
    return got, {k: getattr(st, k) for k, _ in fa._Stats._fields_}


def allocating(ctx, v, p, route, orbit):
    with fa.ResidentFrame(v, p, route, ctx=ctx, orbit=orbit) as f:
        f.run()
        _, cols, st = f.read(raw=True)
    return cols["raw"], st


@pytest.mark.parametrize("route", [fa.ROUTE_BLA, fa.ROUTE_PLAIN])
def test_noalloc_equals_the_allocating_entry_point_twice(route):
    with fa.Context(0) as ctx:
        # call 1: fx tier, 128x72, 50k iterations; call 2: scaled tier, smaller grid
        for name, nx, ny in (("valley-1e-28", None, None), ("i-1e-300", 96, 54)):
            v, p = frame(name, nx, ny)
            orbit = fa.reference(v, p)
            assert len(orbit[0]) <= (1 << 17) + 2
            want, wst = allocating(ctx, v, p, route, orbit)
            got, gst = noalloc(ctx, v, p, route, orbit)
            # fdg_frame_read reports n = 0 for Unresolved samples; the raw kernel outcome
            # keeps the iterate it stopped at
            got["n"] = np.where(got["kind"] == fa.UNRESOLVED, 0, got["n"]).astype(np.uint64)
            for k, t in FIELDS:
                a, b = got[k], np.asarray(want[k]).view(t)
                assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), (name, k)
            for k in COUNTERS:
                assert gst[k] == wst[k], (name, k, gst[k], wst[k])
            assert (route == fa.ROUTE_PLAIN) or gst["blocks"] > 0


def test_noalloc_refuses_an_orbit_beyond_its_scratch():
    with fa.Context(0) as ctx:
        v, p = frame("valley-1e-28", 16, 9)
        re, im, bits = fa.reference(v, p)
        big = (1 << 17) + 3
        re2 = np.zeros(big)
        im2 = np.zeros(big)
        n = min(len(re), big)
        re2[:n], im2[:n] = re[:n], im[:n]
        dev = DeviceArrays({k: 8 * 16 * 9 for k, _ in FIELDS})
        try:
            out = fa._Outcomes(*(dev.ptr[k] for k, _ in FIELDS))
            rc = fa.lib().fdg_render_orbit_noalloc(ctx._h, C.byref(fa._view(v)),
                                                   C.byref(fa._params(p)), fa.ROUTE_BLA,
                                                   fa.EPS, fa._p(re2), fa._p(im2), big, bits,
                                                   C.byref(out), None)
        finally:
            dev.free()
        assert rc == -1
