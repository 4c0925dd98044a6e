"""CPU oracle for the deep-zoom perturbation path -- TEST INFRASTRUCTURE ONLY.

A ctypes binding to ``liboracle.so`` built from ``fdo.c``: a plain-C restatement of
fractodactyl's CPU renderer (crates/fd-kernel/src/{grid,sample,bla,reference,view,store}.rs,
crates/fd-fixed, crates/fd-samples).  Only tests, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may use it, and only as the checker or the timed CPU
baseline; the product library never routes through it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

COL_CLASS, COL_NU, COL_DE, COL_NORMAL = 1, 2, 4, 8
TIER_AUTO, TIER_F64, TIER_FIXED, TIER_SCALED = -1, 0, 1, 2
ESCAPED, INTERIOR, UNRESOLVED = 0, 1, 2
EPS_EXACT = 1.0 / 1125899906842624.0  # 2^-50, fd-cli/src/orbit.rs:16


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


class _Plane(C.Structure):
    _fields_ = [("c_re", C.c_double), ("c_im", C.c_double), ("h_m", C.c_double),
                ("h_e", C.c_int64), ("tier", C.c_int32), ("bits", C.c_uint64),
                ("cos_r", C.c_double), ("sin_r", C.c_double), ("half_x", C.c_double),
                ("half_y", C.c_double)]


class _Stats(C.Structure):
    _fields_ = [("reference_len", C.c_uint64), ("iterations", C.c_uint64),
                ("equivalent", C.c_uint64), ("blocks", C.c_uint64), ("skipped", C.c_uint64),
                ("fallback_samples", C.c_uint64), ("closed_form_samples", C.c_uint64),
                ("bla_levels", C.c_uint32), ("reference_seconds", C.c_double),
                ("bla_seconds", C.c_double), ("render_seconds", C.c_double)]


class _Outcomes(C.Structure):
    _fields_ = [("kind", C.c_void_p), ("n", C.c_void_p), ("zr", C.c_void_p),
                ("zi", C.c_void_p), ("dr", C.c_void_p), ("di", C.c_void_p),
                ("dexp", C.c_void_p)]


class _Columns(C.Structure):
    _fields_ = [("class_", C.c_void_p), ("nu", C.c_void_p), ("de", C.c_void_p),
                ("normal", C.c_void_p)]


class _SeriesConfig(C.Structure):
    _fields_ = [("start_order", C.c_uint32), ("max_order", C.c_uint32),
                ("grow_at", C.c_double), ("shift_px", C.c_double),
                ("derivative_rel", C.c_double)]


class _Series(C.Structure):
    _fields_ = [("n", C.c_uint64), ("rho", C.c_double), ("points", C.c_uint64),
                ("b", C.c_void_p), ("b_cap", C.c_uint32), ("b_len", C.c_uint32),
                ("state_bound", C.c_double), ("derivative_bound", C.c_double),
                ("shift_px", C.c_double), ("built_steps", C.c_uint64),
                ("coefficient_pairs", C.c_uint64), ("growth_steps", C.c_uint64),
                ("stop", C.c_int32)]


# series.rs:43-55 SeriesConfig::default
SERIES_DEFAULT = dict(start_order=8, max_order=64, grow_at=1e-4, shift_px=1e-8,
                      derivative_rel=1e-6)
SERIES_STOPS = ("reference", "radius", "non-finite", "escape", "budget", "stall")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.fdo_exp2i.restype = C.c_double
        L.fdo_exp2i.argtypes = [C.c_int64]
        L.fdo_plane_new.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, C.c_double, C.c_uint32,
                                    C.c_uint32, C.POINTER(_Plane), C.c_char_p, C.c_size_t]
        L.fdo_reference.restype = C.c_int64
        L.fdo_reference.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, C.c_double, C.c_uint32,
                                    C.c_uint32, C.c_uint64, C.c_double, C.c_int, C.c_void_p,
                                    C.c_void_p, C.c_uint64, C.POINTER(C.c_uint32), C.c_char_p,
                                    C.c_size_t]
        L.fdo_reference_fixed.restype = C.c_int64
        L.fdo_reference_fixed.argtypes = [C.c_char_p, C.c_char_p, C.c_uint64, C.c_uint64,
                                          C.c_void_p, C.c_void_p, C.c_uint64]
        L.fdo_direct.restype = C.c_int64
        L.fdo_direct.argtypes = [C.c_char_p, C.c_char_p, C.c_uint64, C.c_uint64, C.c_double,
                                 C.c_void_p, C.POINTER(C.c_int64)]
        L.fdo_zero_step.restype = None
        L.fdo_zero_step.argtypes = [C.c_double, C.c_double, C.c_int64, C.c_double, C.c_double,
                                    C.c_int64, C.c_void_p, C.POINTER(C.c_int64)]
        L.fdo_nucleus_period.restype = C.c_uint64
        L.fdo_nucleus_period.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64]
        L.fdo_bla_build.restype = C.c_int
        L.fdo_bla_build.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64, C.c_double, C.c_double,
                                    C.c_int64, C.c_void_p, C.c_uint64]
        L.fdo_render.restype = C.c_int
        L.fdo_render.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, C.c_double, C.c_uint32,
                                 C.c_uint32, C.c_uint32, C.c_uint64, C.c_double, C.c_uint32,
                                 C.c_int, C.c_int, C.c_double, C.c_int, C.c_void_p,
                                 C.POINTER(_Outcomes), C.POINTER(_Columns), C.POINTER(_Stats),
                                 C.c_char_p, C.c_size_t, C.c_char_p, C.c_size_t]
        L.fdo_render_with.restype = C.c_int
        L.fdo_render_with.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, C.c_double,
                                      C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint64, C.c_double,
                                      C.c_uint32, C.c_int, C.c_int, C.c_double, C.c_void_p,
                                      C.c_void_p, C.c_uint64, C.c_uint32, C.c_double, C.c_int64,
                                      C.c_int, C.c_void_p, C.POINTER(_Outcomes),
                                      C.POINTER(_Columns), C.POINTER(_Stats), C.c_char_p,
                                      C.c_size_t, C.c_char_p, C.c_size_t]
        L.fdo_series_build.restype = C.c_int
        L.fdo_series_build.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64, C.c_double,
                                       C.c_double, C.c_uint64, C.POINTER(_SeriesConfig),
                                       C.c_uint64, C.POINTER(_Series)]
        L.fdo_series_eval.restype = None
        L.fdo_series_eval.argtypes = [C.POINTER(_Series), C.c_int, C.c_double, C.c_double,
                                      C.c_void_p]
        L.fdo_render_series.restype = C.c_int
        L.fdo_render_series.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, C.c_double,
                                        C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint64,
                                        C.c_double, C.c_uint32, C.c_int, C.c_int, C.c_double,
                                        C.POINTER(_Series), C.c_int, C.c_void_p,
                                        C.POINTER(_Outcomes), C.POINTER(_Columns),
                                        C.POINTER(_Stats), C.c_char_p, C.c_size_t, C.c_char_p,
                                        C.c_size_t]
        L.fdo_jetaa_first.restype = None
        L.fdo_jetaa_first.argtypes = [C.c_uint64, C.c_uint64, C.c_void_p]
        L.fdo_jetaa_rebuild.restype = None
        L.fdo_jetaa_rebuild.argtypes = [C.c_void_p] * 4 + [C.c_uint64, C.c_uint64, C.c_void_p]
        L.fdo_jetaa_merge.restype = None
        L.fdo_jetaa_merge.argtypes = [C.c_void_p] * 9 + [C.c_uint64]
        L.fdo_sample_many.restype = None
        L.fdo_sample_many.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p, C.c_int,
                                      C.c_double, C.c_double, C.c_int, C.c_void_p, C.c_void_p,
                                      C.c_uint64, C.c_int, C.c_int64, C.c_int, C.c_uint64,
                                      C.c_double, C.c_uint64, C.c_void_p, C.c_void_p,
                                      C.c_void_p, C.c_void_p, C.c_void_p]
        L.fdo_error_radius.restype = None
        L.fdo_error_radius.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64, C.c_int64, C.c_double,
                                       C.c_double, C.c_void_p]
        L.fdo_guarded_build.restype = C.c_int
        L.fdo_guarded_build.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64,
                                        C.c_double, C.c_double, C.c_void_p, C.c_uint64]
        L.fdo_external_escape_ok.restype = C.c_int
        L.fdo_external_escape_ok.argtypes = [C.c_uint64, C.c_double, C.c_double, C.c_double,
                                             C.c_double, C.c_int64, C.c_double, C.c_double,
                                             C.c_double, C.c_int64]
        L.fdo_sample.restype = C.c_int
        L.fdo_sample.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p, C.c_int, C.c_int,
                                 C.c_double, C.c_double, C.c_double, C.c_double, C.c_int,
                                 C.c_int64, C.c_int, C.c_uint64, C.c_double, C.c_uint64,
                                 C.POINTER(C.c_uint64), C.c_void_p, C.c_void_p,
                                 C.POINTER(C.c_int64), C.POINTER(C.c_uint64),
                                 C.POINTER(C.c_uint64)]
        L.fdo_write_fds.restype = C.c_uint64
        L.fdo_write_fds.argtypes = [C.c_void_p, C.c_uint64, C.c_char_p, C.c_char_p, C.c_char_p,
                                    C.c_double, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint64,
                                    C.c_double, C.c_uint32, C.c_char_p, C.POINTER(_Columns)]
        L.fdo_bla_find_many.restype = C.c_int
        L.fdo_bla_find_many.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64, C.c_double,
                                        C.c_double, C.c_int64, C.c_void_p, C.c_uint64,
                                        C.c_double, C.c_uint64, C.c_void_p]
        L.fdo_fixed_parse.restype = C.c_int
        L.fdo_fixed_parse.argtypes = [C.c_char_p, C.c_uint64, C.c_void_p]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data


def exp2i(k):
    return lib().fdo_exp2i(k)


def plane(re, im, width, rotation, nx, ny):
    p = _Plane()
    err = C.create_string_buffer(256)
    if lib().fdo_plane_new(re.encode(), im.encode(), width.encode(), rotation, nx, ny,
                           C.byref(p), err, 256) != 0:
        raise ValueError(err.value.decode())
    return {k: getattr(p, k) for k, _ in _Plane._fields_}


def reference(re, im, width, rotation, nx, ny, max_iter, escape_radius, tier=TIER_AUTO):
    """grid.rs `reference`: (re, im, bits) of the orbit the CPU renderer uses."""
    err = C.create_string_buffer(256)
    bits = C.c_uint32()
    n = lib().fdo_reference(re.encode(), im.encode(), width.encode(), rotation, nx, ny, max_iter,
                            escape_radius, tier, None, None, 0, C.byref(bits), err, 256)
    if n < 0:
        raise ValueError(err.value.decode())
    zr = np.zeros(n, np.float64)
    zi = np.zeros(n, np.float64)
    lib().fdo_reference(re.encode(), im.encode(), width.encode(), rotation, nx, ny, max_iter,
                        escape_radius, tier, _p(zr), _p(zi), n, C.byref(bits), err, 256)
    return zr, zi, bits.value


def reference_fixed(re, im, bits, max_iter):
    cap = min(max_iter, (1 << 20) - 1) + 1
    zr = np.zeros(cap, np.float64)
    zi = np.zeros(cap, np.float64)
    n = lib().fdo_reference_fixed(re.encode(), im.encode(), bits, max_iter, _p(zr), _p(zi), cap)
    if n < 0:
        raise ValueError("bad decimal")
    return zr[:n].copy(), zi[:n].copy()


def nucleus_period(zr, zi):
    zr = np.ascontiguousarray(zr, np.float64)
    zi = np.ascontiguousarray(zi, np.float64)
    return int(lib().fdo_nucleus_period(_p(zr), _p(zi), len(zr)))


def bla_build(zr, zi, eps, dc_max, dc_exp=0):
    """Bla::build_scaled: list of (count, 7) float64 arrays, one per level 1, 2, ..."""
    zr = np.ascontiguousarray(zr, np.float64)
    zi = np.ascontiguousarray(zi, np.float64)
    steps = max(len(zr) - 2, 0)
    total = sum(steps >> j for j in range(1, 64))
    out = np.zeros((max(total, 1), 7), np.float64)
    nl = lib().fdo_bla_build(_p(zr), _p(zi), len(zr), eps, dc_max, dc_exp, _p(out), total)
    if nl < 0:
        raise ValueError("BLA needs 0 < eps <= 2^-24 and a finite dc_max > 0")
    levels, off = [], 0
    for j in range(1, nl + 1):
        c = steps >> j
        levels.append(out[off:off + c].copy())
        off += c
    return levels


class PackedLevels:
    """bla_build()'s levels concatenated once, for many sample() calls over one table."""

    def __init__(self, levels):
        self.blocks = np.ascontiguousarray(np.concatenate(levels, axis=0)) if levels else None
        self.count = len(levels)


def sample(zr, zi, ar, ai, max_iter, r2, c=None, extended=False, offset_exp=0,
           derivative=True, bla_levels=None, nucleus=0):
    """sample.rs iterate for one sample; returns a dict with the outcome."""
    zr = np.ascontiguousarray(zr, np.float64)
    zi = np.ascontiguousarray(zi, np.float64)
    blocks = None
    nl = 0
    if isinstance(bla_levels, PackedLevels):
        blocks, nl = bla_levels.blocks, bla_levels.count
    elif bla_levels:
        blocks = np.ascontiguousarray(np.concatenate(bla_levels, axis=0))
        nl = len(bla_levels)
    n = C.c_uint64()
    z = np.zeros(2)
    d = np.zeros(2)
    dexp = C.c_int64()
    nb = C.c_uint64()
    ns = C.c_uint64()
    kind = lib().fdo_sample(_p(zr), _p(zi), len(zr), _p(blocks), nl, c is not None,
                            c[0] if c else 0.0, c[1] if c else 0.0, ar, ai, int(extended),
                            offset_exp, int(derivative), max_iter, r2, nucleus, C.byref(n),
                            _p(z), _p(d), C.byref(dexp), C.byref(nb), C.byref(ns))
    return dict(kind=kind, n=n.value, zr=z[0], zi=z[1], dr=d[0], di=d[1], dexp=dexp.value,
                blocks=nb.value, skipped=ns.value)


def error_radius(zr, zi, bits, c_re, c_im):
    """reference.rs Reference::error_radius; bits None for an f64 orbit."""
    zr = np.ascontiguousarray(zr, np.float64)
    zi = np.ascontiguousarray(zi, np.float64)
    q = np.zeros(zr.size)
    lib().fdo_error_radius(_p(zr), _p(zi), zr.size, -1 if bits is None else int(bits), c_re,
                           c_im, _p(q))
    return q


def guarded_build(zr, zi, q, eps, dc_max):
    """bla/guarded.rs GuardedBla::build: a list of (count, 9) arrays per level (a.re a.im
    b.re b.im r alpha beta dalpha dbeta); raises ValueError where the reference refuses."""
    zr = np.ascontiguousarray(zr, np.float64)
    zi = np.ascontiguousarray(zi, np.float64)
    q = np.ascontiguousarray(q, np.float64)
    steps = max(zr.size - 2, 0)
    cap = max(steps, 1)
    out = np.zeros((cap, 9))
    nl = lib().fdo_guarded_build(_p(zr), _p(zi), _p(q), zr.size, eps, dc_max, _p(out), cap)
    if nl < 0:
        raise ValueError("guarded BLA refused")
    levels, off = [], 0
    for j in range(1, nl + 1):
        c = steps >> j
        levels.append(out[off:off + c].copy())
        off += c
    return levels


def direct(re, im, bits, max_iter, r2):
    """Direct iteration of z^2 + c at `bits` fraction bits (fdo_direct): (n, z_n, dz/dc at n
    as (re, im, e), |z_{n-1}|^2, |z_n|^2); n = -1 when |z|^2 stays <= r2 through max_iter."""
    out = np.zeros(6)
    e = C.c_int64()
    n = lib().fdo_direct(re.encode(), im.encode(), bits, max_iter, r2, _p(out), C.byref(e))
    if n == -2:
        raise ValueError("bad decimal")
    return int(n), complex(out[0], out[1]), (float(out[2]), float(out[3]), e.value), \
        float(out[4]), float(out[5])


def zero_step(w, a):
    """sample.rs:1119-1122 zero_step: the scaled step from Z_0 = 0, z' = w^2 2^2e + a 2^ae,
    for w = (re, im, e) and a = (re, im, e); returns (re, im, e)."""
    out = np.zeros(2)
    oe = C.c_int64()
    lib().fdo_zero_step(float(w[0]), float(w[1]), int(w[2]), float(a[0]), float(a[1]),
                        int(a[2]), _p(out), C.byref(oe))
    return float(out[0]), float(out[1]), oe.value


def external_escape_ok(n, z, d, dexp, ez, ed, px_m, px_e):
    """store.rs external_escape_ok for an escaped state (n, z, dz/dc = d 2^dexp)."""
    return bool(lib().fdo_external_escape_ok(n, z[0], z[1], d[0], d[1], dexp, ez, ed, px_m,
                                             px_e))


def sample_many(zr, zi, ar, ai, max_iter, r2, c=None, extended=False, offset_exp=0,
                derivative=True, bla_levels=None, nucleus=0):
    """sample() at many offsets (arrays ar, ai) over one table built once; returns a dict
    of arrays kind, n, zr, zi, dr, di, dexp."""
    zr = np.ascontiguousarray(zr, np.float64)
    zi = np.ascontiguousarray(zi, np.float64)
    ar = np.ascontiguousarray(ar, np.float64)
    ai = np.ascontiguousarray(ai, np.float64)
    pk = bla_levels if isinstance(bla_levels, PackedLevels) else PackedLevels(bla_levels or [])
    m = ar.size
    kind = np.zeros(m, np.uint8)
    n = np.zeros(m, np.uint64)
    z = np.zeros((m, 2))
    d = np.zeros((m, 2))
    dexp = np.zeros(m, np.int64)
    lib().fdo_sample_many(_p(zr), _p(zi), len(zr), _p(pk.blocks), pk.count,
                          c[0] if c else 0.0, c[1] if c else 0.0, c is not None, _p(ar), _p(ai),
                          m, int(extended), offset_exp, int(derivative), max_iter, r2, nucleus,
                          _p(kind), _p(n), _p(z), _p(d), _p(dexp))
    return dict(kind=kind, n=n, zr=z[:, 0], zi=z[:, 1], dr=d[:, 0], di=d[:, 1], dexp=dexp)


def series_build(zr, zi, rho, pixel, max_iter, config=None, stalled_steps=32):
    """series.rs Series::build (stalled_steps 32; 0 = the unbounded control) over an orbit.
    Returns a dict: n, rho, points, b (complex coefficients b[1..]), bounds, stop."""
    zr = np.ascontiguousarray(zr, np.float64)
    zi = np.ascontiguousarray(zi, np.float64)
    cfg = dict(SERIES_DEFAULT, **(config or {}))
    c = _SeriesConfig(**cfg)
    cap = max(int(cfg["max_order"]), 1)
    buf = np.zeros(2 * cap, np.float64)
    out = _Series(b=_p(buf), b_cap=cap)
    if lib().fdo_series_build(_p(zr), _p(zi), len(zr), rho, pixel, max_iter, C.byref(c),
                              stalled_steps, C.byref(out)) != 0:
        raise ValueError("coefficient buffer too small")
    b = buf[:2 * out.b_len].reshape(-1, 2)
    return dict(n=out.n, rho=out.rho, points=out.points, b=b[:, 0] + 1j * b[:, 1],
                state_bound=out.state_bound, derivative_bound=out.derivative_bound,
                shift_px=out.shift_px, built_steps=out.built_steps,
                coefficient_pairs=out.coefficient_pairs, growth_steps=out.growth_steps,
                stop=SERIES_STOPS[out.stop])


def _series_struct(sr):
    b = np.ascontiguousarray(np.stack([np.real(sr["b"]), np.imag(sr["b"])], axis=1)
                             .reshape(-1), np.float64)
    st = _Series(n=sr["n"], rho=sr["rho"], points=sr["points"], b=_p(b), b_cap=len(sr["b"]),
                 b_len=len(sr["b"]), state_bound=sr["state_bound"],
                 derivative_bound=sr["derivative_bound"], shift_px=sr["shift_px"])
    return st, b


def series_eval(sr, ar, ai, derivative=True):
    """Series::eval: (delta, dz/dc) as complex numbers at the offset (ar, ai)."""
    st, keep = _series_struct(sr)
    out = np.zeros(4)
    lib().fdo_series_eval(C.byref(st), int(derivative), ar, ai, _p(out))
    return complex(out[0], out[1]), complex(out[2], out[3])


def render(re, im, width, rotation, nx, ny, max_iter, escape_radius=1e10,
           columns=COL_NU | COL_DE | COL_NORMAL, ss=1, tier=TIER_AUTO, bla=False,
           bla_eps=EPS_EXACT, threads=None, mask=None, outcomes=True, orbit=None,
           table_dc=None, series=None):
    """grid.rs render_with / render_bla (table built per frame as `fd control --bla
    per-frame` does, control.rs:356-366).  Returns a dict of numpy columns, outcomes,
    stats and the header kernel id.  orbit = (re, im, bits) supplies the reference orbit
    (grid.rs:95-109); table_dc = (dc_max, dc_exp) builds the table for that |dc| bound
    instead of the frame's own (after grid.rs:378-413 bla_checks).  series = a
    series_build() result renders with its start (grid.rs:218-226 render_series, over the
    frame's own orbit)."""
    n = nx * ny
    threads = threads or os.cpu_count() or 1
    cols = {"class": np.zeros(n, np.uint8)}
    if columns & COL_NU:
        cols["nu"] = np.zeros(n, np.float64)
    if columns & COL_DE:
        cols["de"] = np.zeros(n, np.float32)
    if columns & COL_NORMAL:
        cols["normal"] = np.zeros(n, np.uint16)
    cs = _Columns(_p(cols["class"]), _p(cols.get("nu")), _p(cols.get("de")),
                  _p(cols.get("normal")))
    oc = None
    res = {}
    if outcomes:
        res = {"kind": np.zeros(n, np.uint8), "n": np.zeros(n, np.uint64),
               "zr": np.zeros(n), "zi": np.zeros(n), "dr": np.zeros(n), "di": np.zeros(n),
               "dexp": np.zeros(n, np.int64)}
        oc = _Outcomes(*[_p(res[k]) for k in ("kind", "n", "zr", "zi", "dr", "di", "dexp")])
    st = _Stats()
    kid = C.create_string_buffer(128)
    err = C.create_string_buffer(256)
    mk = None if mask is None else np.ascontiguousarray(mask, np.uint8)
    ore = oim = None
    olen = obits = 0
    if orbit is not None:
        ore = np.ascontiguousarray(orbit[0], np.float64)
        oim = np.ascontiguousarray(orbit[1], np.float64)
        olen, obits = len(ore), int(orbit[2])
    tdc, tde = table_dc if table_dc is not None else (0.0, 0)
    if series is not None:
        sst, keep = _series_struct(series)
        rc = lib().fdo_render_series(re.encode(), im.encode(), width.encode(), rotation, nx, ny,
                                     ss, max_iter, escape_radius, columns, tier, int(bla),
                                     bla_eps, C.byref(sst), threads, _p(mk),
                                     C.byref(oc) if oc is not None else None, C.byref(cs),
                                     C.byref(st), kid, 128, err, 256)
    else:
        rc = lib().fdo_render_with(re.encode(), im.encode(), width.encode(), rotation, nx, ny,
                                   ss, max_iter, escape_radius, columns, tier, int(bla), bla_eps,
                                   _p(ore), _p(oim), olen, obits, tdc, tde, threads, _p(mk),
                                   C.byref(oc) if oc is not None else None, C.byref(cs),
                                   C.byref(st), kid, 128, err, 256)
    if rc != 0:
        raise ValueError(err.value.decode())
    return dict(columns=cols, outcomes=res, kernel=kid.value.decode(),
                stats={k: getattr(st, k) for k, _ in _Stats._fields_},
                view=(re, im, width, rotation), grid=(nx, ny, ss), max_iter=max_iter,
                escape_radius=escape_radius, colmask=columns | COL_CLASS)


def write_fds(frame, minor=1):
    """fd_samples::write of a render() result; minor rewrites bytes 10..12."""
    re, im, width, rot = frame["view"]
    nx, ny, ss = frame["grid"]
    cols = frame["columns"]
    cs = _Columns(_p(cols["class"]), _p(cols.get("nu")), _p(cols.get("de")),
                  _p(cols.get("normal")))
    args = (re.encode(), im.encode(), width.encode(), rot, nx, ny, ss, frame["max_iter"],
            frame["escape_radius"], frame["colmask"], frame["kernel"].encode(), C.byref(cs))
    size = lib().fdo_write_fds(None, 0, *args)
    buf = np.zeros(size, np.uint8)
    lib().fdo_write_fds(_p(buf), size, *args)
    b = bytearray(buf.tobytes())
    b[10:12] = int(minor).to_bytes(2, "little")
    return bytes(b)


def bla_find(zr, zi, eps, dc_max, ms, d2, budget, dc_exp=0):
    """Bla::find at each m: the block length (2 << level), or None."""
    zr = np.ascontiguousarray(zr, np.float64)
    zi = np.ascontiguousarray(zi, np.float64)
    ms = np.ascontiguousarray(ms, np.uint64)
    out = np.zeros(len(ms), np.int32)
    if lib().fdo_bla_find_many(_p(zr), _p(zi), len(zr), eps, dc_max, dc_exp, _p(ms), len(ms),
                               d2, min(int(budget), (1 << 64) - 1), _p(out)) < 0:
        raise ValueError("bad eps / dc_max")
    return [None if v < 0 else 2 << int(v) for v in out]


def fixed_parse(s, limbs):
    """Fixed::parse: (neg, little-endian limbs)."""
    out = np.zeros(limbs, np.uint64)
    neg = lib().fdo_fixed_parse(s.encode(), limbs, _p(out))
    if neg < 0:
        raise ValueError("not a decimal number / too large: %r" % s)
    return bool(neg), [int(v) for v in out]


def jetaa_first(w, h):
    """jetaa.rs `first`: the pass-1 mask of a w x h pixel ss 2 grid (uint8)."""
    m = np.zeros(4 * w * h, np.uint8)
    lib().fdo_jetaa_first(w, h, _p(m))
    return m


def jetaa_rebuild(cols, w, h):
    """jetaa.rs `rebuild` in place on a columns dict (class, nu, de, normal); returns the
    pass-2 mask (uint8)."""
    for k, t in (("class", np.uint8), ("nu", np.float64), ("de", np.float32),
                 ("normal", np.uint16)):
        assert cols[k].dtype == t and cols[k].flags.c_contiguous and cols[k].size == 4 * w * h
    todo = np.zeros(4 * w * h, np.uint8)
    lib().fdo_jetaa_rebuild(_p(cols["class"]), _p(cols["nu"]), _p(cols["de"]),
                            _p(cols["normal"]), w, h, _p(todo))
    return todo


def jetaa_merge(to, frm, mask):
    """jetaa.rs `merge`: copy the masked samples of `frm` into `to` (columns dicts)."""
    m = np.ascontiguousarray(mask, np.uint8)
    lib().fdo_jetaa_merge(*[_p(to[k]) for k in ("class", "nu", "de", "normal")],
                          *[_p(frm[k]) for k in ("class", "nu", "de", "normal")], _p(m), m.size)
