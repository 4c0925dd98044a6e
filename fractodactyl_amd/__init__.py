"""fractodactyl_amd: fractodactyl's deep-zoom perturbation/BLA path on AMD MI355X.

A ctypes mirror of the C ABI in ``include/fdg.h`` (``libfdg.so``, built in-tree by
``python -m fractodactyl_amd.build``).  The API follows the reference's GPU route and
sample API:

* :func:`enabled` -- ``perturb_gpu::enabled`` (crates/fd-cli/src/perturb_gpu/frame.rs:62-68)
* :func:`render` / :func:`render_bla` -- ``perturb_gpu::render`` / ``render_bla``
  (frame.rs:70-84): ``(header, samples, stats)`` for a :class:`View` and :class:`Params`
* :func:`reference` -- ``fd_kernel::reference`` (crates/fd-kernel/src/grid.rs:84-87)
* :func:`build_bla` -- ``Bla::build_scaled`` (crates/fd-kernel/src/bla.rs:161-181), on the GPU
* :func:`write` -- ``fd_samples::write`` (crates/fd-samples/src/write.rs:8-57)

There is no CPU fallback: without the HIP library or a device these raise.
"""
import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# FDG_LIB points at another build of the same library (e.g. an experimental variant).
LIB_PATH = os.environ.get("FDG_LIB") or os.path.join(HERE, "libfdg.so")

COL_CLASS, COL_NU, COL_DE, COL_NORMAL, COL_BOUND, COL_STRIPE = 1, 2, 4, 8, 16, 32
TIER_AUTO, TIER_F64, TIER_FIXED, TIER_SCALED = -1, 0, 1, 2
ROUTE_PLAIN, ROUTE_BLA = 0, 1
FORM_F64 = 0
FORM_FE32 = 1  # FP32 mantissas + int32 exponents: close to, not equal to, FORM_F64
FORM_MIXED = 2  # FE32 pass, then its escaped samples recomputed in FP64
ESCAPED, INTERIOR, UNRESOLVED = 0, 1, 2
EPS = 1.0 / 1125899906842624.0          # 2^-50, crates/fd-cli/src/orbit.rs:16
FILM_EPS = 1.0 / (1 << 26)              # orbit.rs:22
ERRORS = {-1: "invalid request", -2: "declined", -3: "device failure", -4: "out of memory"}


class FdgError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


@dataclass
class View:
    """fd_samples::View (crates/fd-samples/src/header.rs:263-273)."""
    center_re: str
    center_im: str
    width: str
    rotation: float = 0.0


@dataclass
class Params:
    """fd_kernel::Params (crates/fd-kernel/src/grid.rs:17-38)."""
    nx: int
    ny: int
    ss: int = 1
    max_iter: int = 100_000
    escape_radius: float = 1e10
    columns: int = COL_NU | COL_DE | COL_NORMAL
    tier: int = TIER_AUTO


class _View(C.Structure):
    _fields_ = [("center_re", C.c_char_p), ("center_im", C.c_char_p), ("width", C.c_char_p),
                ("rotation", C.c_double)]


class _Params(C.Structure):
    _fields_ = [("nx", C.c_uint32), ("ny", C.c_uint32), ("ss", C.c_uint32),
                ("max_iter", C.c_uint64), ("escape_radius", C.c_double),
                ("columns", C.c_uint32), ("tier", C.c_int32)]


class _Columns(C.Structure):
    _fields_ = [("class_", C.c_void_p), ("nu", C.c_void_p), ("de", C.c_void_p),
                ("normal", C.c_void_p)]


class _Outcomes(C.Structure):
    _fields_ = [("kind", C.c_void_p), ("n", C.c_void_p), ("zr", C.c_void_p),
                ("zi", C.c_void_p), ("dr", C.c_void_p), ("di", C.c_void_p),
                ("dexp", C.c_void_p), ("ops", C.c_void_p)]


class _EscapedState(C.Structure):
    _fields_ = [("n", C.c_uint64), ("zr", C.c_double), ("zi", C.c_double), ("dr", C.c_double),
                ("di", C.c_double), ("dexp", C.c_int64)]


class _Launch(C.Structure):
    _fields_ = [("groups_per_cu", C.c_uint32), ("claim_div", C.c_uint32),
                ("claim_cap", C.c_uint32), ("refill", C.c_uint32)]


class _Stats(C.Structure):
    _fields_ = [("reference_len", C.c_uint64), ("reference_bits", C.c_uint32),
                ("bla_levels", C.c_uint32), ("nucleus_period", C.c_uint64),
                ("iterations", C.c_uint64), ("equivalent", C.c_uint64),
                ("blocks", C.c_uint64), ("skipped", C.c_uint64), ("plain_steps", C.c_uint64),
                ("fallback_samples", C.c_uint64), ("closed_form_samples", C.c_uint64),
                ("max_sample_ops", C.c_uint64), ("wave_iterations", C.c_uint64),
                ("lane_iterations", C.c_uint64), ("reference_seconds", C.c_double), ("bla_seconds", C.c_double),
                ("kernel_seconds", C.c_double), ("transfer_seconds", C.c_double),
                ("encode_seconds", C.c_double), ("uncertified_samples", C.c_uint64),
                ("reference_period", C.c_uint64), ("reference_repeat_at", C.c_uint64)]


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


SERIES_STOPS = ("reference", "radius", "non-finite", "escape", "budget", "stall")

# Every entry point of include/fdg.h with its ctypes signature.
SIGNATURES = {
    "fdg_abi_version": (C.c_int, []),
    "fdg_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int]),
    "fdg_destroy": (None, [C.c_void_p]),
    "fdg_last_error": (C.c_char_p, [C.c_void_p]),
    "fdg_enabled": (C.c_int, [C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "fdg_render": (C.c_int, [C.c_void_p, C.POINTER(_View), C.POINTER(_Params), C.c_int, C.c_int,
                             C.c_double, C.POINTER(_Columns), C.POINTER(_Outcomes),
                             C.POINTER(_Stats), C.c_char_p, C.c_size_t]),
    "fdg_nucleus": (C.c_int, [C.c_void_p, C.c_char_p, C.c_char_p, C.c_uint64, C.c_uint32,
                              C.c_uint32, C.c_char_p, C.c_char_p, C.c_size_t,
                              C.POINTER(C.c_uint32), C.POINTER(C.c_double)]),
    "fdg_reference": (C.c_int, [C.c_void_p, C.POINTER(_View), C.POINTER(_Params), C.c_void_p,
                                C.c_void_p, C.c_uint64, C.POINTER(C.c_uint64),
                                C.POINTER(C.c_uint32)]),
    "fdg_build_bla": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64, C.c_double,
                                C.c_double, C.c_int64, C.c_void_p, C.c_uint64,
                                C.POINTER(C.c_uint32)]),
    "fdg_frame_open": (C.c_int, [C.c_void_p, C.POINTER(_View), C.POINTER(_Params), C.c_int,
                                 C.c_int, C.c_double, C.POINTER(C.c_void_p)]),
    "fdg_frame_run": (C.c_int, [C.c_void_p, C.POINTER(_Stats)]),
    "fdg_frame_read": (C.c_int, [C.c_void_p, C.POINTER(_Columns), C.POINTER(_Outcomes),
                                 C.POINTER(_Stats), C.c_char_p, C.c_size_t]),
    "fdg_frame_close": (None, [C.c_void_p]),
    "fdg_frame_read_columns": (C.c_int, [C.c_void_p, C.POINTER(_Columns), C.POINTER(_Stats),
                                         C.c_char_p, C.c_size_t]),
    "fdg_frame_mask": (C.c_int, [C.c_void_p, C.c_void_p]),
    "fdg_frame_launch": (C.c_int, [C.c_void_p, C.POINTER(_Launch)]),
    "fdg_frame_tune": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(_Launch),
                                 C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "fdg_render_masked": (C.c_int, [C.c_void_p, C.POINTER(_View), C.POINTER(_Params), C.c_int,
                                    C.c_int, C.c_double, C.c_void_p, C.POINTER(_Columns),
                                    C.POINTER(_Outcomes), C.POINTER(_Stats), C.c_char_p,
                                    C.c_size_t]),
    "fdg_series_config_default": (None, [C.POINTER(_SeriesConfig)]),
    "fdg_series_build": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_double, C.c_double,
                                   C.c_uint64, C.POINTER(_SeriesConfig), C.POINTER(_Series)]),
    "fdg_bla_dc_max": (C.c_int, [C.POINTER(_View), C.POINTER(_Params),
                                 C.POINTER(C.c_double)]),
    "fdg_frame_series": (C.c_int, [C.c_void_p, C.POINTER(_Series)]),
    "fdg_frame_series_prepare": (C.c_int, [C.c_void_p, C.POINTER(_SeriesConfig),
                                           C.POINTER(_Series), C.POINTER(C.c_int)]),
    "fdg_render_series": (C.c_int, [C.c_void_p, C.POINTER(_View), C.POINTER(_Params), C.c_int,
                                    C.c_double, C.POINTER(_Series), C.c_void_p,
                                    C.POINTER(_Columns), C.POINTER(_Outcomes),
                                    C.POINTER(_Stats), C.c_char_p, C.c_size_t]),
    "fdg_jetaa_first": (C.c_int, [C.c_uint32, C.c_uint32, C.c_void_p]),
    "fdg_jetaa_rebuild": (C.c_int, [C.POINTER(_Columns), C.c_uint32, C.c_uint32, C.c_void_p]),
    "fdg_jetaa_merge": (C.c_int, [C.POINTER(_Columns), C.POINTER(_Columns), C.c_void_p,
                                  C.c_uint64]),
    "fdg_frame_jet": (C.c_int, [C.c_void_p, C.POINTER(_Columns), C.POINTER(_Stats), C.c_char_p,
                                C.c_size_t, C.POINTER(C.c_uint64)]),
    "fdg_frame_open_orbit": (C.c_int, [C.c_void_p, C.POINTER(_View), C.POINTER(_Params),
                                       C.c_int, C.c_int, C.c_double, C.c_void_p, C.c_void_p,
                                       C.c_uint64, C.c_uint32, C.POINTER(C.c_void_p)]),
    "fdg_render_orbit_noalloc": (C.c_int, [C.c_void_p, C.POINTER(_View), C.POINTER(_Params),
                                           C.c_int, C.c_double, C.c_void_p, C.c_void_p,
                                           C.c_uint64, C.c_uint32, C.POINTER(_Outcomes),
                                           C.POINTER(_Stats)]),
    "fdg_error_radius": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_int64, C.c_double,
                                   C.c_double, C.c_void_p]),
    "fdg_guarded_bla": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64, C.c_double,
                                  C.c_double, C.c_void_p, C.c_uint64, C.POINTER(C.c_uint32)]),
    "fdg_external_escape_ok": (C.c_int, [C.POINTER(_EscapedState), C.c_double, C.c_double,
                                         C.c_double, C.c_int64]),
    "fdg_overlay_escaped": (C.c_int, [C.POINTER(_View), C.POINTER(_Params), C.POINTER(_Columns),
                                      C.c_void_p, C.c_void_p]),
    "fdg_chunk_id": (C.c_int, [C.c_void_p, C.c_uint64, C.c_void_p]),
    "fdg_chunk_orbit_slab": (C.c_int64, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint32,
                                         C.c_uint32, C.c_uint32, C.c_void_p, C.c_uint64]),
    "fdg_chunk_orbit_slab_read": (C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(C.c_uint64),
                                            C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                            C.POINTER(C.c_uint32), C.c_void_p, C.c_void_p,
                                            C.c_uint64]),
    "fdg_chunk_orbit_manifest": (C.c_int64, [C.c_char_p, C.c_char_p, C.c_uint32, C.c_void_p,
                                             C.c_uint64, C.c_void_p, C.c_uint64]),
    "fdg_chunk_orbit_manifest_read": (C.c_int, [C.c_void_p, C.c_uint64, C.c_char_p,
                                                C.c_char_p, C.c_uint64,
                                                C.POINTER(C.c_uint32), C.c_void_p, C.c_uint64,
                                                C.POINTER(C.c_uint64)]),
    "fdg_chunk_bla": (C.c_int64, [C.c_void_p, C.c_double, C.c_double, C.c_uint64, C.c_uint32,
                                  C.c_void_p, C.c_void_p, C.c_uint64]),
    "fdg_chunk_bla_read": (C.c_int, [C.c_void_p, C.c_uint64, C.c_void_p,
                                     C.POINTER(C.c_double), C.POINTER(C.c_double),
                                     C.POINTER(C.c_uint64), C.POINTER(C.c_uint32), C.c_void_p,
                                     C.c_uint64]),
    "fdg_film_open": (C.c_int, [C.c_void_p, C.POINTER(_View), C.c_uint32, C.POINTER(_Params),
                                C.c_int, C.c_int, C.c_double, C.c_int, C.POINTER(C.c_void_p)]),
    "fdg_film_run": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.POINTER(_Stats)]),
    "fdg_film_read": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(_Columns),
                                C.POINTER(_Outcomes), C.POINTER(_Stats), C.c_char_p,
                                C.c_size_t]),
    "fdg_film_close": (None, [C.c_void_p]),
    "fdg_film_mask": (C.c_int, [C.c_void_p, C.c_void_p]),
    "fdg_film_jet": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(C.c_uint64)]),
    "fdg_jacobian": (C.c_int, [C.POINTER(_View), C.POINTER(_Params), C.POINTER(_Outcomes),
                               C.c_void_p, C.c_void_p]),
    "fdg_write_fds": (C.c_int64, [C.POINTER(_View), C.POINTER(_Params), C.c_char_p,
                                  C.POINTER(_Columns), C.c_void_p, C.c_uint64]),
}

_lib = None


def lib():
    """The loaded libfdg.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FdgError(-3, f"{LIB_PATH} is missing: run python -m fractodactyl_amd.build")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data


def _view(v):
    return _View(v.center_re.encode(), v.center_im.encode(), v.width.encode(), float(v.rotation))


def _params(p):
    return _Params(p.nx, p.ny, p.ss, p.max_iter, p.escape_radius, p.columns, p.tier)


class Context:
    """One HIP device, one stream and its resident buffers (fdg_create/fdg_destroy)."""

    def __init__(self, device=0):
        h = C.c_void_p()
        rc = lib().fdg_create(C.byref(h), device)
        if rc != 0:
            raise FdgError(rc, lib().fdg_last_error(None).decode())
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().fdg_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc < 0:
            raise FdgError(rc, lib().fdg_last_error(self._h).decode())
        return rc


_default = None


def default_context():
    global _default
    if _default is None:
        _default = Context(0)
    return _default


def enabled(value):
    """perturb_gpu::enabled: ``(on, route)`` for "on" / "bla" / "off" / None."""
    e, r = C.c_int(0), C.c_int(0)
    if lib().fdg_enabled(None if value is None else value.encode(), C.byref(e), C.byref(r)) != 0:
        raise ValueError(f"--gpu-perturb: expected on, bla or off, got {value!r}")
    return bool(e.value), r.value


def render(view, params, route=ROUTE_PLAIN, form=FORM_F64, bla_eps=EPS, ctx=None, raw=False):
    """perturb_gpu::render (frame.rs:70-76): ``(header, samples, stats)``.

    ``header`` is a dict of the .fds header fields, ``samples`` a dict of numpy columns
    (``class`` always; ``nu``/``de``/``normal`` as requested) plus the raw outcomes under
    ``"raw"`` when ``raw`` is set."""
    ctx = ctx or default_context()
    n = params.nx * params.ny
    cols = {"class": np.zeros(n, np.uint8)}
    if params.columns & COL_NU:
        cols["nu"] = np.zeros(n, np.float64)
    if params.columns & COL_DE:
        cols["de"] = np.zeros(n, np.float32)
    if params.columns & COL_NORMAL:
        cols["normal"] = np.zeros(n, np.uint16)
    cs = _Columns(_p(cols["class"]), _p(cols.get("nu")), _p(cols.get("de")),
                  _p(cols.get("normal")))
    oc = None
    if raw:
        rw = {"kind": np.zeros(n, np.uint8), "n": np.zeros(n, np.uint64), "zr": np.zeros(n),
              "zi": np.zeros(n), "dr": np.zeros(n), "di": np.zeros(n),
              "dexp": np.zeros(n, np.int64), "ops": np.zeros(n, np.uint32)}
        oc = _Outcomes(*[_p(rw[k]) for k in ("kind", "n", "zr", "zi", "dr", "di", "dexp",
                                             "ops")])
        cols["raw"] = rw
    st = _Stats()
    kid = C.create_string_buffer(256)
    ctx._check(lib().fdg_render(ctx._h, C.byref(_view(view)), C.byref(_params(params)), route,
                                form, bla_eps, C.byref(cs),
                                C.byref(oc) if oc is not None else None, C.byref(st), kid, 256))
    header = dict(view=view, params=params, kernel=kid.value.decode(),
                  columns=params.columns | COL_CLASS)
    return header, cols, {k: getattr(st, k) for k, _ in _Stats._fields_}


def _out_buffers(p, raw):
    """Caller-owned column (and raw outcome) arrays for one frame, and their C views."""
    n = p.nx * p.ny
    cols = {"class": np.zeros(n, np.uint8)}
    if p.columns & COL_NU:
        cols["nu"] = np.zeros(n, np.float64)
    if p.columns & COL_DE:
        cols["de"] = np.zeros(n, np.float32)
    if p.columns & COL_NORMAL:
        cols["normal"] = np.zeros(n, np.uint16)
    cs = _Columns(_p(cols["class"]), _p(cols.get("nu")), _p(cols.get("de")),
                  _p(cols.get("normal")))
    oc = None
    if raw:
        rw = {"kind": np.zeros(n, np.uint8), "n": np.zeros(n, np.uint64),
              "zr": np.zeros(n), "zi": np.zeros(n), "dr": np.zeros(n), "di": np.zeros(n),
              "dexp": np.zeros(n, np.int64), "ops": np.zeros(n, np.uint32)}
        oc = _Outcomes(*[_p(rw[k]) for k in ("kind", "n", "zr", "zi", "dr", "di", "dexp",
                                             "ops")])
        cols["raw"] = rw
    return cols, cs, oc


class ResidentFrame:
    """A frame whose orbit, table and outcomes stay in HBM between passes
    (fdg_frame_open / run / read / close).  orbit = (re, im, bits) supplies the reference
    orbit instead of computing it (fdg_frame_open_orbit)."""

    def __init__(self, view, params, route=ROUTE_BLA, form=FORM_F64, bla_eps=EPS, ctx=None,
                 orbit=None):
        self.ctx = ctx or default_context()
        self.view, self.params = view, params
        self._v, self._p = _view(view), _params(params)
        h = C.c_void_p()
        if orbit is None:
            self.ctx._check(lib().fdg_frame_open(self.ctx._h, C.byref(self._v),
                                                 C.byref(self._p), route, form, bla_eps,
                                                 C.byref(h)))
        else:
            re = np.ascontiguousarray(orbit[0], np.float64)
            im = np.ascontiguousarray(orbit[1], np.float64)
            if re.size != im.size:
                raise FdgError(-1, "supplied reference orbit must start at Z_0 = 0 with equal "
                                   "re/im lengths")
            self.ctx._check(lib().fdg_frame_open_orbit(self.ctx._h, C.byref(self._v),
                                                       C.byref(self._p), route, form, bla_eps,
                                                       _p(re), _p(im), re.size, int(orbit[2]),
                                                       C.byref(h)))
        self._h = h

    def run(self):
        """One pass of the hot path (table build + perturbation kernel); returns stats."""
        st = _Stats()
        self.ctx._check(lib().fdg_frame_run(self._h, C.byref(st)))
        return {k: getattr(st, k) for k, _ in _Stats._fields_}

    def mask(self, mask):
        """Compute only the samples whose entry is set in the following runs (None: all):
        fd_kernel::render_with_masked's mask (crates/fd-kernel/src/grid.rs:173-181)."""
        self._mask = None if mask is None else np.ascontiguousarray(mask, np.uint8)
        if self._mask is not None and self._mask.size != self.params.nx * self.params.ny:
            raise FdgError(-1, "sample mask must cover the sample grid")
        self.ctx._check(lib().fdg_frame_mask(self._h, _p(self._mask)))

    def launch(self, groups_per_cu=0, claim_div=0, claim_cap=0, refill=0):
        """Launch tuning of the following runs (fdg_frame_launch; 0 = the default)."""
        self.ctx._check(lib().fdg_frame_launch(self._h, C.byref(_Launch(
            groups_per_cu, claim_div, claim_cap, refill))))

    def tune(self, reps=2):
        """The autotuner (fdg_frame_tune): times each candidate launch over `reps` runs and
        keeps the fastest for this frame and later frames of its shape.  Returns
        (dict of the chosen launch, its kernel ms, the default launch's kernel ms)."""
        best, bms, dms = _Launch(), C.c_double(), C.c_double()
        self.ctx._check(lib().fdg_frame_tune(self._h, reps, C.byref(best), C.byref(bms),
                                             C.byref(dms)))
        return ({k: getattr(best, k) for k, _ in _Launch._fields_}, bms.value, dms.value)

    def series(self, series):
        """Start every sample of the following runs from `series` (a Series; None clears
        it): Series::render_masked (crates/fd-kernel/src/series.rs:342-374)."""
        self._series = series
        if series is None:
            self.ctx._check(lib().fdg_frame_series(self._h, None))
        else:
            st, self._series_b = series._struct()
            self.ctx._check(lib().fdg_frame_series(self._h, C.byref(st)))

    def series_prepare(self, config=None):
        """Series::prepare on this frame (fdg_frame_series_prepare): builds the series and
        sets it when a probe of every 61st sample predicts a net win.  Returns (Series,
        admitted)."""
        cfg = dict(series_config_default(), **(config or {}))
        cap = max(int(cfg["max_order"]), 1)
        buf = np.zeros(2 * cap)
        out = _Series(b=_p(buf), b_cap=cap)
        ok = C.c_int()
        self.ctx._check(lib().fdg_frame_series_prepare(self._h, C.byref(_SeriesConfig(**cfg)),
                                                       C.byref(out), C.byref(ok)))
        sr = _series_from(out, buf)
        self._series = sr if ok.value else None
        return sr, bool(ok.value)

    def read_columns(self):
        """The columns encoded on the GPU (fdg_frame_read_columns): 15 bytes per sample
        read back; nu/de/normal within an ulp or a quantum of read()'s host encoding."""
        cols, cs, _ = _out_buffers(self.params, False)
        st = _Stats()
        kid = C.create_string_buffer(256)
        self.ctx._check(lib().fdg_frame_read_columns(self._h, C.byref(cs), C.byref(st), kid,
                                                     256))
        header = dict(view=self.view, params=self.params, kernel=kid.value.decode(),
                      columns=self.params.columns | COL_CLASS)
        return header, cols, {k: getattr(st, k) for k, _ in _Stats._fields_}

    def jet(self):
        """A jet-AA frame (fdg_frame_jet; crates/fd-cli/src/film.rs:1359-1457): pass 1 on
        each pixel's top-left sample, the host rebuild, pass 2 on the rejected samples.
        Returns (header, columns, stats, samples computed)."""
        cols, cs, _ = _out_buffers(self.params, False)
        st = _Stats()
        kid = C.create_string_buffer(256)
        done = C.c_uint64()
        self.ctx._check(lib().fdg_frame_jet(self._h, C.byref(cs), C.byref(st), kid, 256,
                                            C.byref(done)))
        header = dict(view=self.view, params=self.params, kernel=kid.value.decode(),
                      columns=self.params.columns | COL_CLASS)
        return header, cols, {k: getattr(st, k) for k, _ in _Stats._fields_}, done.value

    def read(self, raw=False):
        cols, cs, oc = _out_buffers(self.params, raw)
        st = _Stats()
        kid = C.create_string_buffer(256)
        self.ctx._check(lib().fdg_frame_read(self._h, C.byref(cs),
                                             C.byref(oc) if oc is not None else None,
                                             C.byref(st), kid, 256))
        header = dict(view=self.view, params=self.params, kernel=kid.value.decode(),
                      columns=self.params.columns | COL_CLASS)
        return header, cols, {k: getattr(st, k) for k, _ in _Stats._fields_}

    def close(self):
        if getattr(self, "_h", None):
            lib().fdg_frame_close(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


TABLES_PER_FRAME, TABLES_SHARED = 0, 1
LAUNCH_BATCH, LAUNCH_PER_FRAME = 0, 1


class Film:
    """Frames at one centre (a zoom's keyframes) rendered against one reference orbit at
    the deepest frame's precision, as fd_kernel::render_with / render_bla with a supplied
    orbit (crates/fd-kernel/src/grid.rs:95-109,190-210) renders each of them
    (fdg_film_open / run / read / close)."""

    def __init__(self, views, params, route=ROUTE_BLA, form=FORM_F64, bla_eps=EPS,
                 tables=TABLES_PER_FRAME, ctx=None):
        self.ctx = ctx or default_context()
        self.views, self.params = list(views), params
        arr = (_View * len(self.views))(*[_view(v) for v in self.views])
        self._keep = arr
        self._p = _params(params)
        h = C.c_void_p()
        self.ctx._check(lib().fdg_film_open(self.ctx._h, arr, len(self.views),
                                            C.byref(self._p), route, form, bla_eps, tables,
                                            C.byref(h)))
        self._h = h

    def run(self, launch=LAUNCH_BATCH, grid=0):
        """Render every frame (one graph of persistent batch launches, or one launch per
        frame); returns the film's aggregate stats."""
        st = _Stats()
        self.ctx._check(lib().fdg_film_run(self._h, launch, grid, C.byref(st)))
        return {k: getattr(st, k) for k, _ in _Stats._fields_}

    def tune(self, grids=(1024, 2048, 4096), reps=2, launch=LAUNCH_BATCH):
        """The film's autotuner: times `reps` runs under each persistent grid size (in
        workgroups; 0 is the default, 8 per CU) by the kernel time and returns (the fastest
        grid, its kernel ms, the default grid's kernel ms).  A candidate replaces the
        default only when 1% faster.  Outputs do not depend on the grid."""
        def best_of(g):
            return min(self.run(launch, g)["kernel_seconds"] for _ in range(reps)) * 1e3
        default = best_of(0)
        win, win_ms = 0, default
        for g in grids:
            ms = best_of(g)
            if ms < (0.99 * win_ms if win == 0 else win_ms):
                win, win_ms = g, ms
        return win, win_ms, default

    def mask(self, masks):
        """Compute only the set samples in the following runs (fdg_film_mask): masks is
        nframes x (nx ny) bytes, frame-major; None clears."""
        self._mask = None if masks is None else np.ascontiguousarray(masks, np.uint8).ravel()
        n = len(self.views) * self.params.nx * self.params.ny
        if self._mask is not None and self._mask.size != n:
            raise FdgError(-1, "film mask must cover every frame's samples")
        self.ctx._check(lib().fdg_film_mask(self._h, _p(self._mask)))

    def jet(self):
        """Jet-AA for every frame (fdg_film_jet): returns ([columns per frame], computed)."""
        outs = [_out_buffers(self.params, False) for _ in self.views]
        arr = (_Columns * len(outs))(*[cs for _, cs, _ in outs])
        done = C.c_uint64()
        self.ctx._check(lib().fdg_film_jet(self._h, arr, C.byref(done)))
        return [cols for cols, _, _ in outs], done.value

    def read(self, k, raw=False):
        cols, cs, oc = _out_buffers(self.params, raw)
        st = _Stats()
        kid = C.create_string_buffer(256)
        self.ctx._check(lib().fdg_film_read(self._h, k, C.byref(cs),
                                            C.byref(oc) if oc is not None else None,
                                            C.byref(st), kid, 256))
        header = dict(view=self.views[k], params=self.params, kernel=kid.value.decode(),
                      columns=self.params.columns | COL_CLASS)
        return header, cols, {k: getattr(st, k) for k, _ in _Stats._fields_}

    def close(self):
        if getattr(self, "_h", None):
            lib().fdg_film_close(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def render_masked(view, params, mask, route=ROUTE_PLAIN, form=FORM_F64, bla_eps=EPS, ctx=None,
                  raw=False):
    """fd_kernel::render_with_masked / render_bla_masked (crates/fd-kernel/src/grid.rs:
    173-209): only the samples whose mask entry is set are computed; the others are left
    unset (zero columns; raw kind 255)."""
    ctx = ctx or default_context()
    m = np.ascontiguousarray(mask, np.uint8)
    if m.size != params.nx * params.ny:
        raise FdgError(-1, "sample mask must cover the sample grid")
    cols, cs, oc = _out_buffers(params, raw)
    st = _Stats()
    kid = C.create_string_buffer(256)
    ctx._check(lib().fdg_render_masked(ctx._h, C.byref(_view(view)), C.byref(_params(params)),
                                       route, form, bla_eps, _p(m), C.byref(cs),
                                       C.byref(oc) if oc is not None else None, C.byref(st),
                                       kid, 256))
    header = dict(view=view, params=params, kernel=kid.value.decode(),
                  columns=params.columns | COL_CLASS)
    return header, cols, {k: getattr(st, k) for k, _ in _Stats._fields_}


class Series:
    """An admitted parameter-series start (fd_kernel::Series, crates/fd-kernel/src/
    series.rs:57-82): every sample with |dc| <= rho starts at iterate n with delta
    sum b[k] (dc / rho)^k.  Built on the host by series_build."""

    def __init__(self, n, rho, points, b, state_bound=0.0, derivative_bound=0.0,
                 shift_px=0.0, built_steps=0, coefficient_pairs=0, growth_steps=0,
                 stop="reference"):
        self.n, self.rho, self.points = int(n), float(rho), int(points)
        self.b = np.asarray(b, complex)
        self.state_bound, self.derivative_bound = state_bound, derivative_bound
        self.shift_px, self.built_steps = shift_px, built_steps
        self.coefficient_pairs, self.growth_steps, self.stop = (coefficient_pairs,
                                                                growth_steps, stop)

    def _struct(self):
        b = np.ascontiguousarray(np.stack([self.b.real, self.b.imag], axis=1).reshape(-1))
        st = _Series(n=self.n, rho=self.rho, points=self.points, b=_p(b), b_cap=len(self.b),
                     b_len=len(self.b), state_bound=self.state_bound,
                     derivative_bound=self.derivative_bound, shift_px=self.shift_px)
        return st, b

    def replace(self, **kw):
        d = dict(self.__dict__)
        d.update(kw)
        return Series(**d)


def series_config_default():
    """SeriesConfig::default (series.rs:43-55) as a dict."""
    c = _SeriesConfig()
    lib().fdg_series_config_default(C.byref(c))
    return {k: getattr(c, k) for k, _ in _SeriesConfig._fields_}


def series_build(re, im, rho, pixel, max_iter, config=None):
    """Series::build (series.rs:88-234) over an orbit, on the host (fdg_series_build)."""
    re = np.ascontiguousarray(re, np.float64)
    im = np.ascontiguousarray(im, np.float64)
    cfg = dict(series_config_default(), **(config or {}))
    c = _SeriesConfig(**cfg)
    cap = max(int(cfg["max_order"]), 1)
    buf = np.zeros(2 * cap)
    out = _Series(b=_p(buf), b_cap=cap)
    rc = lib().fdg_series_build(_p(re), _p(im), len(re), rho, pixel, max_iter, C.byref(c),
                                C.byref(out))
    if rc != 0:
        raise FdgError(rc, "series build: bad arguments")
    return _series_from(out, buf)


def _series_from(out, buf):
    b = buf[:2 * out.b_len].reshape(-1, 2)
    return Series(out.n, out.rho, out.points, b[:, 0] + 1j * b[:, 1], out.state_bound,
                  out.derivative_bound, out.shift_px, out.built_steps, out.coefficient_pairs,
                  out.growth_steps, SERIES_STOPS[out.stop])


def bla_dc_max(view, params):
    """fd_kernel::bla_dc_max (grid.rs:432-440): the frame's |dc| bound (not on the scaled
    tier)."""
    d = C.c_double()
    rc = lib().fdg_bla_dc_max(C.byref(_view(view)), C.byref(_params(params)), C.byref(d))
    if rc != 0:
        raise FdgError(rc, "bla_dc_max: invalid view, or the scaled tier")
    return d.value


def render_series(view, params, series, route=ROUTE_PLAIN, bla_eps=EPS, mask=None, ctx=None,
                  raw=False):
    """fd_kernel::render_series / Series::render_masked (grid.rs:218-226, series.rs:342-374)
    on the GPU: every sample starts from the series state; unresolved ones retry from
    zero."""
    ctx = ctx or default_context()
    m = None if mask is None else np.ascontiguousarray(mask, np.uint8)
    if m is not None and m.size != params.nx * params.ny:
        raise FdgError(-1, "sample mask must cover the sample grid")
    cols, cs, oc = _out_buffers(params, raw)
    st = _Stats()
    kid = C.create_string_buffer(256)
    sst, keep = series._struct()
    ctx._check(lib().fdg_render_series(ctx._h, C.byref(_view(view)), C.byref(_params(params)),
                                       route, bla_eps, C.byref(sst), _p(m), C.byref(cs),
                                       C.byref(oc) if oc is not None else None, C.byref(st),
                                       kid, 256))
    header = dict(view=view, params=params, kernel=kid.value.decode(),
                  columns=params.columns | COL_CLASS)
    return header, cols, {k: getattr(st, k) for k, _ in _Stats._fields_}


def jetaa_first(w, h):
    """jetaa.rs `first` (fdg_jetaa_first): the pass-1 mask of a w x h pixel ss 2 grid."""
    m = np.zeros(4 * w * h, np.uint8)
    rc = lib().fdg_jetaa_first(w, h, _p(m))
    if rc != 0:
        raise FdgError(rc, "jetaa_first")
    return m


def _cols_struct(cols):
    return _Columns(_p(cols["class"]), _p(cols.get("nu")), _p(cols.get("de")),
                    _p(cols.get("normal")))


def jetaa_rebuild(cols, w, h):
    """jetaa.rs `rebuild` (fdg_jetaa_rebuild) in place on ss 2 columns (class, nu, de,
    normal); returns the pass-2 mask."""
    for k, t in (("class", np.uint8), ("nu", np.float64), ("de", np.float32),
                 ("normal", np.uint16)):
        if cols[k].dtype != t or not cols[k].flags.c_contiguous or cols[k].size != 4 * w * h:
            raise FdgError(-1, f"column {k} must be {t.__name__}[{4 * w * h}]")
    todo = np.zeros(4 * w * h, np.uint8)
    rc = lib().fdg_jetaa_rebuild(C.byref(_cols_struct(cols)), w, h, _p(todo))
    if rc != 0:
        raise FdgError(rc, "jetaa_rebuild")
    return todo


def jetaa_merge(to, frm, mask):
    """jetaa.rs `merge` (fdg_jetaa_merge): copy the masked samples of frm into to."""
    m = np.ascontiguousarray(mask, np.uint8)
    rc = lib().fdg_jetaa_merge(C.byref(_cols_struct(to)), C.byref(_cols_struct(frm)), _p(m),
                               m.size)
    if rc != 0:
        raise FdgError(rc, "jetaa_merge")


def render_bla(view, params, form=FORM_F64, bla_eps=EPS, ctx=None, raw=False):
    """perturb_gpu::render_bla (frame.rs:78-84)."""
    return render(view, params, ROUTE_BLA, form, bla_eps, ctx, raw)


def nucleus(re, im, period, bits, max_iters=60):
    """Newton on z_P(c) = 0 from (re, im) at `bits` fraction bits (fdg_nucleus, host-only;
    the reference's zone_build.rs:215-233 iteration): returns (re, im, Newton steps, log2
    of the last step), the nucleus as decimal strings; FdgError when it escapes or does not
    converge."""
    cap = int(bits * 0.302) + 64
    ro, io = C.create_string_buffer(cap), C.create_string_buffer(cap)
    its, step = C.c_uint32(), C.c_double()
    rc = lib().fdg_nucleus(None, re.encode(), im.encode(), int(period), int(bits),
                           int(max_iters), ro, io, cap, C.byref(its), C.byref(step))
    if rc < 0:
        raise FdgError(rc, lib().fdg_last_error(None).decode())
    return ro.value.decode(), io.value.decode(), its.value, step.value


def reference(view, params, ctx=None):
    """fd_kernel::reference: ``(re, im, bits)`` of the orbit the renderer uses."""
    ctx = ctx or default_context()
    ln, bits = C.c_uint64(), C.c_uint32()
    v, p = _view(view), _params(params)
    ctx._check(lib().fdg_reference(ctx._h, C.byref(v), C.byref(p), None, None, 0, C.byref(ln),
                                   C.byref(bits)))
    re = np.zeros(ln.value)
    im = np.zeros(ln.value)
    ctx._check(lib().fdg_reference(ctx._h, C.byref(v), C.byref(p), _p(re), _p(im), ln.value,
                                   C.byref(ln), C.byref(bits)))
    return re, im, bits.value


def build_bla(re, im, eps, dc_max, dc_exp=0, ctx=None):
    """Bla::build_scaled on the GPU: a list of (count, 7) arrays for levels 1, 2, ..."""
    ctx = ctx or default_context()
    re = np.ascontiguousarray(re, np.float64)
    im = np.ascontiguousarray(im, np.float64)
    steps = max(len(re) - 2, 0)
    total = sum(steps >> j for j in range(1, 64))
    out = np.zeros((max(total, 1), 7))
    nl = C.c_uint32()
    ctx._check(lib().fdg_build_bla(ctx._h, _p(re), _p(im), len(re), eps, dc_max, dc_exp,
                                   _p(out), total, C.byref(nl)))
    levels, off = [], 0
    for j in range(1, nl.value + 1):
        c = steps >> j
        levels.append(out[off:off + c].copy())
        off += c
    return levels


def jacobian(header, samples):
    """Screen-space 2x2 Jacobian of z at escape (fdg_jacobian): (j[n, 2, 2], exponent[n]),
    J = j * 2^exponent, rows d(Re z, Im z), columns d/d(x, y) in samples."""
    raw = samples["raw"]
    n = len(raw["kind"])
    oc = _Outcomes(*[_p(raw[k]) for k in ("kind", "n", "zr", "zi", "dr", "di", "dexp", "ops")])
    j = np.zeros((n, 2, 2))
    e = np.zeros(n, np.int64)
    rc = lib().fdg_jacobian(C.byref(_view(header["view"])), C.byref(_params(header["params"])),
                            C.byref(oc), _p(j), _p(e))
    if rc != 0:
        raise FdgError(rc, "jacobian")
    return j, e


def write(header, samples, minor=1):
    """fd_samples::write: the canonical .fds bytes (``minor`` rewrites bytes 10..12)."""
    v, p = _view(header["view"]), _params(header["params"])
    cs = _Columns(_p(samples["class"]), _p(samples.get("nu")), _p(samples.get("de")),
                  _p(samples.get("normal")))
    kid = header["kernel"].encode()
    size = lib().fdg_write_fds(C.byref(v), C.byref(p), kid, C.byref(cs), None, 0)
    if size < 0:
        raise FdgError(size, "write failed")
    buf = np.zeros(size, np.uint8)
    lib().fdg_write_fds(C.byref(v), C.byref(p), kid, C.byref(cs), _p(buf), size)
    b = bytearray(buf.tobytes())
    b[10:12] = int(minor).to_bytes(2, "little")
    return bytes(b)


# ---------------------------------------------------------------- atlas chunks


def chunk_id(data):
    """SHA-256 of a chunk's bytes (fd_atlas::ChunkId::of)."""
    b = np.frombuffer(bytes(data), np.uint8)
    out = np.zeros(32, np.uint8)
    rc = lib().fdg_chunk_id(_p(b) if b.size else None, b.size, _p(out))
    if rc != 0:
        raise FdgError(rc, "chunk_id")
    return out.tobytes()


def _sized(call):
    n = call(None, 0)
    if n < 0:
        raise FdgError(int(n), "refused by the chunk format")
    buf = np.zeros(n, np.uint8)
    call(_p(buf), n)
    return buf.tobytes()


def orbit_slabs(re, im, bits, size=4096):
    """OrbitSlab::split + to_chunk: the orbit's slab chunks (bytes), in order."""
    re = np.ascontiguousarray(re, np.float64)
    im = np.ascontiguousarray(im, np.float64)
    out = []
    for start in range(0, re.size, size):
        ln = min(size, re.size - start)
        out.append(_sized(lambda b, c: lib().fdg_chunk_orbit_slab(
            _p(re[start:]), _p(im[start:]), start, ln, size, bits, b, c)))
    return out


def orbit_slab_read(data):
    """OrbitSlab::from_chunk: (start, size, bits, re, im)."""
    b = np.frombuffer(bytes(data), np.uint8)
    st, ln, sz, bits = C.c_uint64(), C.c_uint32(), C.c_uint32(), C.c_uint32()
    rc = lib().fdg_chunk_orbit_slab_read(_p(b), b.size, C.byref(st), C.byref(ln), C.byref(sz),
                                         C.byref(bits), None, None, 0)
    if rc != 0:
        raise FdgError(rc, "not a canonical orbit slab chunk")
    re, im = np.zeros(ln.value), np.zeros(ln.value)
    lib().fdg_chunk_orbit_slab_read(_p(b), b.size, None, None, None, None, _p(re), _p(im),
                                    ln.value)
    return st.value, sz.value, bits.value, re, im


def orbit_manifest(center_re, center_im, bits, slab_ids):
    """OrbitManifest::to_chunk."""
    ids = np.frombuffer(b"".join(slab_ids), np.uint8)
    return _sized(lambda b, c: lib().fdg_chunk_orbit_manifest(
        center_re.encode(), center_im.encode(), bits, _p(ids), len(slab_ids), b, c))


def orbit_manifest_read(data):
    """OrbitManifest::from_chunk: (center_re, center_im, bits, [slab ids])."""
    b = np.frombuffer(bytes(data), np.uint8)
    bits, n = C.c_uint32(), C.c_uint64()
    cap = b.size + 1
    sre, sim = C.create_string_buffer(cap), C.create_string_buffer(cap)
    rc = lib().fdg_chunk_orbit_manifest_read(_p(b), b.size, sre, sim, cap, C.byref(bits),
                                             None, 0, C.byref(n))
    if rc != 0:
        raise FdgError(rc, "not a canonical orbit manifest chunk")
    ids = np.zeros(32 * n.value, np.uint8)
    lib().fdg_chunk_orbit_manifest_read(_p(b), b.size, None, None, 0, None, _p(ids), n.value,
                                        None)
    raw = ids.tobytes()
    return (sre.value.decode(), sim.value.decode(), bits.value,
            [raw[32 * k:32 * k + 32] for k in range(n.value)])


def bla_chunk(orbit_id, eps, dc_max, points, levels):
    """BlaTable::to_chunk from build_bla()'s levels (list of (count, 7) arrays)."""
    blocks = np.ascontiguousarray(np.concatenate(levels, axis=0) if levels else np.zeros((0, 7)))
    oid = np.frombuffer(bytes(orbit_id), np.uint8)
    return _sized(lambda b, c: lib().fdg_chunk_bla(_p(oid), eps, dc_max, points, len(levels),
                                                   _p(blocks), b, c))


def bla_chunk_read(data):
    """BlaTable::from_chunk: (orbit id, eps, dc_max, points, levels)."""
    b = np.frombuffer(bytes(data), np.uint8)
    oid = np.zeros(32, np.uint8)
    eps, dc, pts, nl = C.c_double(), C.c_double(), C.c_uint64(), C.c_uint32()
    rc = lib().fdg_chunk_bla_read(_p(b), b.size, _p(oid), C.byref(eps), C.byref(dc),
                                  C.byref(pts), C.byref(nl), None, 0)
    if rc != 0:
        raise FdgError(rc, "not a canonical BLA table chunk")
    steps = max(pts.value - 2, 0)
    total = sum(steps >> j for j in range(1, nl.value + 1))
    blocks = np.zeros((max(total, 1), 7))
    lib().fdg_chunk_bla_read(_p(b), b.size, None, None, None, None, None, _p(blocks), total)
    levels, off = [], 0
    for j in range(1, nl.value + 1):
        c = steps >> j
        levels.append(blocks[off:off + c].copy())
        off += c
    return oid.tobytes(), eps.value, dc.value, pts.value, levels


def error_radius(re, im, bits, c_re, c_im):
    """Reference::error_radius (fdg_error_radius); bits None for an f64 orbit."""
    re = np.ascontiguousarray(re, np.float64)
    im = np.ascontiguousarray(im, np.float64)
    q = np.zeros(re.size)
    rc = lib().fdg_error_radius(_p(re), _p(im), re.size, -1 if bits is None else int(bits),
                                c_re, c_im, _p(q))
    if rc != 0:
        raise FdgError(rc, "error_radius")
    return q


def guarded_bla(re, im, q, eps, dc_max):
    """GuardedBla::build (fdg_guarded_bla): a list of (count, 9) arrays, levels 1.."""
    re = np.ascontiguousarray(re, np.float64)
    im = np.ascontiguousarray(im, np.float64)
    q = np.ascontiguousarray(q, np.float64)
    steps = max(re.size - 2, 0)
    cap = max(steps, 1)
    out = np.zeros((cap, 9))
    nl = C.c_uint32()
    rc = lib().fdg_guarded_bla(_p(re), _p(im), _p(q), re.size, eps, dc_max, _p(out), cap,
                               C.byref(nl))
    if rc != 0:
        raise FdgError(rc, lib().fdg_last_error(None).decode())
    levels, off = [], 0
    for j in range(1, nl.value + 1):
        c = steps >> j
        levels.append(out[off:off + c].copy())
        off += c
    return levels


def external_escape_ok(n, z, d, dexp, ez, ed, px_m, px_e):
    """external_escape_ok (fdg_external_escape_ok) for an escaped state."""
    st = _EscapedState(n, z[0], z[1], d[0], d[1], dexp)
    rc = lib().fdg_external_escape_ok(C.byref(st), ez, ed, px_m, px_e)
    if rc < 0:
        raise FdgError(rc, "external_escape_ok")
    return rc == 1


def overlay_escaped(view, params, cols, states):
    """overlay_escaped (fdg_overlay_escaped): states is a dict of per-sample arrays n, zr,
    zi, dr, di, dexp and a boolean `present`; writes the present ones into cols."""
    n = params.nx * params.ny
    arr = np.zeros(n, dtype=[("n", "<u8"), ("zr", "<f8"), ("zi", "<f8"), ("dr", "<f8"),
                             ("di", "<f8"), ("dexp", "<i8")])
    for k in ("n", "zr", "zi", "dr", "di", "dexp"):
        arr[k] = states[k]
    present = np.ascontiguousarray(states["present"], np.uint8)
    v, p = _view(view), _params(params)
    rc = lib().fdg_overlay_escaped(C.byref(v), C.byref(p), C.byref(_cols_struct(cols)),
                                   arr.ctypes.data, _p(present))
    if rc != 0:
        raise FdgError(rc, lib().fdg_last_error(None).decode())
