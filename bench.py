"""Benchmark: effective pixel-iterations/s of the deep-zoom perturbation/BLA path on one GPU.

Workload (default, config 3): the 1920x1080 frame of width 1e-1000 at the period-16116
ladder nucleus of the reference's rung-7676 frame (its minibrot is ~1e-1000 wide, so the
frame holds its interior and its exterior), max_iter 1,000,000, escape radius 1e10, BLA on
(eps 2^-50), columns nu/de/normal (so dz/dc is carried), FP64 mantissa x 2^e deltas (the
scaled tier).  A step is one pass of the hot path over the frame with its reference orbit
resident in HBM: the GPU BLA table build plus the persistent perturbation kernel over all
2,073,600 samples.  value = plain-equivalent iterations (BLA-skipped steps included, the
reference's `iterations.equivalent`) per second of wall time over the timed steps.  The
FE32 and mixed forms run the same frame (with their agreement), and the 1e-100 frame, the
escape-dominated frame, the 64-keyframe film and the 1e-3000 jet-AA frame run as
companions.

Prints ONE JSON line.  --gpus above 1 exits: this path runs on one GPU here.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_TFLOPS = 78.6      # MI355X vector FP64 (and FP32 without packed instructions)
PEAK_FP32_TFLOPS = 157.3     # MI355X vector FP32 with packed (v_pk_*) instructions
F_PLAIN, F_BLOCK = 18, 22    # algorithmic flops per plain step / applied block
F_PLAIN_D, F_BLOCK_D = 27, 30  # with dz/dc carried


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_cpu():
    """The host's CPU: model name, online cores and the cores this process may run on."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    online = os.cpu_count() or 1
    allowed = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else online
    return model, online, allowed


def cpu_threads():
    """Worker threads for the CPU baseline: every core this process may use, capped at 16,
    the host CPU share that comes with one GPU on the box."""
    return max(1, min(16, host_cpu()[2]))


def cpu_baseline(view, params, route_bla, target_seconds, threads, orbit=None, runs=3):
    """The oracle (C port of the reference CPU BLA renderer, its dynamic row queue on
    `threads` workers) on a row sample of the frame sized for ~target_seconds of render
    per run: `runs` runs of the same sample, the first labelled cold, the median reported
    (render-only; end to end adds the host reference orbit and the table build, each
    timed apart).  orbit: a supplied (re, im, bits) reference, as the film's frames use."""
    import numpy as np
    import oracle
    nx, ny = params.nx, params.ny

    def run(stride, offset=0):
        mask = np.zeros(nx * ny, np.uint8)
        rows = list(range(offset, ny, stride))
        for j in rows:
            mask[j * nx:(j + 1) * nx] = 1
        f = oracle.render(view.center_re, view.center_im, view.width, view.rotation, nx, ny,
                          params.max_iter, params.escape_radius, columns=params.columns,
                          bla=route_bla, threads=threads, mask=mask, outcomes=False,
                          orbit=orbit)
        return f["stats"], len(rows)

    st, rows = run(ny // 4, ny // 8)   # sizing probe: 4 rows
    per_row = st["render_seconds"] / rows
    want_rows = max(4, min(ny, int(target_seconds / max(per_row, 1e-6))))
    stride = max(1, ny // want_rows)
    each = []
    for k in range(runs):
        st, rows = run(stride, stride // 2)
        e2e = st["render_seconds"] + st["reference_seconds"] + st["bla_seconds"]
        each.append(dict(run=k + 1, cold=k == 0, render_seconds=st["render_seconds"],
                         reference_seconds=st["reference_seconds"],
                         bla_seconds=st["bla_seconds"],
                         value=st["equivalent"] / st["render_seconds"],
                         end_to_end_value=st["equivalent"] / e2e))
    med = float(np.median([r["value"] for r in each]))
    model, online, allowed = host_cpu()
    return dict(value=med, unit="pixel-iter/s", cores=threads, kind="port",
                end_to_end_value=float(np.median([r["end_to_end_value"] for r in each])),
                runs=each, statistic=f"median of {runs} runs (run 1 cold)",
                cpu_model=model, nproc=online, affinity_cores=allowed,
                sample=f"{rows} of {ny} rows (every {stride}th) x {nx} samples of the same "
                       f"frame, {'BLA eps 2^-50' if route_bla else 'plain'}, per run: oracle "
                       f"render {np.median([r['render_seconds'] for r in each]):.2f}s "
                       f"(median; reference orbit and table build timed apart), {threads} "
                       f"threads of {allowed} allowed / {online} online cores")


LAUNCH = "auto"   # --launch: auto (autotune each frame), default, or G,D,C


def measure(fa, view, params, route, steps, warmup, label, form=None, keep=None):
    """Open the frame (orbit resident in HBM), autotune its launch (untimed), warm up, time
    `steps` passes.  keep: a list that receives the frame's columns and raw outcomes."""
    form = fa.FORM_F64 if form is None else form
    t = time.perf_counter()
    frame = fa.ResidentFrame(view, params, route, form, fa.EPS)
    log(f"{label}: frame open (host reference orbit + upload) {time.perf_counter() - t:.2f}s")
    launch = None
    if LAUNCH not in ("auto", "default"):
        g, d, c, *rf = (int(x) for x in LAUNCH.split(","))
        rf = rf[0] if rf else 0
        frame.launch(g, d, c, rf)
        launch = dict(chosen=dict(groups_per_cu=g, claim_div=d, claim_cap=c, refill=rf),
                      given=True)
    if LAUNCH == "auto":
        t = time.perf_counter()
        chosen, tuned_ms, default_ms = frame.tune(reps=2)
        launch = dict(chosen=chosen, tuned_kernel_ms=tuned_ms, default_kernel_ms=default_ms,
                      tuning_seconds=time.perf_counter() - t,
                      candidates="groups/CU {4,8,16} x claim divisor {2,4,8} x claim cap "
                                 "{256,1024,4096} with whole-wave refills, plus the default "
                                 "(8, 4, 1024, 64), then refills at 1 or 16 idle lanes under "
                                 "the best; min of 2 runs each")
        log(f"{label}: tuned launch {chosen} {tuned_ms:.3f} ms (default {default_ms:.3f} ms)")
    for i in range(warmup):
        st = frame.run()
        log(f"{label} warmup {i}: kernel {st['kernel_seconds'] * 1e3:.2f} ms, table "
            f"{st['bla_seconds'] * 1e3:.2f} ms")
    stats = []
    t0 = time.perf_counter()
    for i in range(steps):
        stats.append(frame.run())
    t1 = time.perf_counter()
    # (logged after the timed region, which then holds the passes alone)
    for i, st in enumerate(stats):
        log(f"{label} step {i}: kernel {st['kernel_seconds'] * 1e3:.2f} ms")
    h, cols, _ = frame.read(raw=True)
    frame.close()
    if keep is not None:
        keep.append(cols)
    import numpy as np
    ops = cols["raw"]["ops"].astype(np.float64)
    ops_pct = {f"p{q}": float(np.percentile(ops, q)) for q in (50, 90, 99, 99.9, 100)}
    st = stats[-1]
    step_s = (t1 - t0) / steps
    kernel_s = sum(s["kernel_seconds"] for s in stats) / len(stats)
    deriv = bool(params.columns & (fa.COL_DE | fa.COL_NORMAL))
    fp, fb = (F_PLAIN_D, F_BLOCK_D) if deriv else (F_PLAIN, F_BLOCK)
    flops = st["plain_steps"] * fp + st["blocks"] * fb
    achieved = flops / kernel_s / 1e12
    n = params.nx * params.ny
    cls = cols["class"] & 3
    classes = {k: float((cls == v).sum()) / n for k, v in (("escaped", 0), ("interior", 1),
                                                           ("unresolved", 2))}
    return dict(value=st["equivalent"] / step_s, ms_per_step=step_s * 1e3,
                kernel_ms=kernel_s * 1e3, table_ms=st["bla_seconds"] * 1e3,
                kernel_id=h["kernel"], reference_len=st["reference_len"],
                reference_bits=st["reference_bits"], bla_levels=st["bla_levels"],
                nucleus_period=st["nucleus_period"], plain_steps=st["plain_steps"],
                blocks=st["blocks"], skipped=st["skipped"],
                equivalent_iterations=st["equivalent"], classes=classes,
                sample_ops_percentiles=ops_pct, max_sample_ops=st["max_sample_ops"],
                wave_iterations=st["wave_iterations"],
                lane_utilisation=st["lane_iterations"] / max(1, 64 * st["wave_iterations"]),
                reference_seconds=st["reference_seconds"],
                uncertified_samples=st.get("uncertified_samples", 0), launch=launch,
                roofline={"bound": "valu-fp64", "achieved": achieved, "peak": PEAK_FP64_TFLOPS,
                          "unit": "TFLOP/s", "frac": achieved / PEAK_FP64_TFLOPS,
                          "traffic": None, "flops_per_launch": flops,
                          "flops_model": f"{fp}/plain step + {fb}/applied block"})


def measured_traffic(config):
    """HBM bytes per launch of the dominant kernel for this workload, from the rocprofv3
    FETCH_SIZE / WRITE_SIZE runs committed under profiles/ (each counter scaled by what it
    reads for a plain copy of known size), or None when no profile covers the workload."""
    path = os.path.join(ROOT, "profiles", f"traffic_config{config}.json")
    if not os.path.exists(path):
        return None
    t = json.load(open(path))
    kib = t["fetch_kib_per_launch"] * t["fetch_scale"] + t["write_kib_per_launch"] * t["write_scale"]
    return kib * 1024.0


def end_to_end(fa, view, params, route, reps=2):
    """The drop-in call as the route's caller makes it (fdg_render: host reference orbit,
    upload, table, kernel, readback of the outcomes, host column encoding), wall time per
    frame over `reps` calls after one warm call, with its parts from the call's stats."""
    fa.render(view, params, route)
    walls, parts = [], None
    for _ in range(reps):
        t = time.perf_counter()
        _, _, st = fa.render(view, params, route)
        walls.append(time.perf_counter() - t)
        parts = {k: st[k] * 1e3 for k in ("reference_seconds", "bla_seconds", "kernel_seconds",
                                          "transfer_seconds", "encode_seconds")}
    w = sum(walls) / len(walls)
    return dict(ms_per_frame=w * 1e3, value=st["equivalent"] / w,
                parts_ms={k.replace("_seconds", ""): v for k, v in parts.items()},
                reference_period=st["reference_period"],
                reference_repeat_at=st["reference_repeat_at"])


def end_to_end_device_encode(fa, view, params, route, reps=2):
    """end_to_end with Store::put on the GPU (fdg_frame_read_columns): the frame opened,
    run and read as 15-byte columns."""
    def once():
        t = time.perf_counter()
        with fa.ResidentFrame(view, params, route) as f:
            f.run()
            _, _, st = f.read_columns()
        return time.perf_counter() - t, st
    once()
    walls, st = [], None
    for _ in range(reps):
        w, st = once()
        walls.append(w)
    w = sum(walls) / len(walls)
    return dict(ms_per_frame=w * 1e3, value=st["equivalent"] / w,
                parts_ms={k.replace("_seconds", ""): st[k] * 1e3
                          for k in ("reference_seconds", "bla_seconds", "kernel_seconds",
                                    "transfer_seconds", "encode_seconds")})


def agreement(ref, other):
    """Per-sample agreement of another delta form with the FP64 form's outcomes."""
    import numpy as np
    ka, kb = ref["raw"]["kind"], other["raw"]["kind"]
    esc = (ka == 0) & (kb == 0)
    out = {"samples": int(ka.size), "class_agreement": float((ka == kb).mean()),
           "exterior_agreement": float(((ka == 0) == (kb == 0)).mean()),
           "escaped_in_both": int(esc.sum())}
    if esc.any():
        dn = other["raw"]["n"][esc].astype(np.int64) - ref["raw"]["n"][esc].astype(np.int64)
        dnu = np.abs(other["nu"][esc] - ref["nu"][esc])
        out.update(escape_count_equal=float((dn == 0).mean()),
                   escape_count_within_1=float((np.abs(dn) <= 1).mean()),
                   smooth_count_err_median=float(np.median(dnu)),
                   smooth_count_err_p99=float(np.percentile(dnu, 99)),
                   smooth_count_within_1e_3=float((dnu <= 1e-3).mean()))
    return out


_DIRECT = None


def vs_direct(k, cols):
    """A frame's outcomes at the 256 seeded pixels of tests/golden/direct_hp.json (direct
    high-precision iteration of z^2 + c) against that fixture, each miss as the displacement
    of c in pixels that explains it (the reference's tools/oracle.py rule): class misses by
    the escaped side's distance estimate, smooth-count misses by |d nu| de ln2 / 2.  None
    when the fixture has no entry for frame k."""
    global _DIRECT
    import numpy as np
    if _DIRECT is None:
        path = os.path.join(ROOT, "tests", "golden", "direct_hp.json")
        _DIRECT = json.load(open(path)) if os.path.exists(path) else {}
    d = _DIRECT.get(str(k))
    if d is None:
        return None
    idx = np.array(d["j"]) * d["nx"] + np.array(d["i"])
    tn = np.array(d["n"], np.int64)
    ok = np.array(d["converged"])
    tz2 = np.array(d["zr"]) ** 2 + np.array(d["zi"]) ** 2
    tde = np.array([2.0 ** x if x is not None else np.nan for x in d["log2_de_px"]])
    kind, n = cols["raw"]["kind"][idx], cols["raw"]["n"][idx].astype(np.int64)
    nu, de = cols["nu"][idx], cols["de"][idx].astype(np.float64)
    esc_t, esc_r = tn >= 0, kind == 0
    both, miss = ok & esc_t & esc_r, ok & (esc_t != esc_r)
    with np.errstate(divide="ignore", invalid="ignore"):
        tnu = tn + 1.0 - np.log2(0.5 * np.log2(tz2))
    dnu = np.abs(nu - tnu)[both]
    px = dnu * tde[both] * math.log(2.0) / 2.0
    return dict(pixels=int(ok.sum()), escaped=int((ok & esc_t).sum()),
                class_mismatch=int(miss.sum()),
                class_px_err=float(np.where(esc_t, tde, de)[miss].max()) if miss.any() else 0.0,
                nu_px_err=float(px.max()) if both.any() else 0.0,
                escape_count_equal=float((n[both] == tn[both]).mean()) if both.any() else 1.0,
                smooth_count_within_1e_3=float((dnu <= 1e-3).mean()) if both.any() else 1.0,
                bar="class and smooth-count misses within 1e-3 pixel of displacement")


def measure_forms(fa, view, params, route, steps, warmup, label, ref_metrics, ref_cols,
                  direct_k=None):
    """The FP32 floatexp form and the mixed form (FP32 pass + FP64 recompute of its escaped
    samples) on the same frame: speed against the FP64 form, per-sample agreement with
    it (the FP64 form equals the oracle bit for bit), and agreement with direct
    high-precision iteration at the fixture's pixels of frame direct_k."""
    out = {}
    for form, name in ((fa.FORM_FE32, "fe32"), (fa.FORM_MIXED, "mixed")):
        keep = []
        m = measure(fa, view, params, route, steps, warmup, f"{label}-{name}", form, keep)
        out[name] = {"kernel_ms": m["kernel_ms"], "ms_per_step": m["ms_per_step"],
                     "value": m["value"], "kernel_id": m["kernel_id"],
                     "speedup_over_fp64_kernel": ref_metrics["kernel_ms"] / m["kernel_ms"],
                     "uncertified_samples": m["uncertified_samples"],
                     "agreement_with_fp64": agreement(ref_cols, keep[0])}
        if direct_k is not None:
            out[name]["vs_direct_high_precision"] = vs_direct(direct_k, keep[0])
    if direct_k is not None:
        out["fp64_vs_direct_high_precision"] = vs_direct(direct_k, ref_cols)
    return out


def measure_film(fa, sc, cols, steps, warmup, threads, cpu_seconds):
    """The 64-keyframe film: one shared orbit, per-frame or shared BLA tables, rendered by
    one graph of persistent batch launches or by one launch per frame."""
    import numpy as np
    from fractodactyl_amd.scenes import film_widths
    cfg = sc["config"]
    widths = film_widths(sc)
    views = [fa.View(sc["center_re"], sc["center_im"], w, 0.0) for w in widths]
    params = fa.Params(cfg["nx"], cfg["ny"], 1, cfg["max_iter"], cfg["escape_radius"], cols)
    out = {"workload": f"film: {len(views)} keyframes {cfg['nx']}x{cfg['ny']}, widths "
                       f"{widths[0]} .. {widths[-1]} (geometric) toward the period-"
                       f"{sc['nucleus']['period']} nucleus, one shared orbit, max_iter "
                       f"{cfg['max_iter']}, BLA eps 2^-50, FP64 deltas, nu/de/normal"}
    for tables, tname in ((fa.TABLES_PER_FRAME, "per_frame_tables"),
                          (fa.TABLES_SHARED, "shared_table")):
        t = time.perf_counter()
        film = fa.Film(views, params, tables=tables)
        open_s = time.perf_counter() - t
        res = {"open_seconds": open_s}
        # the persistent grid of the batch launch, autotuned (untimed)
        grid, grid_ms, default_ms = film.tune()
        res["batch_launch_tuning"] = dict(grid=grid or "default (8 per CU)",
                                          tuned_kernel_ms=grid_ms,
                                          default_kernel_ms=default_ms,
                                          candidates="grids 1024, 2048, 4096 workgroups")
        for launch, lname in ((fa.LAUNCH_BATCH, "one_graph_batch_launch"),
                              (fa.LAUNCH_PER_FRAME, "launch_per_frame")):
            g = grid if launch == fa.LAUNCH_BATCH else 0
            for i in range(warmup):
                st = film.run(launch, g)
            walls, kerns = [], []
            for i in range(steps):
                t0 = time.perf_counter()
                st = film.run(launch, g)
                walls.append(time.perf_counter() - t0)
                kerns.append(st["kernel_seconds"])
                log(f"film {tname} {lname} step {i}: kernel {kerns[-1] * 1e3:.2f} ms")
            w = sum(walls) / len(walls)
            res[lname] = dict(ms_per_film=w * 1e3, kernel_ms=sum(kerns) / len(kerns) * 1e3,
                              value=st["equivalent"] / w, equivalent_iterations=st["equivalent"],
                              plain_steps=st["plain_steps"], blocks=st["blocks"],
                              max_sample_ops=st["max_sample_ops"],
                              lane_utilisation=st["lane_iterations"] /
                              max(1, 64 * st["wave_iterations"]))
        res["reference_seconds"] = st["reference_seconds"]
        res["reference_bits"] = st["reference_bits"]
        res["reference_len"] = st["reference_len"]
        res["tables_seconds"] = st["bla_seconds"]
        res["batch_speedup_over_per_frame_launches"] = (
            res["launch_per_frame"]["ms_per_film"] / res["one_graph_batch_launch"]["ms_per_film"])
        if tables == fa.TABLES_PER_FRAME:
            eq, esc = [], []
            for k in range(len(views)):
                _, c, fs = film.read(k)
                eq.append(fs["equivalent"])
                esc.append(float(((c["class"] & 3) == 0).mean()))
            res["per_frame_equivalent_min_median_max"] = [min(eq), float(np.median(eq)),
                                                          max(eq)]
            res["per_frame_escaped_fraction_min_max"] = [min(esc), max(esc)]
        film.close()
        out[tname] = res
    if cpu_seconds > 0:
        # the oracle on rows of every 8th keyframe, from the same shared orbit
        import oracle
        deepest = views[-1]
        zr, zi, bits = oracle.reference(deepest.center_re, deepest.center_im, deepest.width,
                                        0.0, params.nx, params.ny, params.max_iter,
                                        params.escape_radius)
        vals = [cpu_baseline(views[k], params, True, cpu_seconds / 8, threads,
                             orbit=(zr, zi, bits), runs=1) for k in range(0, len(views), 8)]
        out["cpu_baseline"] = dict(value=float(np.mean([v["value"] for v in vals])),
                                   unit="pixel-iter/s", cores=threads, kind="port",
                                   sample="mean over keyframes 0, 8, .., 56 of: " +
                                          vals[0]["sample"].split(",")[0] + " (first frame)")
    return out


# The reference's series-initializer frames (docs/2026-10-20/series-frames/results/
# frames.txt): one centre on the PROB-22 zoom path, 24 widths from 5e-15 to 2.4e-29.
SERIES_CENTRE = ("-0.8296874999999996981089858669738646984522311568651036619209",
                 "0.1859374999999989417838556112859064221737945709381783514212")
SERIES_WIDTHS = ["5.0118723e-15", "1.2589254e-15", "3.1622777e-16", "7.9432823e-17",
                 "1.9952623e-17", "5.0118723e-18", "1.2589254e-18", "3.1622777e-19",
                 "7.9432823e-20", "1.9952623e-20", "5.0118723e-21", "1.2589254e-21",
                 "3.1622777e-22", "7.9432823e-23", "1.9952623e-23", "5.0118723e-24",
                 "1.2589254e-24", "3.1622777e-25", "7.9432823e-26", "1.9952623e-26",
                 "4.216965e-27", "7.4989421e-28", "1.3335214e-28", "2.3713737e-29"]


def measure_series(fa, cols, reps=3):
    """The four arms of the reference's series study (plain, bla, series, series-bla) on
    its 24 frames at 480x270, max_iter 1e5, on the GPU: per arm the median device time
    (table build + kernel) per frame, the host series build, and the series arms'
    agreement with the plain render (class mismatches, worst nu shift in output pixels
    where de > 1e-3, fd compare's contract)."""
    import numpy as np
    arms = ("plain", "bla", "series", "series_bla", "auto", "auto_bla")
    rows, tot = [], {k: 0.0 for k in arms + ("prepare",)}
    for w in SERIES_WIDTHS:
        v = fa.View(SERIES_CENTRE[0], SERIES_CENTRE[1], w, 0.0)
        p = fa.Params(480, 270, 1, 100_000, 1e10, cols)
        re, im, _ = fa.reference(v, p)
        t = time.perf_counter()
        sr = fa.series_build(re, im, fa.bla_dc_max(v, p), float(w) / p.nx, p.max_iter)
        build_ms = (time.perf_counter() - t) * 1e3
        row = dict(width=w, series_n=sr.n, orders=len(sr.b), stop=sr.stop,
                   series_build_ms=build_ms)
        outs = {}
        for arm, route, use in (("plain", fa.ROUTE_PLAIN, False), ("bla", fa.ROUTE_BLA, False),
                                ("series", fa.ROUTE_PLAIN, True),
                                ("series_bla", fa.ROUTE_BLA, True),
                                ("auto", fa.ROUTE_PLAIN, "auto"),
                                ("auto_bla", fa.ROUTE_BLA, "auto")):
            with fa.ResidentFrame(v, p, route) as f:
                if use == "auto":
                    t = time.perf_counter()
                    _, ok = f.series_prepare()
                    prep = (time.perf_counter() - t) * 1e3
                    row[arm + "_admitted"] = ok
                    row[arm + "_prepare_ms"] = prep
                    tot["prepare"] += prep
                elif use and sr.n > 0:
                    f.series(sr)
                times = []
                for _ in range(reps):
                    st = f.run()
                    times.append(st["kernel_seconds"] + st["bla_seconds"])
                _, c, _ = f.read()
            outs[arm] = c
            ms = float(np.median(times)) * 1e3
            row[arm + "_ms"] = ms
            tot[arm] += ms
        a, de = outs["plain"]["nu"], outs["plain"]["de"].astype(np.float64)
        sel = de > 1e-3
        for arm in ("series", "series_bla", "auto", "auto_bla"):
            b = outs[arm]["nu"]
            row[arm + "_class_mismatch"] = int((outs[arm]["class"] != outs["plain"]["class"])
                                              .sum())
            shift = np.abs(a - b) * de * math.log(2) / 2.0
            row[arm + "_worst_nu_px"] = float(shift[sel].max()) if sel.any() else 0.0
        log(json.dumps(row))
        rows.append(row)
    tot["series_build"] = sum(r["series_build_ms"] for r in rows)
    return dict(workload="24 frames of the reference's series study (frames.txt), 480x270, "
                         "max_iter 1e5, nu/de/normal; device ms = table build + kernel, "
                         f"median of {reps}; auto = fdg_frame_series_prepare's choice "
                         "(its probe runs reported apart as prepare_ms)",
                total_ms=tot, frames=rows)


def copy_probe(nbytes=1 << 28, reps=3):
    """A plain device-to-device copy of known size (torch's copy kernel): profiling its
    FETCH_SIZE / WRITE_SIZE counters gives the scale that turns those counters into HBM
    bytes for the perturbation kernel's traffic figure."""
    import torch
    x = torch.ones(nbytes, dtype=torch.uint8, device="cuda")
    y = torch.empty_like(x)
    for _ in range(reps):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": nbytes, "reps": reps}), flush=True)


def scene_view(fa, sc, cols):
    c = sc["config"]
    return (fa.View(sc["center_re"], sc["center_im"], c["width"], 0.0),
            fa.Params(c["nx"], c["ny"], 1, c["max_iter"], c["escape_radius"], cols))


def workload_name(k, sc, params):
    c = sc["config"]
    if k == 3:
        return (f"config3: {params.nx}x{params.ny} width {c['width']} at the period-"
                f"{sc['nucleus']['period']} ladder nucleus (its minibrot ~1e-1000), max_iter "
                f"{params.max_iter}, BLA eps 2^-50, scaled tier (f64 mantissa x 2^e deltas), "
                f"dz/dc carried, nu/de/normal")
    return (f"config{k}: {params.nx}x{params.ny} width {c['width']} at nucleus+{c['shift']}w "
            f"(period {sc['nucleus']['period']}, minibrot ~1e{sc['nucleus']['log10_size']:.0f}), "
            f"max_iter {params.max_iter}, BLA eps 2^-50, FP64 deltas, nu/de/normal")


def measure_config5(fa, sc, cols, steps, warmup):
    """The 1e-3000 jet-AA frame at full size: 3840x2160 pixels at 2x2 samples (7680x4320
    samples), max_iter 1e7, BLA, dz/dc, through fdg_frame_jet (pass 1 on each pixel's
    top-left sample, the host rebuild, pass 2 on the rejected samples, merge) with the
    frame's 2^20-point orbit (~10,100 bits) and BLA table resident in HBM.  The host
    reference orbit is timed apart (frame open).  Every sample of this frame stays bounded:
    it rebases onto the orbit's exact f64 zero at the nucleus period and restarts from dc,
    so it is Unresolved at max_iter with 1e7 equivalent iterations, nearly all BLA-skipped."""
    c = sc["config"]
    v = fa.View(sc["center_re"], sc["center_im"], c["width"], 0.0)
    p = fa.Params(c["nx"], c["ny"], c["ss"], c["max_iter"], c["escape_radius"], cols)
    t = time.perf_counter()
    f = fa.ResidentFrame(v, p, fa.ROUTE_BLA, fa.FORM_F64, fa.EPS)
    open_s = time.perf_counter() - t
    log(f"config5: frame open (orbit + upload) {open_s:.2f}s")
    for i in range(warmup):
        _, _, st, done = f.jet()
        log(f"config5 warmup {i}: kernel {st['kernel_seconds'] * 1e3:.1f} ms, {done} samples")
    walls, stats = [], []
    for i in range(steps):
        t = time.perf_counter()
        h, out, st, done = f.jet()
        walls.append(time.perf_counter() - t)
        stats.append(st)
    for i, (w, st) in enumerate(zip(walls, stats)):
        log(f"config5 step {i}: jet frame {w * 1e3:.1f} ms, kernel "
            f"{st['kernel_seconds'] * 1e3:.1f} ms")
    f.close()
    st = stats[-1]
    wall = sum(walls) / len(walls)
    cls = out["class"] & 3
    return dict(
        workload=(f"config5: {c['nx'] // 2}x{c['ny'] // 2} pixels at 2x2 samples, width "
                  f"{c['width']} at a seeded satellite nucleus of period "
                  f"{sc['nucleus']['period']} ({sc['nucleus']['bits']} bits), max_iter "
                  f"{c['max_iter']}, BLA eps 2^-50, scaled tier with dz/dc, jet-AA frame "
                  f"(fdg_frame_jet)"),
        ms_per_frame=wall * 1e3, value=st["equivalent"] / wall, unit="pixel-iter/s",
        kernel_ms=st["kernel_seconds"] * 1e3, table_ms=st["bla_seconds"] * 1e3,
        host_rebuild_and_encode_ms=st["encode_seconds"] * 1e3,
        transfer_ms=st["transfer_seconds"] * 1e3,
        samples_computed=done, samples=c["nx"] * c["ny"],
        equivalent_iterations=st["equivalent"], plain_steps=st["plain_steps"],
        blocks=st["blocks"], skipped=st["skipped"], kernel_id=h["kernel"],
        reference_len=st["reference_len"], reference_bits=st["reference_bits"],
        reference_seconds=st["reference_seconds"], frame_open_seconds=open_s,
        classes={k: float((cls == val).mean()) for k, val in (("escaped", 0), ("interior", 1),
                                                              ("unresolved", 2))})


def measure_plain_fp64(fa, view, params, label, row_stride, forms_ms):
    """The plain FP64 GPU form: perturbation with FP64 deltas and no BLA (ROUTE_PLAIN),
    every step executed, on every row_stride-th row of the same frame (a sample mask; one
    pass), its kernel time scaled to the whole frame by the rows computed.  forms_ms: the
    BLA forms' kernel ms on the whole frame, reported as speed-ups over this form."""
    import numpy as np
    t = time.perf_counter()
    frame = fa.ResidentFrame(view, params, fa.ROUTE_PLAIN, fa.FORM_F64, fa.EPS)
    log(f"{label}-plain: frame open {time.perf_counter() - t:.2f}s")
    mask = np.zeros((params.ny, params.nx), np.uint8)
    mask[::row_stride] = 1
    frame.mask(mask.ravel())
    st = frame.run()
    frame.close()
    rows = int(mask[:, 0].sum())
    ms = st["kernel_seconds"] * 1e3
    est = ms * params.ny / rows
    log(f"{label}-plain: {rows} rows, kernel {ms:.1f} ms, frame estimate {est:.0f} ms")
    out = {"route": "plain perturbation (no BLA), FP64 deltas",
           "sample": f"{rows} of {params.ny} rows (every {row_stride}th) x {params.nx} samples, "
                     f"one pass",
           "kernel_ms_sample": ms, "ms_per_frame_estimate": est,
           "value": st["equivalent"] / st["kernel_seconds"],
           "equivalent_iterations_sample": st["equivalent"], "plain_steps_sample": st["plain_steps"]}
    out["speedup_of_bla_forms"] = {k: est / v for k, v in forms_ms.items()}
    return out


def fe32_roofline(m, forms):
    """The FE32 form's kernel against the FP32 vector peak (same algorithmic flops)."""
    r = m["roofline"]
    out = {}
    for name in ("fe32", "mixed"):
        if name in forms:
            ach = r["flops_per_launch"] / (forms[name]["kernel_ms"] * 1e-3) / 1e12
            out[name] = {"bound": "valu-fp32", "achieved": ach, "peak": PEAK_FP32_TFLOPS,
                         "unit": "TFLOP/s", "frac": ach / PEAK_FP32_TFLOPS,
                         "frac_of_unpacked_fp32": ach / PEAK_FP64_TFLOPS,
                         "flops_model": r["flops_model"]}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", type=int, default=3, choices=(2, 3),
                    help="the headline frame: 3 (1e-1000, 1M iterations, the north star's "
                         "target) or 2 (1e-100)")
    ap.add_argument("--cpu-seconds", type=float, default=8.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--frames", default="config,other,escape,film,config5",
                    help="which frames to run: config (the JSON line); companions other "
                         "(the other of configs 2 and 3), escape (reference zoom path at "
                         "2e-49), film (64 keyframes sharing one orbit), config5 (the "
                         "1e-3000 jet-AA frame); series (the reference's 24 "
                         "series-study frames, companion mode only)")
    ap.add_argument("--forms", action="store_true",
                    help="with companions only: also run the FE32 and mixed forms")
    ap.add_argument("--launch", default="auto",
                    help="launch of the frames' perturbation kernel: auto (the autotuner, "
                         "untimed, per frame), default, or G,D,C[,R] = groups per CU, claim "
                         "divisor, claim cap, refill (for profiling a chosen launch)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a 256 MiB device copy (the traffic counters' calibration)")
    a = ap.parse_args()
    global LAUNCH
    LAUNCH = a.launch
    if a.copy_probe:
        copy_probe()
        return
    if a.gpus != 1:
        print(f"bench.py: this path runs on one GPU; --gpus {a.gpus} is not supported "
              "(a frame shards by tiles of samples; run one process per GPU on a share of the "
              "tiles)", file=sys.stderr)
        sys.exit(2)

    import fractodactyl_amd as fa
    from fractodactyl_amd.scenes import scene as get_scene
    cols = fa.COL_NU | fa.COL_DE | fa.COL_NORMAL
    sc = get_scene(a.config)
    view, params = scene_view(fa, sc, cols)
    route = fa.ROUTE_BLA if sc["config"]["bla"] else fa.ROUTE_PLAIN
    other = 2 if a.config == 3 else 3
    fa.default_context()
    frames = a.frames.split(",")
    threads = cpu_threads()
    cpu_s = 0.0 if a.no_cpu_baseline else a.cpu_seconds
    ka = json.load(open(os.path.join(ROOT, "tests", "golden", "known_answers.json")))
    # Escape-dominated companion: the reference's own deep-zoom path (bench/path-atlas-v0.txt,
    # the period-764 minibrot's neighbourhood) at its frame-749 width, at 1080p.
    v0 = fa.View(ka["v0_centre"]["re"], ka["v0_centre"]["im"], "2e-49", 0.0)
    p0 = fa.Params(1920, 1080, 1, 100_000, 1e10, cols)
    if "config" not in frames:
        # profiling aid: companions alone, each printed as JSON on stderr, no bench line
        todo = []
        if "escape" in frames:
            todo.append(("escape", v0, p0))
        if "deep" in frames or "config3" in frames:
            todo.append(("deep",) + scene_view(fa, get_scene(3), cols))
        if "config2" in frames:
            todo.append(("config2",) + scene_view(fa, get_scene(2), cols))
        for name, v, p in todo:
            keep = []
            m = measure(fa, v, p, fa.ROUTE_BLA, a.steps, a.warmup, name, keep=keep)
            if a.forms:
                m["precision_forms"] = measure_forms(fa, v, p, fa.ROUTE_BLA, a.steps, a.warmup,
                                                     name, m, keep[0])
            log(json.dumps(m))
        if "film" in frames:
            log(json.dumps(measure_film(fa, get_scene(4), cols, a.steps, a.warmup, 1, 0.0)))
        if "series" in frames:
            log(json.dumps(measure_series(fa, cols)))
        if "config5" in frames:
            log(json.dumps(measure_config5(fa, get_scene(5), cols, a.steps, a.warmup)))
        return
    keep = []
    m = measure(fa, view, params, route, a.steps, a.warmup, f"config{a.config}", keep=keep)
    forms = measure_forms(fa, view, params, route, a.steps, a.warmup, f"config{a.config}", m,
                          keep[0], direct_k=a.config)
    comp = {}
    if "other" in frames:
        so = get_scene(other)
        vo, po = scene_view(fa, so, cols)
        keepo = []
        mo = measure(fa, vo, po, fa.ROUTE_BLA, a.steps, a.warmup, f"config{other}", keep=keepo)
        mo["precision_forms"] = measure_forms(fa, vo, po, fa.ROUTE_BLA, a.steps, a.warmup,
                                              f"config{other}", mo, keepo[0], direct_k=other)
        mo["roofline"]["traffic"] = measured_traffic(other)
        mo["workload"] = workload_name(other, so, po)
        if cpu_s > 0:
            mo["cpu_baseline"] = cpu_baseline(vo, po, True, cpu_s / 2, threads, runs=1)
        mo["end_to_end"] = end_to_end(fa, vo, po, fa.ROUTE_BLA)
        mo["end_to_end_device_encode"] = end_to_end_device_encode(fa, vo, po, fa.ROUTE_BLA)
        comp[f"config{other}_frame"] = mo
    if "escape" in frames:
        keep0 = []
        m0 = measure(fa, v0, p0, fa.ROUTE_BLA, a.steps, a.warmup, "v0-2e-49", keep=keep0)
        m0["precision_forms"] = measure_forms(fa, v0, p0, fa.ROUTE_BLA, a.steps, a.warmup,
                                              "v0-2e-49", m0, keep0[0])
        m0["workload"] = ("reference deep-zoom path (v0 centre, period-764 minibrot) at width "
                          "2e-49, 1920x1080, max_iter 1e5, BLA")
        if cpu_s > 0:
            m0["cpu_baseline"] = cpu_baseline(v0, p0, True, cpu_s / 2, threads, runs=1)
        comp["escape_frame"] = m0
    if "film" in frames:
        comp["film"] = measure_film(fa, get_scene(4), cols, a.steps, a.warmup, threads,
                                    cpu_s / 2)
    if "config5" in frames:
        comp["config5_frame"] = measure_config5(fa, get_scene(5), cols, min(a.steps, 3), 1)
    # the plain FP64 GPU form (no BLA) on a row sample of the headline frame
    plain = measure_plain_fp64(fa, view, params, f"config{a.config}", 16 if a.config == 3 else 1,
                               {"fp64_bla": m["kernel_ms"], "fe32_bla": forms["fe32"]["kernel_ms"],
                                "mixed_bla": forms["mixed"]["kernel_ms"]})
    out = {
        "metric": "effective pixel-iter/s at 1080p deep zoom (BLA-skipped steps included)",
        "value": m["value"],
        "unit": "pixel-iter/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": m["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (scene fractodactyl_amd/scenes.json: the reference's ladder nucleus "
                "for configs 3 and 4, seeded Newton nuclei for configs 1, 2 and 5)",
        "config": {
            "workload": workload_name(a.config, sc, params),
            **{k: v for k, v in m.items() if k not in ("value", "ms_per_step", "roofline")},
            "precision_forms": forms,
            "precision_forms_roofline": fe32_roofline(m, forms),
            "plain_fp64_gpu": plain,
            **comp,
        },
        "roofline": dict(m["roofline"], traffic=measured_traffic(a.config)),
    }
    # PCIe- and host-inclusive: the whole fdg_render call on the headline frame
    out["config"]["end_to_end"] = end_to_end(fa, view, params, route)
    out["config"]["end_to_end_device_encode"] = end_to_end_device_encode(fa, view, params, route)
    if cpu_s > 0:
        t = time.perf_counter()
        out["cpu_baseline"] = cpu_baseline(view, params, route == fa.ROUTE_BLA, cpu_s, threads)
        log(f"cpu baseline: {time.perf_counter() - t:.1f}s")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
