"""Benchmark: effective pixel-iterations/s of the deep-zoom perturbation/BLA path on one GPU.

Workload (default): the 1920x1080 frame of width 1e-100 centred 0.3 widths from a seeded
Newton nucleus, max_iter 100,000, escape radius 1e10, BLA on (eps 2^-50), columns
nu/de/normal (so dz/dc is carried), FP64 deltas.  A step is one pass of the hot path over
the frame with its reference orbit resident in HBM: the GPU BLA table build plus the
persistent perturbation kernel over all 2,073,600 samples.  value = plain-equivalent
iterations (BLA-skipped steps included, the reference's `iterations.equivalent`) per
second of wall time over the timed steps.

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

PEAK_FP64_TFLOPS = 78.6      # MI355X vector FP64
F_PLAIN, F_BLOCK = 18, 22    # algorithmic flops per plain step / applied block
F_PLAIN_D, F_BLOCK_D = 27, 30  # with dz/dc carried


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(view, params, route_bla, target_seconds, threads):
    """The oracle (C port of the reference CPU BLA renderer) on a row sample of the frame."""
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
                          bla=route_bla, threads=threads, mask=mask, outcomes=False)
        return f["stats"], len(rows)

    st, rows = run(ny // 4, ny // 8)   # probe: 4 rows
    per_row = st["render_seconds"] / rows
    want_rows = max(4, min(ny, int(target_seconds / max(per_row, 1e-6))))
    stride = max(1, ny // want_rows)
    st, rows = run(stride, stride // 2)
    return dict(value=st["equivalent"] / st["render_seconds"], unit="pixel-iter/s",
                cores=threads, kind="port",
                sample=f"{rows} of {ny} rows (every {stride}th) x {nx} samples of the same "
                       f"frame, {'BLA eps 2^-50' if route_bla else 'plain'}, oracle render "
                       f"{st['render_seconds']:.2f}s (reference orbit "
                       f"{st['reference_seconds']:.3f}s, table {st['bla_seconds']:.3f}s "
                       f"reported apart)")


def measure(fa, view, params, route, steps, warmup, label):
    """Open the frame (orbit resident in HBM), warm up, time `steps` passes."""
    t = time.perf_counter()
    frame = fa.ResidentFrame(view, params, route, fa.FORM_F64, fa.EPS)
    log(f"{label}: frame open (host reference orbit + upload) {time.perf_counter() - t:.2f}s")
    for i in range(warmup):
        st = frame.run()
        log(f"{label} warmup {i}: kernel {st['kernel_seconds'] * 1e3:.2f} ms, table "
            f"{st['bla_seconds'] * 1e3:.2f} ms")
    stats = []
    t0 = time.perf_counter()
    for i in range(steps):
        stats.append(frame.run())
        log(f"{label} step {i}: kernel {stats[-1]['kernel_seconds'] * 1e3:.2f} ms")
    t1 = time.perf_counter()
    h, cols, _ = frame.read(raw=True)
    frame.close()
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
                roofline={"bound": "valu-fp64", "achieved": achieved, "peak": PEAK_FP64_TFLOPS,
                          "unit": "TFLOP/s", "frac": achieved / PEAK_FP64_TFLOPS,
                          "traffic": None, "flops_per_launch": flops,
                          "flops_model": f"{fp}/plain step + {fb}/applied block"})


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", type=int, default=2)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--frames", default="config,escape",
                    help="which frames to run: config (the JSON line), escape (companion)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py: this path runs on one GPU; --gpus {a.gpus} is not supported "
              "(the frame shards by tiles, see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import fractodactyl_amd as fa
    from fractodactyl_amd.scenes import scene as get_scene
    sc = get_scene(a.config)
    cfg = sc["config"]
    cols = fa.COL_NU | fa.COL_DE | fa.COL_NORMAL
    view = fa.View(sc["center_re"], sc["center_im"], cfg["width"], 0.0)
    params = fa.Params(cfg["nx"], cfg["ny"], 1, cfg["max_iter"], cfg["escape_radius"], cols)
    route = fa.ROUTE_BLA if cfg["bla"] else fa.ROUTE_PLAIN
    fa.default_context()
    frames = a.frames.split(",")
    if "config" not in frames:
        # profiling aid: the escape companion alone, no JSON line
        ka = json.load(open(os.path.join(ROOT, "tests", "golden", "known_answers.json")))
        v0 = fa.View(ka["v0_centre"]["re"], ka["v0_centre"]["im"], "2e-49", 0.0)
        p0 = fa.Params(1920, 1080, 1, 100_000, 1e10, cols)
        m0 = measure(fa, v0, p0, fa.ROUTE_BLA, a.steps, a.warmup, "v0-2e-49")
        log(json.dumps(m0))
        return
    m = measure(fa, view, params, route, a.steps, a.warmup, f"config{a.config}")
    # Escape-dominated companion: the reference's own deep-zoom path (bench/path-atlas-v0.txt,
    # the period-764 minibrot's neighbourhood) at its frame-749 width, at 1080p.
    ka = json.load(open(os.path.join(ROOT, "tests", "golden", "known_answers.json")))
    v0 = fa.View(ka["v0_centre"]["re"], ka["v0_centre"]["im"], "2e-49", 0.0)
    p0 = fa.Params(1920, 1080, 1, 100_000, 1e10, cols)
    m0 = (measure(fa, v0, p0, fa.ROUTE_BLA, a.steps, a.warmup, "v0-2e-49")
          if "escape" in frames else {})
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
        "data": "synthetic (seeded Newton nucleus scene, fractodactyl_amd/scenes.json)",
        "config": {
            "workload": f"config{a.config}: {params.nx}x{params.ny} width {cfg['width']} at "
                        f"nucleus+{cfg['shift']}w (period {sc['nucleus']['period']}, minibrot "
                        f"~1e{sc['nucleus']['log10_size']:.0f}), max_iter {params.max_iter}, BLA "
                        f"eps 2^-50, FP64 deltas, nu/de/normal",
            **{k: v for k, v in m.items() if k not in ("value", "ms_per_step", "roofline")},
            "escape_frame": {"workload": "reference deep-zoom path (v0 centre, period-764 "
                                         "minibrot) at width 2e-49, 1920x1080, max_iter 1e5, BLA",
                             **{k: v for k, v in m0.items()}},
        },
        "roofline": m["roofline"],
    }
    if not a.no_cpu_baseline:
        threads = min(16, os.cpu_count() or 1)
        t = time.perf_counter()
        out["cpu_baseline"] = cpu_baseline(view, params, route == fa.ROUTE_BLA, a.cpu_seconds,
                                           threads)
        if "escape" in frames:
            out["config"]["escape_frame"]["cpu_baseline"] = cpu_baseline(
                v0, p0, True, a.cpu_seconds, threads)
        log(f"cpu baselines: {time.perf_counter() - t:.1f}s")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
