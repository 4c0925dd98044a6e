"""Loader and comparator for the reference's frozen mpmath truth pack (bench/truth/v1),
restating crates/fd-kernel/examples/support/mod.rs:29-372 (load, compare, displacement)."""
import gzip
import math
import os

import numpy as np

DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "truth_v1")
CASES = ["exterior", "filament", "period764", "parabolic", "extreme"]


def load(name):
    with gzip.open(os.path.join(DIR, name + ".tsv.gz"), "rt") as f:
        lines = f.read().splitlines()
    assert lines[0] == "fd-truth\t1"
    v = lines[1].split("\t")
    assert v[0] == "view" and len(v) == 12
    case = dict(name=v[1], re=v[2], im=v[3], width=v[4], nx=int(v[5]), ny=int(v[6]),
                ss=int(v[7]), max_iter=int(v[8]), escape_radius=float(v[9]),
                rotation=float(v[10]), bits=int(v[11]))
    pts = []
    for line in lines[2:]:
        if line.startswith("#"):
            continue
        f = line.split("\t")
        assert int(f[0]) == len(pts)
        pts.append((f[1], int(f[2]), [float(x) for x in f[3:8]]))
    case["points"] = pts
    return case


def edges():
    out = []
    with open(os.path.join(DIR, "precision-edges.txt")) as f:
        for line in f:
            if line.startswith("#") or line.startswith("fd-truth-edges"):
                continue
            p = line.split("\t")
            out.append((p[0], int(p[1])))
    return out


def _disp(excess, rate):
    if excess <= 0.0:
        return 0.0
    return excess / rate if rate > 0.0 else math.inf


def compare(case, cls, nu, de, normal, known_edges=()):
    """support/mod.rs:237-320 compare: returns a dict score."""
    s = dict(class_mismatches=0, known_edges=0, scalar_failures=0, nu_px=0.0, de_px=0.0,
             normal_px=0.0, mismatches=[])
    for i, (tc, it, vals) in enumerate(case["points"]):
        kind = int(cls[i]) & 3
        if tc == "E":
            ok = kind == 0
        elif tc == "A":
            ok = kind == 1
        else:
            ok = kind in (1, 2)
        if not ok:
            s["class_mismatches"] += 1
            s["mismatches"].append(i)
            if kind == 2 and tc == "E" and (case["name"], i) in known_edges:
                s["known_edges"] += 1
            continue
        if tc != "E":
            continue
        tnu, tde, tang, de_rate, n_rate = vals
        gnu = float(nu[i])
        gde = float(de[i])
        gang = float(normal[i]) * math.tau / 65536.0
        if not (math.isfinite(gnu) and math.isfinite(gde)):
            s["scalar_failures"] += 1
            continue
        nu_px = abs(gnu - tnu) * tde * math.log(2) / 2.0
        dd = abs(gde - tde)
        de_px = _disp(dd - tde * 2.0 ** -24, de_rate) if dd / tde > 1e-2 else 0.0
        nd = abs(((gang - tang + math.pi) % math.tau) - math.pi)
        n_px = _disp(nd - math.pi / 65536.0, n_rate) if nd > 1e-3 else 0.0
        s["nu_px"] = max(s["nu_px"], nu_px)
        s["de_px"] = max(s["de_px"], de_px)
        s["normal_px"] = max(s["normal_px"], n_px)
        if nu_px > 1e-3 or de_px > 1e-3 or n_px > 1e-3:
            s["scalar_failures"] += 1
    return s
