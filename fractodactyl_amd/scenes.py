"""Seeded deep-zoom scenes: Newton nuclei of chosen period and the five benchmark frames.

A nucleus of period P is a root of z_P(c) = 0 (z_0 = 0, z_{k+1} = z_k^2 + c).  The search
is seeded and reproducible:

1. draw the target period P* and two points of the seed box;
2. bisect between an escaping and a bounded point towards the boundary (fixed point at
   the working precision) until the escaping end survives past the period range;
3. take that point's orbit's record-minimum ("partial") iterates |z_p| and pick the p in
   the period range closest to P*;
4. Newton on z_p(c) = 0 from that point, then check the period is exactly p and estimate
   the minibrot size (size ~ 1/|beta lambda^2|, lambda = prod 2 z_k, beta = sum 1/lambda_k).

Arithmetic is Python integers as fixed point with ``bits`` fraction bits.  Results are
cached in ``scenes.json`` next to this file so benchmark runs never repeat the search.
"""
import json
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
CACHE = os.path.join(HERE, "scenes.json")


def _fx(x, b):
    """float or decimal string -> fixed point int."""
    from decimal import Decimal, getcontext
    getcontext().prec = b // 3 + 50
    return int((Decimal(str(x)) * (Decimal(2) ** b)).to_integral_value())


def _dec(v, b, digits):
    from decimal import Decimal, getcontext
    getcontext().prec = digits + 30
    d = Decimal(v) / (Decimal(2) ** b)
    return format(d, "." + str(digits) + "f") if abs(d) >= 1e-5 else format(d, "." + str(digits) + "e")


def escape_time(cr, ci, b, nmax):
    """Iterations until |z|^2 > 4 (None: bounded for nmax)."""
    one4 = 4 << b
    x = y = 0
    for n in range(1, nmax + 1):
        x, y = ((x * x - y * y) >> b) + cr, ((x * y) >> (b - 1)) + ci
        if x * x + y * y > (one4 << b):
            return n
    return None


def partials(cr, ci, b, nmax):
    """Iterates at which |z_n| reaches a new minimum (atom-domain periods)."""
    x = y = 0
    best = None
    out = []
    for n in range(1, nmax + 1):
        x, y = ((x * x - y * y) >> b) + cr, ((x * y) >> (b - 1)) + ci
        m = x * x + y * y
        if m > (4 << (2 * b)):
            break
        if best is None or m < best:
            best = m
            out.append(n)
    return out


def newton(cr, ci, p, b, iters=80):
    """Newton on z_p(c) = 0 in fixed point; returns (cr, ci, converged)."""
    one = 1 << b
    tol = 1 << 24
    for _ in range(iters):
        x = y = 0
        dx = dy = 0
        for _k in range(p):
            # dz' = 2 z dz + 1, z' = z^2 + c
            dx, dy = ((x * dx - y * dy) >> (b - 1)) + one, ((x * dy + y * dx) >> (b - 1))
            x, y = ((x * x - y * y) >> b) + cr, ((x * y) >> (b - 1)) + ci
            if abs(x) > (1 << (b + 40)) or abs(y) > (1 << (b + 40)):
                return cr, ci, False
        den = (dx * dx + dy * dy) >> b
        if den == 0:
            return cr, ci, False
        # (x + iy) / (dx + i dy)
        sr = ((x * dx + y * dy)) // den
        si = ((y * dx - x * dy)) // den
        cr -= sr
        ci -= si
        if abs(sr) < tol and abs(si) < tol:
            return cr, ci, True
    return cr, ci, False


def nucleus_check(cr, ci, p, b):
    """(|z_p| as log2, smallest divisor period returning within 2^-b/4, size estimate)."""
    import mpmath
    mpmath.mp.prec = 64
    x = y = 0
    lam = mpmath.mpc(1)
    beta = mpmath.mpc(1)
    zmin_div = None
    for k in range(1, p + 1):
        x, y = ((x * x - y * y) >> b) + cr, ((x * y) >> (b - 1)) + ci
        z = mpmath.mpc(mpmath.mpf(x) / 2 ** b, mpmath.mpf(y) / 2 ** b)
        if k < p:
            lam = 2 * z * lam
            beta = beta + 1 / lam
            if p % k == 0 and abs(z) < mpmath.mpf(2) ** (-b // 4):
                zmin_div = k
    zp = mpmath.mpf(x * x + y * y) ** 0.5 / 2 ** b
    size = 1 / abs(beta * lam * lam)
    return float(mpmath.log(zp + mpmath.mpf(2) ** (-b - 10), 2)), zmin_div, float(mpmath.log10(size))


def find_nucleus(seed, pmin, pmax, bits, box=((-1.8, 0.4), (0.05, 0.9)), tries=400,
                 max_log10_size=None):
    """A seeded nucleus with period in [pmin, pmax] (closest partial to a seeded target),
    optionally with a minibrot smaller than 10^max_log10_size."""
    rng = np.random.default_rng(seed)
    target = int(rng.integers(pmin, pmax + 1))
    b = bits + 32
    for attempt in range(tries):
        a = (rng.uniform(*box[0]), rng.uniform(*box[1]))
        c = (rng.uniform(*box[0]), rng.uniform(*box[1]))
        ar, ai, cr_, ci_ = (_fx(v, b) for v in (*a, *c))
        ea = escape_time(ar, ai, b, 64)
        ec = escape_time(cr_, ci_, b, 64)
        if (ea is None) == (ec is None):
            continue
        out, inn = ((ar, ai), (cr_, ci_)) if ea is not None else ((cr_, ci_), (ar, ai))
        # bisect towards the boundary until the escaping end survives past pmax
        lim = int(pmax * 1.5)
        for _ in range(b - 8):
            mr, mi = (out[0] + inn[0]) >> 1, (out[1] + inn[1]) >> 1
            e = escape_time(mr, mi, b, lim)
            if e is None:
                inn = (mr, mi)
            else:
                out = (mr, mi)
                if e > lim * 2 // 3:
                    break
        parts = [q for q in partials(out[0], out[1], b, lim) if pmin <= q <= pmax]
        for p in sorted(parts, key=lambda q: abs(q - target))[:3]:
            nr, ni, ok = newton(out[0], out[1], p, b)
            if not ok:
                continue
            # Newton's last step was below 2^(24 - b): c is the nucleus to ~bits bits.
            lz, div, lsize = nucleus_check(nr, ni, p, b)
            if div is not None:
                continue
            if max_log10_size is not None and lsize > max_log10_size:
                continue
            digits = int(bits * math.log10(2)) + 4
            return dict(seed=seed, target_period=target, period=p, bits=bits,
                        attempts=attempt + 1, re=_dec(nr, b, digits), im=_dec(ni, b, digits),
                        log2_zp=lz, log10_size=lsize, max_log10_size=max_log10_size)
    raise RuntimeError("no nucleus found")


# The five benchmark frames (the seeds are the frame numbers).
CONFIGS = {
    1: dict(pmin=200, pmax=2000, bits=256, width="1e-20", nx=480, ny=270, max_iter=20_000,
            escape_radius=1e3, bla=False, shift=0.0, max_log10_size=-21),
    2: dict(pmin=1000, pmax=10000, bits=512, width="1e-100", nx=1920, ny=1080,
            max_iter=100_000, escape_radius=1e10, bla=True, shift=0.3),
    5: dict(pmin=20_000, pmax=200_000, bits=10_500, width="1e-3000", nx=7680, ny=4320, ss=2,
            max_iter=10_000_000, escape_radius=1e10, bla=True, shift=0.0),
}


def satellite_nucleus(seed, pmin, pmax, bits):
    """A seeded nucleus of period P in [pmin, pmax]: the satellite component of the main
    cardioid at internal angle p/P (p seeded, coprime with P) has exact period P and its
    nucleus lies next to the cardioid point c(t) = e^{it}/2 - e^{2it}/4, t = 2 pi p / P.
    Newton on z_P(c) = 0 from that point at 96 bits, then once more at each doubled
    precision up to `bits` (+32 guard bits); the result's period is checked to be exactly P
    (no return to within 2^-64 at a proper divisor).  Python-integer fixed point, as
    find_nucleus; a 10,500-bit nucleus of period ~1e5 takes about a minute."""
    rng = np.random.default_rng(seed)
    P = int(rng.integers(pmin, pmax + 1))
    while True:
        q = int(rng.integers(1, P))
        if math.gcd(q, P) == 1:
            break
    t = 2 * math.pi * q / P
    c0 = (0.5 * math.cos(t) - 0.25 * math.cos(2 * t), 0.5 * math.sin(t) - 0.25 * math.sin(2 * t))
    b = 96
    cr, ci = _fx(repr(c0[0]), b), _fx(repr(c0[1]), b)
    cr, ci, ok = newton(cr, ci, P, b, iters=60)
    if not ok:
        raise RuntimeError("Newton did not converge at 96 bits")
    target = bits + 32
    while b < target:
        nb = min(2 * b, target)
        cr, ci = cr << (nb - b), ci << (nb - b)
        b = nb
        cr, ci, ok = newton(cr, ci, P, b, iters=3)
        if not ok:
            raise RuntimeError(f"Newton did not converge at {b} bits")
    # exact period: no proper divisor returns near 0
    divs = {d for k in range(1, int(math.isqrt(P)) + 1) if P % k == 0 for d in (k, P // k)}
    divs.discard(P)
    x = y = 0
    near = 1 << (b - 64)
    for k in range(1, P + 1):
        x, y = ((x * x - y * y) >> b) + cr, ((x * y) >> (b - 1)) + ci
        if k in divs and abs(x) < near and abs(y) < near:
            raise RuntimeError(f"period divides {k}")
    digits = int(bits * math.log10(2)) + 4
    return dict(seed=seed, period=P, angle=f"{q}/{P}", bits=bits, re=_dec(cr, b, digits),
                im=_dec(ci, b, digits), log10_size=-2.0 * math.log10(P),
                source=f"satellite of the main cardioid at internal angle {q}/{P}, Newton "
                       f"from the cardioid point, refined by precision doubling")


def _offset_centre(nuc, shift, width, seed, bits):
    """Centre = nucleus + shift * width * e^{i theta}, theta seeded, as exact decimals."""
    from decimal import Decimal, getcontext
    getcontext().prec = int(bits * 0.302) + 40
    th = np.random.default_rng(seed + 1000).uniform(0, 2 * math.pi)
    w = Decimal(width)
    dr = Decimal(repr(shift * math.cos(th))) * w
    di = Decimal(repr(shift * math.sin(th))) * w
    digits = int(bits * 0.302) + 10
    return (format(Decimal(nuc["re"]) + dr, "." + str(digits) + "f"),
            format(Decimal(nuc["im"]) + di, "." + str(digits) + "f"), th)


def scene(k, generate=False):
    """The benchmark frame k: dict with view strings, grid, iterations and provenance.

    Frame 3 (1e-1000) uses the reference's own ladder nucleus of period 16116 (its minibrot
    is ~1e-1000 wide), stored in scenes.json: the bisection search above only reaches
    near-parabolic boundary points, whose minibrots are far larger than 1e-1000.  Scene 4
    (the 64-keyframe film, widths 1e-10 .. 1e-300) zooms toward the same nucleus; see
    film_widths."""
    cache = json.load(open(CACHE)) if os.path.exists(CACHE) else {}
    key = str(k)
    if key not in cache:
        if not generate:
            raise KeyError(f"scene {k} not in {CACHE}; run python -m fractodactyl_amd.scenes {k}")
        cfg = CONFIGS[k]
        if k == 5:
            nuc = satellite_nucleus(k, cfg["pmin"], cfg["pmax"], cfg["bits"])
        else:
            nuc = find_nucleus(k, cfg["pmin"], cfg["pmax"], cfg["bits"],
                               max_log10_size=cfg.get("max_log10_size"), tries=4000)
        re_, im_, th = _offset_centre(nuc, cfg["shift"], cfg["width"], k, cfg["bits"])
        cache[key] = dict(config=cfg, nucleus=nuc, center_re=re_, center_im=im_, theta=th)
        json.dump(cache, open(CACHE, "w"), indent=1)
    return cache[key]


def film_widths(sc):
    """Keyframe widths w_k = w_first (w_last / w_first)^(k / (frames - 1)), as decimals
    (mantissa and exponent apart, so widths below 1e-308 print too)."""
    cfg = sc["config"]
    a, b = math.log10(float(cfg["width_first"])), math.log10(float(cfg["width_last"]))
    n = cfg["frames"]
    out = []
    for k in range(n):
        x = a + (b - a) * k / (n - 1)
        e = math.floor(x)
        out.append(f"{10.0 ** (x - e):.6f}e{e}")
    return out


if __name__ == "__main__":
    import sys
    import time
    for a in sys.argv[1:]:
        t = time.time()
        s = scene(int(a), generate=True)
        print(a, s["nucleus"]["period"], s["nucleus"]["log10_size"], f"{time.time() - t:.1f}s")
