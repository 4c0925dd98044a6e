"""A Python restatement of the reference's ordinary-tier sample loop, with a trace of the
Newton interior confirmation (test infrastructure only: the KERN-05 seahorse trace test
uses it to show why two frozen samples end Unresolved under the code as written).

sample_plain_f64 follows crates/fd-kernel/src/sample.rs:170-575 (`iterate` with Binary64
deltas, no BLA, no error radii, no stripe replay): the closed-form main components
(interior.rs:6-11), the nucleus contraction heuristic (sample.rs:657-698 Contraction), the
plain step (number.rs:95-107 Binary64::delta), escape, rebase onto Z_0 (sample.rs:418-451),
resolvable (sample.rs:700-708) and PERF-08's Brent saves with holds and Newton's cycle-point
prediction (sample.rs:451-575, Multiplier at sample.rs:597-652).  Python floats are IEEE
doubles evaluated left to right without contraction, and math.log2 is the C library's, so
it reproduces the oracle's outcomes; tests/test_seahorse_trace.py checks that it does.
"""
import math

U = 2.220446049250313e-16 / 2
LO, HI = 1.0 / 4294967296.0, 4294967296.0


def exp2i(k):
    """2^k: exact where representable, 0 below the subnormals, inf above
    (crates/fd-fixed/src/lib.rs:30-42)"""
    if k > 1023:
        return math.inf
    if k < -1074:
        return 0.0
    return math.ldexp(1.0, k)


def in_main_components(cr, ci):
    """interior.rs:6-11"""
    y2 = ci * ci
    q = (cr - 0.25) * (cr - 0.25) + y2
    return q * (q + (cr - 0.25)) <= 0.25 * y2 or (cr + 1.0) * (cr + 1.0) + y2 <= 0.0625


class Multiplier:
    """sample.rs:597-652"""
    def __init__(self):
        self.re, self.im, self.exp = 1.0, 0.0, 0

    def multiply(self, re, im):
        self.re, self.im = self.re * re - self.im * im, self.re * im + self.im * re
        a = max(abs(self.re), abs(self.im))
        if not (a > 1e-120 and a < 1e120):
            if a > 0.0 and math.isfinite(a):
                e = math.floor(math.log2(a))
                s = exp2i(-e)
                self.re, self.im, self.exp = self.re * s, self.im * s, self.exp + e

    def value(self):
        s = 0.0 if self.exp < -1000 else exp2i(min(self.exp, 1000))
        return self.re * s, self.im * s

    def below_one(self):
        return (math.isfinite(self.re) and math.isfinite(self.im) and
                (self.exp < -64 or (self.exp <= 64 and
                                    (self.re * self.re + self.im * self.im) * exp2i(2 * self.exp) < 1.0)))


class Contraction:
    """sample.rs:657-698"""
    def __init__(self):
        self.re, self.im, self.exp, self.started = 0.0, 0.0, 0, False

    def multiply(self, re, im, exp):
        if not self.started:
            self.re, self.started = 1.0, True
        a = max(abs(re), abs(im))
        if a != 0.0 and not (LO <= a <= HI):
            e = math.floor(math.log2(a))
            s = exp2i(-e)
            re, im, exp = re * s, im * s, exp + e
        self.re, self.im = self.re * re - self.im * im, self.re * im + self.im * re
        self.exp += exp
        a = max(abs(self.re), abs(self.im))
        if a != 0.0 and not (LO <= a <= HI):
            e = math.floor(math.log2(a))
            s = exp2i(-e)
            self.re, self.im, self.exp = self.re * s, self.im * s, self.exp + e

    def small(self):
        return (self.started and math.isfinite(self.re) and math.isfinite(self.im) and
                (self.exp < -64 or (self.exp <= 64 and
                                    (self.re * self.re + self.im * self.im) * exp2i(2 * self.exp) < 1e-6)))


def sample_plain_f64(zr, zi, c, ar, ai, max_iter, r2, nucleus, log=None):
    """sample.rs:170-575 for the f64 tier's Binary64 deltas, no BLA: ('escaped' | 'interior' |
    'unresolved', n).  log: a list that receives every near-return of the Newton
    confirmation as dict(n, period, repeat (hold == period), below (|mu| < 1), predicted (a
    prediction stood from the last return), log2_mu, z)."""
    if c is not None and in_main_components(c[0], c[1]):
        return "interior", 0
    last = len(zr) - 1
    xr = xi = 0.0
    m = n = 0
    sr = si = 0.0
    chk, sn = 16, 8
    pending = False
    mu = Multiplier()
    hold, predicted = 0, False
    qr = qi = qstep = 0.0
    nuc = nucleus or None
    con = Contraction()
    while n < max_iter:
        fr, fi = zr[m] + 1.0 * xr, zi[m] + 1.0 * xi
        if nuc is not None and n >= 1:
            con.multiply(2.0 * fr, 2.0 * fi, 0)
        if hold > 0:
            mu.multiply(2.0 * fr, 2.0 * fi)
        Zr, Zi = zr[m], zi[m]
        xr, xi = (2.0 * (Zr * xr - Zi * xi) + xr * xr - xi * xi + ar,
                  2.0 * (Zr * xi + Zi * xr + xr * xi) + ai)
        m += 1
        n += 1
        pr, pi = 1.0 * xr, 1.0 * xi
        fr, fi = zr[m] + pr, zi[m] + pi
        f2 = fr * fr + fi * fi
        if f2 > r2:
            return "escaped", n
        if nuc is not None and n == nuc + 1:
            if con.small():
                return "interior", n
            nuc = None
        delta2 = pr * pr + pi * pi
        if f2 < delta2 or m == last:
            xr, xi, m = fr, fi, 0
        if not (xr * xr + xi * xi > 1e-12 * 1e-12 * f2):
            if hold > 0:
                if n >= sn + hold:
                    hold, predicted, pending = 0, False, True
            elif n >= chk:
                pending = True
            continue
        dist = abs(fr - sr) + abs(fi - si)
        size = abs(sr) + abs(si)
        if dist < 1e-3 * size:
            period = n - sn
            was = predicted
            below = mu.below_one()
            if hold == period and below:
                mr, mi = mu.value()
                gr, gi = fr - sr, fi - si
                er, ei = mr - 1.0, mi
                den = er * er + ei * ei
                p_r = sr - (gr * er + gi * ei) / den
                p_i = si - (gi * er - gr * ei) / den
                tol = min(0.1 * math.sqrt(den), 1e-2) * qstep + 16.0 * U * size
                if predicted and abs(p_r - qr) + abs(p_i - qi) <= tol:
                    return "interior", n
                qr, qi, qstep = p_r, p_i, (abs(gr) + abs(gi)) / math.sqrt(den)
                predicted = True
            else:
                predicted = False
            if log is not None:
                a = max(abs(mu.re), abs(mu.im))
                log2_mu = (mu.exp + math.log2(math.hypot(mu.re, mu.im))) if a > 0 else -math.inf
                log.append(dict(n=n, period=period, repeat=hold == period, below=bool(below),
                                predicted=was, log2_mu=log2_mu, z=(fr, fi)))
            hold = period
            sr, si, sn, pending = fr, fi, n, False
            mu = Multiplier()
            continue
        if hold > 0:
            if n < sn + hold:
                continue
            hold, predicted, pending = 0, False, True
        if pending or n >= chk:
            chk = max(chk, 1 << n.bit_length())  # (n + 1).next_power_of_two()
            sr, si, sn, pending = fr, fi, n, False
    return "unresolved", 0
