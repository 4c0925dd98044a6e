// The FP32 "floatexp" delta form for gfx950: the perturbation loop with FP32 mantissas
// and carried int32 binary exponents (delta = x 2^e, dz/dc = d 2^de, dc = a 2^oe), FP32
// orbit and BLA coefficients, FP64 only where precision decides a branch (the BLA
// validity test against the FP64 radii) and in the host's final smooth-iteration log.
//
// It follows the same iteration as the FP64 form (crates/fd-kernel/src/sample.rs:237-575:
// BLA lookup with the step budget, block apply, plain step, escape, nucleus contraction
// heuristic, rebase) with three differences, all stated where they occur: FP32 rounding
// (so results are close to, not equal to, the reference's), no Brent/Newton periodicity
// confirmation (samples the reference proves Interior that way end Unresolved here), and
// no scaled-tier close-return contraction test.  Work scheduling, outputs and statistics
// are those of the FP64 kernel.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "bla_find.hpp"
#include "device_math.hpp"
#include "kernels.hpp"
#include "lane_stats.hpp"

namespace fdg {
namespace {

using namespace fdg::dm;

#define FDG_GLOBAL __attribute__((address_space(1)))
typedef float v2f __attribute__((ext_vector_type(2)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef int32_t v2i __attribute__((ext_vector_type(2)));

constexpr int kBlock32 = 256;
constexpr int kWalk32 = 4;

// x 2^k in FP32 (exact unless it under/overflows)
__device__ inline float scale2(float x, int32_t k) {
    k = k > 300 ? 300 : (k < -300 ? -300 : k);
    return ldexpf(x, k);
}

// Keep a mantissa pair near 1: (re, im) 2^e unchanged in value, max(|re|, |im|) in [1, 2).
__device__ inline void renorm(float& re, float& im, int32_t& e) {
    float a = fmaxf(fabsf(re), fabsf(im));
    if (!(a > 0.0f) || !isfinite(a)) return;
    int32_t k = __builtin_amdgcn_frexp_expf(a) - 1;
    re = ldexpf(re, -k);
    im = ldexpf(im, -k);
    e += k;
}

// interior.rs:6-11 (f64 tier closed form, evaluated in FP64 as the reference does)
__device__ inline bool in_main_components32(double cr, double ci) {
    double y2 = ci * ci;
    double q = (cr - 0.25) * (cr - 0.25) + y2;
    return q * (q + (cr - 0.25)) <= 0.25 * y2 || (cr + 1.0) * (cr + 1.0) + y2 <= 0.0625;
}

__device__ inline uint32_t lanes_below32(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// A complex floatexp value: (re, im) 2^e with max(|re|, |im|) in [1, 2), or zero.
struct Fx {
    float re, im;
    int32_t e;
};

// z = x 2^xe + y 2^ye, exact exponent bookkeeping, FP32 rounding of the sum.
__device__ inline Fx fx_add(float xr, float xi, int32_t xe, float yr, float yi, int32_t ye) {
    bool xz = xr == 0.0f && xi == 0.0f, yz = yr == 0.0f && yi == 0.0f;
    int32_t e = xz ? ye : (yz ? xe : (xe > ye ? xe : ye));
    Fx z;
    z.re = scale2(xr, xe - e) + scale2(yr, ye - e);
    z.im = scale2(xi, xe - e) + scale2(yi, ye - e);
    z.e = e;
    renorm(z.re, z.im, z.e);
    return z;
}

// |x 2^xe|^2 < |y 2^ye|^2 without leaving FP32's range
__device__ inline bool fx_less2(float xr, float xi, int32_t xe, float yr, float yi, int32_t ye) {
    float x2 = xr * xr + xi * xi, y2 = yr * yr + yi * yi;
    if (y2 == 0.0f) return false;
    if (x2 == 0.0f) return true;
    int32_t d = 2 * (xe - ye);  // mantissas are normalised: x2, y2 in [1, 8)
    if (d > 8) return false;
    if (d < -8) return true;
    return scale2(x2, d) < y2;
}

// log2 of a bound that is zero (an exact quantity)
constexpr float kNoError = -1.0e30f;
// log2 of the relative rounding charged to one complex FP32 operation: 16 units of 2^-24
constexpr float kRound = -20.0f;

struct Lane32 {
    float xr, xi, dr, di, ar, ai;
    float Zr, Zi, Z1r, Z1i;     // Z_m and Z_{m+1} mantissas
    float lq, lq1;              // log2 of their error radii (Reference::error_radius)
    float le;                   // log2 of a bound on |delta - the exact delta|
    bool safe;                  // |z| + that bound stayed below the escape radius
    // Brent/Newton periodicity (ordinary tiers, sample.rs:524-558): the saved iterate
    // (sr, si) at sn, the next save point chk, the awaited return period hold, the
    // return multiplier mu = (mr, mi) 2^me since the save, the last cycle-point
    // prediction (qr, qi) and its step; flags kPending | kPredicted
    double sr, si, qr, qi, qstep;
    float mur, mui;
    int32_t mue;
    uint32_t sn, chk, hold, flags;
    float cr, ci;               // nucleus contraction product (mantissa)
    double R0;                  // level-0 r2 of the block starting at m (when any)
    int32_t e, de, ce, ze, z1e; // exponents: delta, dz/dc, product, Z_m, Z_{m+1}
    uint32_t m, n, nucleus, blocks, skipped, plain;
    bool con;                   // contraction product started
};

struct Hot32 {
    const FDG_GLOBAL v4f* orbit;    // Z_k as (re, im) 2^e: (re, im, bits of e, 0)
    const FDG_GLOBAL v4f* ab;       // A, B mantissas: (A.re, A.im) 2^ea, (B.re, B.im) 2^eb
    const FDG_GLOBAL v2i* abexp;    // (ea, eb): |A|, |B| exceed FP32's range on long blocks
    const FDG_GLOBAL double* r2;
    ConstR2 r2c;                    // the same radii, for scalar loads
    const FDG_GLOBAL double* alpha;  // the table's state envelopes (bla.rs:91-111)
    const FDG_GLOBAL double* beta;
    uint32_t last, max_iter, nlevels, steps, count0;
    uint64_t stride;
    float esc2, escR;
    int32_t oe;
    float Z1r, Z1i, Z0q, Z1q;
    int32_t Z1e;
    bool periodic;                 // ordinary tiers: Brent/Newton confirmation on
};

constexpr uint32_t kPending32 = 1, kPredicted32 = 2;
// sample.rs:524-558's tolerance term 16 u |s| with the FP32 unit, as the predictions are
// made from FP32 iterates
constexpr double kU32 = 1.0 / 16777216.0;

// Error bounds are carried as log2 of the bound, in FP32; each sum of bounds uses an upper
// bound on log2(2^a + 2^b), so the slack does not compound over long samples.  The hardware
// log2 / exp2 are accurate to about an ulp; every use is inflated by 2^-17 relative plus
// 2^-20 absolute (in log2 units) to stay an upper bound.
__device__ inline float lse(float a, float b) {
    const float m = fmaxf(a, b), d = fabsf(a - b);
    const float t = __builtin_amdgcn_logf(1.0f + __builtin_amdgcn_exp2f(-d));
    return m + t * (1.0f + 0x1p-17f) + 0x1p-20f;
}
// log2 |(re, im) 2^e| (kNoError for 0), inflated as above
__device__ inline float lmag(float re, float im, int32_t e) {
    const float a = re * re + im * im;
    if (!(a > 0.0f)) return kNoError;
    return (float)e + 0.5f * __builtin_amdgcn_logf(a) * (1.0f + 0x1p-17f) + 0x1p-18f;
}
// an upper bound on log2 v for v >= 0 (kNoError for 0)
__device__ inline float lg64(double v) {
    return v > 0.0 ? (float)__builtin_amdgcn_frexp_exp(v) : kNoError;
}

// A sample's start: delta = 0, D = 0, dc = a 2^aexp with a rounded to FP32 and its
// exponent kept apart from the frame's offset exponent, so any depth fits.
__device__ inline void start32(Lane32& s, const Hot32& h, uint32_t nucleus, double ar, double ai,
                               int32_t& aexp) {
    float fr = (float)ar, fi = (float)ai;
    int32_t ae = 0;
    renorm(fr, fi, ae);
    aexp = h.oe + ae;
    s.ar = fr;
    s.ai = fi;
    s.xr = s.xi = 0.0f;
    s.e = aexp;
    s.dr = s.di = 0.0f;
    s.de = 0;
    s.Zr = s.Zi = 0.0f;
    s.ze = 0;
    s.Z1r = h.Z1r;
    s.Z1i = h.Z1i;
    s.z1e = h.Z1e;
    s.lq = h.Z0q;
    s.lq1 = h.Z1q;
    s.le = kNoError;
    s.safe = true;
    s.sr = s.si = s.qr = s.qi = s.qstep = 0.0;
    s.mur = 1.0f;
    s.mui = 0.0f;
    s.mue = 0;
    s.sn = 8;    // the ordinary tiers' first save point (sample.rs:216-230)
    s.chk = 16;
    s.hold = 0;
    s.flags = 0;
    s.cr = 1.0f;
    s.ci = 0.0f;
    s.ce = 0;
    s.con = false;
    s.R0 = 0.0;
    s.m = s.n = 0;
    s.nucleus = nucleus;
    s.blocks = s.skipped = s.plain = 0;
}

// One block or plain step plus the checks after it.  Returns -1 running, or the outcome
// kind (0 escaped, z in ofr/ofi; 1 interior; 2 unresolved).
template <bool D, bool L, bool CERT>
__device__ inline int advance32(Lane32& s, const Hot32& h, int32_t aexp, float& ofr, float& ofi) {
    if (s.n >= h.max_iter) return 2;
    v4f pz2 = h.orbit[s.m + 2];
    double pr0 = 0.0;
    if (L && (s.m & 1u) == 0 && (s.m >> 1) < h.count0) pr0 = h.r2[s.m >> 1];
    float Znr = s.Z1r, Zni = s.Z1i, Zxr = pz2.x, Zxi = pz2.y;
    float Znq = s.lq1, Zxq = pz2.w;
    int32_t Zne = s.z1e, Zxe = __float_as_int(pz2.z);
    // log2 of this step's input magnitudes (certifying pass only): delta, dc, Z_m
    const float ld = CERT ? lmag(s.xr, s.xi, s.e) : 0.0f;
    const float ldc = CERT ? lmag(s.ar, s.ai, aexp) : 0.0f;
    const float lZ = CERT ? lmag(s.Zr, s.Zi, s.ze) : 0.0f;
    double r0next = pr0;
    int lvl = -1;
    if (L) {
        uint32_t k = s.m - 1;
        if ((k & 1u) == 0 && (s.nucleus == 0 || s.m != 0)) {
            double xr = (double)s.xr, xi = (double)s.xi;
            double d2 = (xr * xr + xi * xi) * exp2i(2 * (int64_t)s.e);
            uint32_t b = h.max_iter;
            // with periodicity, a block ends at the next save point or awaited return at
            // the latest (sample.rs:238-262), so saves land on the plain kernel's iterates
            if (h.periodic) {
                uint32_t pb = s.hold > 0 ? s.sn + s.hold
                                         : ((s.flags & kPending32) ? 0xffffffffu : s.chk);
                b = b < pb ? b : pb;
            }
            uint32_t nb = s.nucleus ? (s.n == 0 ? 1u : s.nucleus + 1) : 0xffffffffu;
            b = b < nb ? b : nb;
            b = b > s.n ? b - s.n : 0;
            if (d2 < s.R0) {
                // lanes in lockstep (one start and budget in every searching lane) take the
                // cap and the radii through the scalar unit
                const uint32_t k0 = __builtin_amdgcn_readfirstlane(k),
                               b0 = __builtin_amdgcn_readfirstlane(b);
                if (__ballot(k != k0 || b != b0) == 0) {
                    int cap = find_cap(k0, h.nlevels, b0);
                    if (cap >= 0)
                        lvl = find_level_down_uniform(h.r2c, 0u, (uint32_t)h.stride, k0, cap, d2);
                } else {
                    int cap = find_cap(k, h.nlevels, b);
                    if (cap >= 0) lvl = find_level_down<false>(h.r2, 0u, (uint32_t)h.stride, k, cap, d2);
                }
            }
        }
    }
    float mr, mi;  // this step's multiplier (for the contraction product), times 2^mexp
    int32_t mexp = 0;
    if (lvl >= 0) {
        uint32_t at = bla_block_index(0u, (uint32_t)h.stride, s.m - 1, lvl);
        v4f ab = h.ab[at];
        v2i ex = h.abexp[at];
        uint32_t blen = 2u << lvl;
        v4f zn = h.orbit[s.m + blen], zx = h.orbit[s.m + blen + 1];
        Znr = zn.x;
        Zni = zn.y;
        Zne = __float_as_int(zn.z);
        Znq = zn.w;
        Zxr = zx.x;
        Zxi = zx.y;
        Zxe = __float_as_int(zx.z);
        Zxq = zx.w;
        uint32_t i0 = (s.m + blen - 1) >> 1;
        r0next = i0 < h.count0 ? h.r2[i0] : 0.0;
        // error bound through the block (the reference route's bla_advance,
        // crates/fd-cli/shaders/perturb_bla.wgsl:30-35): (|A| + alpha) e + alpha |delta|
        // + beta |dc| + (|B| + beta) rc + rounding(|A| |delta| + |B| |dc|), rc the FP32
        // rounding of dc; each sum of k terms bounded by k times its largest, in log2
        if (CERT) {
            const float lA = lmag(ab.x, ab.y, ex.x), lB = lmag(ab.z, ab.w, ex.y);
            const float la = lg64(h.alpha[at]), lb = lg64(h.beta[at]);
            const float lrc = ldc - 23.0f;
            float t = lse(lse(lA, la) + s.le, la + ld);
            t = lse(t, lse(lb + ldc, lse(lB, lb) + lrc));
            s.le = lse(t, kRound + lse(lA + ld, lB + ldc));
        }
        // A delta + B dc with A = a 2^ea, B = b 2^eb, each term scaled into the units of
        // the larger (one ldexp each, so neither overflows nor needlessly underflows).
        mr = ab.x;
        mi = ab.y;
        mexp = ex.x;
        if (D) {
            // D' = A D + B
            int32_t E = max(s.de + ex.x, ex.y);
            float fa = scale2(1.0f, s.de + ex.x - E), fb = scale2(1.0f, ex.y - E);
            float dr = s.dr * fa, di = s.di * fa;
            float t = ab.x * dr - ab.y * di + ab.z * fb;
            s.di = ab.x * di + ab.y * dr + ab.w * fb;
            s.dr = t;
            s.de = E;
            renorm(s.dr, s.di, s.de);
        }
        int32_t E = max(s.e + ex.x, ex.y + aexp);
        float fa = scale2(1.0f, s.e + ex.x - E), fb = scale2(1.0f, ex.y + aexp - E);
        float xr = s.xr * fa, xi = s.xi * fa, bdr = s.ar * fb, bdi = s.ai * fb;
        float t = ab.x * xr - ab.y * xi + ab.z * bdr - ab.w * bdi;
        s.xi = ab.x * xi + ab.y * xr + ab.z * bdi + ab.w * bdr;
        s.xr = t;
        s.e = E;
        s.m += blen;
        s.n += blen;
        s.blocks += 1;
        s.skipped += blen;
    } else {
        // z = Z_m + delta, for the derivative and the contraction multiplier
        Fx z = fx_add(s.Zr, s.Zi, s.ze, s.xr, s.xi, s.e);
        mr = 2.0f * z.re;
        mi = 2.0f * z.im;
        mexp = z.e;
        if (D) {
            // D' = 2 z D + 1, in the units of the larger term
            int32_t E = max(s.de + z.e + 1, 0);
            float fz = scale2(1.0f, s.de + z.e - E), f1 = scale2(1.0f, -E);
            float dr = s.dr * fz, di = s.di * fz;
            float t = 2.0f * (z.re * dr - z.im * di) + f1;
            s.di = 2.0f * (z.re * di + z.im * dr);
            s.dr = t;
            s.de = E;
            renorm(s.dr, s.di, s.de);
        }
        // error bound through the step (the reference route's frame loop,
        // crates/fd-cli/shaders/perturb_frame.wgsl:33-36): e' = e (2w + e) + rho with
        // w = |z| + q + rounding(|Z| + |delta|), rho = rounding(2|Z||delta| + |delta|^2 +
        // |dc|) + 2 q |delta| + rc
        if (CERT) {
            const float lz = lmag(z.re, z.im, z.e);
            const float lw = lse(lse(lz, s.lq), kRound + lse(lZ, ld));
            const float lrho = lse(ld + lse(kRound + lse(1.0f + lZ, ld), 1.0f + s.lq),
                                   lse(kRound, -23.0f) + ldc);
            s.le = lse(s.le + lse(1.0f + lw, s.le), lrho);
        }
        // delta' = 2 Z delta + delta^2 + dc, in the units of the largest term
        bool zz = s.Zr == 0.0f && s.Zi == 0.0f;
        int32_t E = max(max(2 * s.e + 1, aexp), zz ? INT32_MIN / 2 : s.e + s.ze + 1);
        // (Z = 0 is left out of E, so its scale factor must not be formed: 0 inf = NaN)
        float fz = zz ? 0.0f : scale2(1.0f, s.e + s.ze - E), fq = scale2(1.0f, 2 * s.e - E),
              fdc = scale2(1.0f, aexp - E);
        float xr = s.xr, xi = s.xi;
        float zr = s.Zr * fz, zi = s.Zi * fz;
        s.xr = 2.0f * (zr * xr - zi * xi) + fq * (xr * xr - xi * xi) + s.ar * fdc;
        s.xi = 2.0f * (zr * xi + zi * xr) + fq * (2.0f * xr * xi) + s.ai * fdc;
        s.e = E;
        s.m += 1;
        s.n += 1;
        s.plain += 1;
    }
    renorm(s.xr, s.xi, s.e);
    if (s.nucleus && s.n >= 1 + (lvl >= 0 ? (2u << lvl) : 1u)) {
        // product of the step multipliers since n = 1 (sample.rs:657-698), floatexp
        if (!s.con) {
            s.cr = 1.0f;
            s.ci = 0.0f;
            s.ce = 0;
            s.con = true;
        }
        float nr = s.cr * mr - s.ci * mi, ni = s.cr * mi + s.ci * mr;
        s.cr = nr;
        s.ci = ni;
        s.ce += mexp;
        renorm(s.cr, s.ci, s.ce);
    }
    if (h.periodic && s.hold > 0) {
        // the return multiplier (sample.rs:597-652), mantissa kept near 1
        float nr = s.mur * mr - s.mui * mi, ni = s.mur * mi + s.mui * mr;
        s.mur = nr;
        s.mui = ni;
        s.mue += mexp;
        renorm(s.mur, s.mui, s.mue);
    }
    // the landing point z = Z_{m'} + delta'
    Fx zl = fx_add(Znr, Zni, Zne, s.xr, s.xi, s.e);
    float fr = scale2(zl.re, zl.e), fi = scale2(zl.im, zl.e);
    // certified below the escape radius: |z| + e < R (|z| bounded by its L1 norm)
    if (CERT) s.safe = s.safe && fabsf(fr) + fabsf(fi) + exp2f(s.le) < h.escR;
    // Once its bound reaches the escape radius the sample is recomputed in FP64 whatever
    // it does next (escaped or not), so the FP32 pass stops it here (3: abandoned).
    if (CERT && !s.safe) return 3;
    if (fr * fr + fi * fi > h.esc2) {
        ofr = fr;
        ofi = fi;
        return 0;
    }
    if (s.nucleus && s.n == s.nucleus + 1) {
        // |product| < 1e-6, with the exponent kept apart
        if (s.con && ((s.cr == 0.0f && s.ci == 0.0f) || s.ce < -21 ||
                      (s.ce <= 21 && (double)(s.cr * s.cr + s.ci * s.ci) *
                                             exp2i(2 * (int64_t)s.ce) < 1e-12)))
            return 1;
        s.nucleus = 0;
    }
    bool rebase = fx_less2(zl.re, zl.im, zl.e, s.xr, s.xi, s.e) || s.m == h.last;
    if (rebase) {
        // the new delta is z = Z + delta: e + q + rounding(|Z| + |delta|)
        if (CERT)
            s.le = lse(lse(s.le, Znq),
                       kRound + lse(lmag(Znr, Zni, Zne), lmag(s.xr, s.xi, s.e)));
        s.xr = zl.re;
        s.xi = zl.im;
        s.e = zl.e;
        s.m = 0;
        s.Zr = s.Zi = 0.0f;
        s.ze = 0;
        s.Z1r = h.Z1r;
        s.Z1i = h.Z1i;
        s.z1e = h.Z1e;
        s.lq = h.Z0q;
        s.lq1 = h.Z1q;
    } else {
        s.Zr = Znr;
        s.Zi = Zni;
        s.ze = Zne;
        s.Z1r = Zxr;
        s.Z1i = Zxi;
        s.z1e = Zxe;
        s.lq = Znq;
        s.lq1 = Zxq;
    }
    s.R0 = r0next;
    if (h.periodic) {
        // Brent saves and Newton's cycle-point prediction (sample.rs:451-575), on the
        // physical z (FP64 here: ordinary-tier iterates are inside its range)
        const double zr = ldexp((double)zl.re, zl.e), zi = ldexp((double)zl.im, zl.e);
        const double f2 = zr * zr + zi * zi;
        const double xr = ldexp((double)s.xr, s.e), xi = ldexp((double)s.xi, s.e);
        const bool res = xr * xr + xi * xi > 1e-24 * f2;  // (after a rebase delta = z)
        const double dist = fabs(zr - s.sr) + fabs(zi - s.si);
        const double size = fabs(s.sr) + fabs(s.si);
        if (res && dist < 1e-3 * size) {
            const uint32_t period = s.n - s.sn;
            // |mu| < 1 (sample.rs scaled_below)
            const double m2 = (double)s.mur * s.mur + (double)s.mui * s.mui;
            const bool below = isfinite(m2) &&
                               (s.mue < -64 || (s.mue <= 64 && m2 * exp2i(2 * s.mue) < 1.0));
            if (s.hold == period && below) {
                const double sc = exp2i(s.mue);
                const double mr2 = (double)s.mur * sc, mi2 = (double)s.mui * sc;
                const double gr = zr - s.sr, gi = zi - s.si;
                const double er = mr2 - 1.0, ei = mi2;
                const double den = er * er + ei * ei;
                const double pr2 = s.sr - (gr * er + gi * ei) / den;
                const double pi2 = s.si - (gi * er - gr * ei) / den;
                const double tol = fmin(0.1 * sqrt(den), 1e-2) * s.qstep + 16.0 * kU32 * size;
                if ((s.flags & kPredicted32) && fabs(pr2 - s.qr) + fabs(pi2 - s.qi) <= tol)
                    return 1;  // Interior: two consecutive predictions agree
                s.qr = pr2;
                s.qi = pi2;
                s.qstep = (fabs(gr) + fabs(gi)) / sqrt(den);
                s.flags |= kPredicted32;
            } else {
                s.flags &= ~kPredicted32;
            }
            s.hold = period;
            s.sr = zr;
            s.si = zi;
            s.sn = s.n;
            s.flags &= ~kPending32;
            s.mur = 1.0f;
            s.mui = 0.0f;
            s.mue = 0;
            return -1;
        }
        const uint32_t h0 = s.hold;
        const bool expired = h0 > 0 && s.n >= s.sn + h0;
        if (expired) {
            s.hold = 0;
            s.flags = (s.flags & ~kPredicted32) | kPending32;
        }
        if (!res && h0 == 0 && s.n >= s.chk) s.flags |= kPending32;  // deferred save
        if (res && (h0 == 0 || expired) && ((s.flags & kPending32) || s.n >= s.chk)) {
            const uint32_t np = 1u << (32 - __builtin_clz(s.n));  // next power of two > n
            s.chk = s.chk > np ? s.chk : np;
            s.sr = zr;
            s.si = zi;
            s.sn = s.n;
            s.flags &= ~kPending32;
        }
    }
    return -1;
}

// CERT: the mixed form's pass, which also carries the error bound and certifies samples.
template <bool D, bool L, bool CERT>
__global__ void __launch_bounds__(kBlock32, 4) perturb_fe32(const FrameDev* __restrict__ fp) {
    __shared__ unsigned long long stat[ST_COUNT];
    if (threadIdx.x < ST_COUNT) stat[threadIdx.x] = 0;
    __shared__ unsigned long long lane_slots[kLaneStats * kBlock32];
    LaneStats<kBlock32> ls{lane_slots};
    ls.clear();
    __syncthreads();
    const uint32_t lane = __lane_id();
    Hot32 h;
    h.orbit = (const FDG_GLOBAL v4f*)(fp->orbit32);
    h.ab = (const FDG_GLOBAL v4f*)(fp->ab32);
    h.abexp = (const FDG_GLOBAL v2i*)(fp->abexp32);
    h.r2 = (const FDG_GLOBAL double*)(fp->bla.r2);
    h.r2c = (ConstR2)fp->bla.r2;
    h.alpha = (const FDG_GLOBAL double*)(fp->bla_alpha);
    h.beta = (const FDG_GLOBAL double*)(fp->bla_beta);
    h.last = (uint32_t)fp->len - 1;
    h.max_iter = (uint32_t)fp->max_iter;
    h.nlevels = fp->bla.nlevels;
    h.steps = (uint32_t)fp->bla.steps;
    h.stride = fp->bla.stride;
    h.count0 = (uint32_t)(fp->bla.steps >> 1);
    h.esc2 = (float)fp->r2;
    h.escR = (float)sqrt(fp->r2);
    h.periodic = fp->tier != 2;
    h.oe = (int32_t)fp->h_e;
    {
        v4f z1 = h.orbit[1];
        h.Z1r = z1.x;
        h.Z1i = z1.y;
        h.Z1e = __float_as_int(z1.z);
        h.Z1q = z1.w;
        h.Z0q = h.orbit[0].w;
    }
    FDG_GLOBAL uint8_t* o_kind = (FDG_GLOBAL uint8_t*)fp->kind;
    FDG_GLOBAL uint64_t* o_n = (FDG_GLOBAL uint64_t*)fp->n;
    FDG_GLOBAL double* o_zr = (FDG_GLOBAL double*)fp->zr;
    FDG_GLOBAL double* o_zi = (FDG_GLOBAL double*)fp->zi;
    FDG_GLOBAL double* o_dr = (FDG_GLOBAL double*)fp->dr;
    FDG_GLOBAL double* o_di = (FDG_GLOBAL double*)fp->di;
    FDG_GLOBAL int64_t* o_dexp = (FDG_GLOBAL int64_t*)fp->dexp;
    FDG_GLOBAL uint32_t* o_ops = (FDG_GLOBAL uint32_t*)fp->ops;
    const uint32_t count = fp->count, nx = fp->nx, nucleus = (uint32_t)fp->nucleus;
    uint32_t* next = fp->next;
    Lane32 s;
    int32_t aexp = 0;  // exponent of this sample's dc mantissa: dc = a 2^aexp
    bool active = false, drained = false;
    uint32_t idx = 0;
    unsigned long long witers = 0, liters = 0;
    // finished: samples this wave has started (closed-form interiors included; a masked-out
    // index or an exhausted claim starts none)
    uint32_t cb = 0, ce = 0, seen = 0, finished = 0;
    const uint32_t waves = gridDim.x * (kBlock32 / 64);
    // claim policy (the launch's tuning; 0 = the defaults)
    const uint32_t cdiv = fp->claim_div ? fp->claim_div : 4u;
    const uint32_t ccap = fp->claim_cap ? fp->claim_cap : 1024u;
    // (whole-wave refills by default: a wave's tile of samples stays in lockstep)
    const uint32_t refill = fp->refill ? fp->refill : 64u;
    for (;;) {
        uint64_t idle = __ballot(!active);
        if (idle && !drained) {
            uint32_t want = __popcll(idle);
            uint32_t avail = ce - cb;
            uint32_t nb = 0, ne = 0;
            if (avail < want) {
                uint32_t left = count > seen ? count - seen : 0;
                // Samples shorter than 64 loop iterations (fewer wave iterations so far
                // than samples this wave finished) make the claims themselves the
                // bottleneck: then the chunk is the work left per wave rather than a
                // quarter of it.
                uint32_t div = witers + 64 < finished ? 1u : cdiv;
                uint32_t size = max(min(left / (div * waves), ccap), want - avail);
                // Every claim is a whole number of 64-sample groups, so every claimed
                // range starts on a group boundary: no two waves write parts of one
                // 64-byte line of the 1-byte outcome kinds.
                size = (size + 63u) & ~63u;
                uint32_t leader = __ffsll((unsigned long long)idle) - 1;
                if (lane == leader) nb = atomicAdd(next, size);
                nb = __builtin_amdgcn_readlane(nb, leader);
                seen = nb + size;
                ne = nb < count ? min(nb + size, count) : nb;
            }
            uint32_t rank = lanes_below32(idle);
            uint32_t my = rank < avail ? cb + rank : nb + (rank - avail);
            bool have = rank < avail || my < ne;
            if (avail >= want) {
                cb += want;
            } else {
                cb = min(nb + (want - avail), ne);
                ce = ne;
                if (nb >= count) drained = true;
            }
            const uint32_t one = sample_of(my, nx, fp->ny);
            bool took = !active && have && (!fp->mask || fp->mask[one]);
            if (took) {
                idx = one;
                uint32_t i = idx % nx, j = idx / nx;
                double x = (double)i + 0.5 - fp->half_x;
                double y = (double)j + 0.5 - fp->half_y;
                double ux = fp->cos_r * x + fp->sin_r * y;
                double uy = fp->sin_r * x - fp->cos_r * y;
                o_ops[idx] = 0;
                if (fp->tier == 0 &&
                    in_main_components32(fp->c_re + ux * fp->h, fp->c_im + uy * fp->h)) {
                    o_kind[idx] = 1;
                    o_n[idx] = 0;
                    o_zr[idx] = o_zi[idx] = o_dr[idx] = o_di[idx] = 0.0;
                    o_dexp[idx] = 0;
                    if (L) ls.add(ST_CLOSED, 1ull);
                } else {
                    start32(s, h, nucleus, ux * fp->h_m, uy * fp->h_m, aexp);
                    active = true;
                }
            }
            finished += __popcll(__ballot(took));
        }
        // Fill the wave before stepping: a masked-out index or a closed-form interior
        // leaves its lane idle, and the steps below refill only at `refill` idle lanes.
        if (__ballot(!active) && !drained) continue;
        if (!__ballot(active)) break;  // (drained)
        do {
            witers += 1;
            liters += __popcll(__ballot(active));
            if (active) {
                float fr = 0.0f, fi = 0.0f;
                int k = advance32<D, L, CERT>(s, h, aexp, fr, fi);
                if (k >= 0) {
                    bool esc = k == 0;
                    // kept: a non-escaped sample whose bound never reached the escape radius
                    bool keep = !esc && (!CERT || s.safe);
                    if (CERT) fp->cert[idx] = keep ? 1 : 0;
                    if (CERT && !esc && !keep) atomicAdd(&stat[ST_UNCERT], 1ull);
                    // (an abandoned sample's fields are placeholders the FP64 pass
                    // overwrites)
                    o_kind[idx] = (uint8_t)(k == 3 ? 2 : k);
                    o_n[idx] = s.n;
                    o_ops[idx] = s.plain + s.blocks;
                    o_zr[idx] = esc ? (double)fr : 0.0;
                    o_zi[idx] = esc ? (double)fi : 0.0;
                    o_dr[idx] = esc ? (double)s.dr : 0.0;
                    o_di[idx] = esc ? (double)s.di : 0.0;
                    // dz/dc = d 2^de per unit of c; the outcome's D is per unit of c too
                    o_dexp[idx] = esc ? (int64_t)s.de : 0;
                    unsigned long long spent = k == 2 ? h.max_iter : s.n;
                    if (!(!keep && fp->defer_escaped)) {
                        ls.add(ST_EQUIVALENT, spent);
                        ls.add(ST_ITERATIONS, L ? spent - s.skipped + s.blocks : spent);
                        if (L) {
                            ls.add(ST_BLOCKS, (unsigned long long)s.blocks);
                            ls.add(ST_SKIPPED, (unsigned long long)s.skipped);
                            if (s.blocks == 0) ls.add(ST_FALLBACK, 1ull);
                        }
                        ls.add(ST_PLAIN, (unsigned long long)s.plain);
                        ls.max(ST_MAXOPS, (unsigned long long)(s.plain + s.blocks));
                    }
                    active = false;
                }
            }
        } while (__popcll(__ballot(!active)) < refill || (drained && __ballot(active)));
    }
    ls.fold(stat);
    if (lane == 0) {
        atomicAdd(&stat[ST_WAVE_ITERS], witers);
        atomicAdd(&stat[ST_LANE_ITERS], liters);
    }
    __syncthreads();
    if (threadIdx.x < ST_COUNT) {
        unsigned long long v = stat[threadIdx.x];
        if (v && threadIdx.x == ST_MAXOPS)
            atomicMax(&fp->stats[ST_MAXOPS], v);
        else if (v)
            atomicAdd(&fp->stats[threadIdx.x], v);
    }
}



// One complex coefficient as an FP32 mantissa pair and a shared exponent.
__device__ inline void split32(double re, double im, float& mr, float& mi, int32_t& e) {
    double a = fmax(fabs(re), fabs(im));
    e = 0;
    if (a > 0.0 && isfinite(a)) e = __builtin_amdgcn_frexp_exp(a) - 1;
    mr = (float)ldexp(re, -e);
    mi = (float)ldexp(im, -e);
}

// The orbit as FP32 mantissa pairs with their exponents: (re, im, bits of e, 0) per point
// (deep returns take |Z| far below FP32's range).
__global__ void orbit_to_f32(const double* __restrict__ orbit, const double* __restrict__ q,
                             float* __restrict__ o32, uint64_t points) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= points) return;
    float r, im;
    int32_t e;
    split32(orbit[2 * i], orbit[2 * i + 1], r, im, e);
    o32[4 * i] = r;
    o32[4 * i + 1] = im;
    o32[4 * i + 2] = __int_as_float(e);
    o32[4 * i + 3] = q ? lg64(q[i]) : kNoError;  // log2 of the point's error radius
}

// Table coefficients (A, B) per block as FP32 mantissas with their exponents.
__global__ void table_to_f32(const double* __restrict__ ab, float* __restrict__ ab32,
                             int32_t* __restrict__ abexp, uint64_t blocks) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= blocks) return;
    float ar, ai, br, bi;
    int32_t ea, eb;
    split32(ab[4 * i], ab[4 * i + 1], ar, ai, ea);
    split32(ab[4 * i + 2], ab[4 * i + 3], br, bi, eb);
    ab32[4 * i] = ar;
    ab32[4 * i + 1] = ai;
    ab32[4 * i + 2] = br;
    ab32[4 * i + 3] = bi;
    abexp[2 * i] = ea;
    abexp[2 * i + 1] = eb;
}

}  // namespace

// The mixed form's second pass covers the samples the FP32 pass found escaped: their
// indices, compacted with one atomic per wave (the order within the list is immaterial,
// every sample being independent).
__global__ void escaped_list(const uint8_t* __restrict__ kind, const uint8_t* __restrict__ cert,
                             uint32_t* __restrict__ list, uint32_t* __restrict__ count,
                             uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    // (a sample outside a mask keeps kind 0xff: never listed)
    bool esc = i < n && kind[i] != 0xff && (cert ? cert[i] == 0 : kind[i] == 0);
    uint64_t b = __ballot(esc);
    if (!b) return;
    uint32_t leader = __ffsll((unsigned long long)b) - 1, base = 0;
    if (__lane_id() == leader) base = atomicAdd(count, (uint32_t)__popcll(b));
    base = __builtin_amdgcn_readlane(base, leader);
    if (esc) list[base + lanes_below32(b)] = i;
}

hipError_t launch_escaped_list(const uint8_t* kind, const uint8_t* cert, uint32_t* list,
                               uint32_t* count, uint32_t n, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(escaped_list, dim3((n + 255) / 256), dim3(256), 0, stream, kind, cert,
                       list, count, n);
    return hipGetLastError();
}

hipError_t launch_orbit_to_f32(const double* orbit, const double* q, float* o32, uint64_t points,
                               hipStream_t stream) {
    if (points == 0) return hipSuccess;
    hipLaunchKernelGGL(orbit_to_f32, dim3((unsigned)((points + 255) / 256)), dim3(256), 0, stream,
                       orbit, q, o32, points);
    return hipGetLastError();
}

hipError_t launch_table_to_f32(const double* ab, float* ab32, int32_t* abexp, uint64_t blocks,
                               hipStream_t stream) {
    if (blocks == 0) return hipSuccess;
    hipLaunchKernelGGL(table_to_f32, dim3((unsigned)((blocks + 255) / 256)), dim3(256), 0, stream,
                       ab, ab32, abexp, blocks);
    return hipGetLastError();
}

hipError_t launch_perturb_fe32(const FrameDev* f, bool deriv, bool bla, int grid,
                               hipStream_t stream, bool cert) {
#define FDG_LAUNCH32(D, L, C) \
    hipLaunchKernelGGL((perturb_fe32<D, L, C>), dim3(grid), dim3(kBlock32), 0, stream, f)
    switch ((cert ? 4 : 0) | (deriv ? 2 : 0) | (bla ? 1 : 0)) {
        case 0: FDG_LAUNCH32(false, false, false); break;
        case 1: FDG_LAUNCH32(false, true, false); break;
        case 2: FDG_LAUNCH32(true, false, false); break;
        case 3: FDG_LAUNCH32(true, true, false); break;
        case 4: FDG_LAUNCH32(false, false, true); break;
        case 5: FDG_LAUNCH32(false, true, true); break;
        case 6: FDG_LAUNCH32(true, false, true); break;
        default: FDG_LAUNCH32(true, true, true); break;
    }
#undef FDG_LAUNCH32
    return hipGetLastError();
}

}  // namespace fdg
