// FP64 deep-zoom kernels for gfx950: the BLA table builder (a level-by-level tree
// reduction) and the persistent per-sample perturbation kernel.
//
// Built with -ffp-contract=off: every floating-point expression keeps the reference's
// operation order (crates/fd-kernel/src/{sample,number,bla,interior}.rs), and gfx950's
// IEEE add/mul/div/sqrt then give the CPU kernel's exact bits.  The "FP64" form is the
// ordinary tiers' Binary64 arithmetic and the scaled tier's Binary64Exponent (f64 mantissa
// times a carried power of two), selected at compile time.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "bla_find.hpp"
#include "device_math.hpp"
#include "kernels.hpp"
#include "lane_stats.hpp"

namespace fdg {
namespace {

using namespace fdg::dm;

// ------------------------------------------------------------------ BLA table build

struct Block {
    double ar, ai, br, bi, r, alpha, beta;
};

// Block::single (bla.rs:72-85): one step at Z_k.
__device__ inline Block block_single(double zr, double zi, double eps) {
    Block b = {0, 0, 0, 0, 0, 0, 0};
    double az = hypot_ref(zr, zi);
    double r = 2.0 * eps * az;
    if (!(r > 0.0 && az <= 2.0)) return b;
    b.ar = 2.0 * zr + 0.0;
    b.ai = 2.0 * zi + 0.0;
    b.br = 1.0;
    b.bi = 0.0;
    b.r = r;
    b.alpha = r;
    b.beta = 0.0;
    return b;
}

// v 2^e for e far below the f64 range (bla.rs:114-121).
__device__ inline double times_exp2(double v, int64_t e) {
    if (e >= -1022) return v * exp2i(e);
    return v * exp2i(-1022) * exp2i(e + 1022);
}

// Block::merge (bla.rs:91-111): x then y, for |dc| <= c 2^ce.
__device__ inline Block block_merge(const Block& x, const Block& y, double c, int64_t ce) {
    Block z = {0, 0, 0, 0, 0, 0, 0};
    if (x.r == 0.0 || y.r == 0.0) return z;
    double ax = hypot_ref(x.ar, x.ai), ay = hypot_ref(y.ar, y.ai);
    double ahat = ax + x.alpha;
    double bhat = hypot_ref(x.br, x.bi) + x.beta;
    double r = fmin(x.r, (y.r - times_exp2(bhat * c, ce)) / ahat);
    Block m;
    m.ar = y.ar * x.ar - y.ai * x.ai + 0.0 + 0.0;
    m.ai = y.ar * x.ai + y.ai * x.ar + 0.0 + 0.0;
    m.br = y.ar * x.br - y.ai * x.bi + y.br + 0.0;
    m.bi = y.ar * x.bi + y.ai * x.br + y.bi + 0.0;
    m.r = r;
    m.alpha = ay * x.alpha + y.alpha * ahat;
    m.beta = ay * x.beta + y.alpha * bhat + y.beta;
    if (r > 0.0 && isfinite(m.ar) && isfinite(m.ai) && isfinite(m.br) && isfinite(m.bi) &&
        isfinite(m.r) && isfinite(m.alpha) && isfinite(m.beta))
        return m;
    return z;
}

struct BlaBuild {
    double* r;        // validity radius
    double* r2;
    double* ab;       // 4 per block
    double* alpha;
    double* beta;
    uint64_t steps, stride;
    double eps, dc_max;
    int64_t dc_exp;
    uint32_t* valid;  // per level: 1 if any block is valid
};

__device__ inline void store_block(const BlaBuild& t, uint64_t at, const Block& b) {
    t.r[at] = b.r;
    t.r2[at] = b.r * b.r;
    reinterpret_cast<double4*>(t.ab)[at] = make_double4(b.ar, b.ai, b.br, b.bi);
    t.alpha[at] = b.alpha;
    t.beta[at] = b.beta;
}

__device__ inline Block load_block(const BlaBuild& t, uint64_t at) {
    double4 ab = reinterpret_cast<const double4*>(t.ab)[at];
    return Block{ab.x, ab.y, ab.z, ab.w, t.r[at], t.alpha[at], t.beta[at]};
}

__device__ inline void flag_valid(const BlaBuild& t, uint32_t level, bool v) {
    if (__ballot(v) && __lane_id() == 0) atomicOr(&t.valid[level], 1u);
}

// Level 0 (2-step blocks) straight from the orbit: singles at k = 1 + 2t and 2 + 2t.  Each
// level's slot holds stride >> (level + 1) entries; the ones past its last block are
// written as zero blocks, so a lookup that reads past the table's end finds no valid block.
__global__ void __launch_bounds__(256) bla_level0(const double* orbit, BlaBuild t) {
    uint64_t cnt = t.steps >> 1, slot = t.stride >> 1;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < slot;
         i += (uint64_t)gridDim.x * blockDim.x) {
        Block m = {0, 0, 0, 0, 0, 0, 0};
        if (i < cnt) {
            uint64_t k = 1 + 2 * i;
            Block x = block_single(orbit[2 * k], orbit[2 * k + 1], t.eps);
            Block y = block_single(orbit[2 * k + 2], orbit[2 * k + 3], t.eps);
            m = block_merge(x, y, t.dc_max, t.dc_exp);
        }
        store_block(t, bla_level_offset(t.stride, 0) + i, m);
        flag_valid(t, 0, m.r > 0.0);
    }
}

// Level `lv` >= 1: pairs of level lv-1 blocks (zero blocks in the padding, as level 0).
__global__ void __launch_bounds__(256) bla_level(BlaBuild t, uint32_t lv) {
    uint64_t cnt = t.steps >> (lv + 1), slot = t.stride >> (lv + 1);
    uint64_t src = bla_level_offset(t.stride, lv - 1), dst = bla_level_offset(t.stride, lv);
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < slot;
         i += (uint64_t)gridDim.x * blockDim.x) {
        Block m = {0, 0, 0, 0, 0, 0, 0};
        if (i < cnt)
            m = block_merge(load_block(t, src + 2 * i), load_block(t, src + 2 * i + 1), t.dc_max,
                            t.dc_exp);
        store_block(t, dst + i, m);
        flag_valid(t, lv, m.r > 0.0);
    }
}

// Ten levels per launch: a workgroup builds one span of 512 blocks of level lv0 (two per
// thread: singles merged straight from the orbit for level 0, pairs of the stored level
// lv0-1 blocks above it) and merges them up through LDS to the single level lv0+9 block
// over the span, storing every level on the way.  Each merge takes the same two blocks
// as the level-by-level build, so the table is the same bit for bit, from one launch per
// ten levels instead of one per level.  Blocks past a level's last one are stored as
// zero blocks (as in bla_level0 / bla_level), also where both halves exist.
constexpr int kSpanThreads = 256;
constexpr int kSpan = 2 * kSpanThreads;  // blocks of level lv0 per workgroup
constexpr int kSpanLevels = 10;          // levels lv0 .. lv0 + 9 per launch

__global__ void __launch_bounds__(kSpanThreads) bla_span(const double* orbit, BlaBuild t,
                                                          uint32_t lv0, uint32_t lv_end) {
    __shared__ Block lds[kSpan];
    const uint32_t tid = threadIdx.x;
    const uint64_t slot0 = t.stride >> (lv0 + 1), nspans = (slot0 + kSpan - 1) / kSpan;
    for (uint64_t span = blockIdx.x; span < nspans; span += gridDim.x) {
        // level lv0: blocks 2 tid and 2 tid + 1 of the span
        const uint64_t cnt0 = t.steps >> (lv0 + 1);
        bool any = false;
        for (int h = 0; h < 2; h++) {
            uint64_t i = span * kSpan + 2 * tid + h;
            Block m = {0, 0, 0, 0, 0, 0, 0};
            if (i < cnt0) {
                if (lv0 == 0) {
                    uint64_t k = 1 + 2 * i;
                    Block x = block_single(orbit[2 * k], orbit[2 * k + 1], t.eps);
                    Block y = block_single(orbit[2 * k + 2], orbit[2 * k + 3], t.eps);
                    m = block_merge(x, y, t.dc_max, t.dc_exp);
                } else {
                    uint64_t src = bla_level_offset(t.stride, lv0 - 1);
                    m = block_merge(load_block(t, src + 2 * i), load_block(t, src + 2 * i + 1),
                                    t.dc_max, t.dc_exp);
                }
            }
            if (i < slot0) store_block(t, bla_level_offset(t.stride, lv0) + i, m);
            any = any || m.r > 0.0;
            lds[2 * tid + h] = m;
        }
        flag_valid(t, lv0, any);
        // levels lv0 + 1 .. : pairs of the level below, through LDS
        for (uint32_t lv = lv0 + 1; lv < lv_end; lv++) {
            const uint32_t j = lv - lv0;            // 1 .. 9
            const uint32_t n = kSpan >> j;          // blocks of this level in the span
            const uint64_t cnt = t.steps >> (lv + 1), slot = t.stride >> (lv + 1);
            __syncthreads();
            Block m = {0, 0, 0, 0, 0, 0, 0};
            Block x = m, y = m;
            if (tid < n) {
                x = lds[2 * tid];
                y = lds[2 * tid + 1];
            }
            __syncthreads();
            uint64_t i = (span * kSpan >> j) + tid;
            if (tid < n) {
                if (i < cnt) m = block_merge(x, y, t.dc_max, t.dc_exp);
                if (i < slot) store_block(t, bla_level_offset(t.stride, lv) + i, m);
                lds[tid] = m;
            }
            flag_valid(t, lv, tid < n && m.r > 0.0);
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------------ perturbation

constexpr double kHi = 4294967296.0;
constexpr double kLo = 1.0 / 4294967296.0;
constexpr double kSquareOk = 1e-290;
constexpr double kTiny = 1.0 / 1e270;
constexpr double kU = 2.220446049250313e-16 / 2.0;
constexpr double kResolvable = 1e-12;

// interior.rs:6-11
__device__ inline bool in_main_components(double cr, double ci) {
    double y2 = ci * ci;
    double q = (cr - 0.25) * (cr - 0.25) + y2;
    return q * (q + (cr - 0.25)) <= 0.25 * y2 || (cr + 1.0) * (cr + 1.0) + y2 <= 0.0625;
}

// interior.rs:16-64 map + contracting: `period` steps from phase m0 with delta x0.
__device__ bool contracting(const double* orbit, uint64_t len, uint64_t m0, double x0r,
                            double x0i, double ar, double ai, uint64_t period) {
    uint64_t last = len - 1, m = m0;
    double xr = x0r, xi = x0i, dr = 1.0, di = 0.0;
    for (uint64_t s = 0; s < period; s++) {
        double Zr = orbit[2 * m], Zi = orbit[2 * m + 1];
        double fr = Zr + xr, fi = Zi + xi;
        double t = 2.0 * (fr * dr - fi * di);
        di = 2.0 * (fr * di + fi * dr);
        dr = t;
        t = 2.0 * (Zr * xr - Zi * xi) + xr * xr - xi * xi + ar;
        xi = 2.0 * (Zr * xi + Zi * xr + xr * xi) + ai;
        xr = t;
        m += 1;
        fr = orbit[2 * m] + xr;
        fi = orbit[2 * m + 1] + xi;
        double f2 = fr * fr + fi * fi;
        if (f2 > 4.0) return false;
        if (f2 < xr * xr + xi * xi || m == last) {
            xr = fr;
            xi = fi;
            m = 0;
        }
    }
    return dr * dr + di * di < 1.0;
}

// sample.rs:597-652 Multiplier and sample.rs:657-698 Contraction: a complex product with
// a separate binary exponent.  The exponent choice only picks an exact power-of-two split.
// The exponent e with a = m 2^e, 1 <= m < 2 (a finite, > 0).  Where the reference takes
// floor(log2(a)) only to pick a power-of-two rescaling, any exponent gives the same value
// (the scaling is exact), so the frexp exponent replaces the libm log2.
__device__ inline int32_t exponent_of(double a) { return __builtin_amdgcn_frexp_exp(a) - 1; }

__device__ inline void mu_multiply(double& re_, double& im_, int32_t& ex, double re, double im) {
    double nr = re_ * re - im_ * im, ni = re_ * im + im_ * re;
    re_ = nr;
    im_ = ni;
    double a = fmax(fabs(nr), fabs(ni));
    if (!(a > 1e-120 && a < 1e120) && a > 0.0 && isfinite(a)) {
        int32_t e = exponent_of(a);
        double sc = exp2i32(-e);
        re_ *= sc;
        im_ *= sc;
        ex += e;
    }
}

__device__ inline bool scaled_below(double re, double im, int32_t ex, double bound) {
    return isfinite(re) && isfinite(im) &&
           (ex < -64 || (ex <= 64 && (re * re + im * im) * exp2i32(2 * ex) < bound));
}

__device__ inline void con_multiply(double& re_, double& im_, int32_t& ex, bool& started,
                                    double re, double im) {
    if (!started) {
        re_ = 1.0;
        started = true;
    }
    int32_t exp = 0;
    double a = fmax(fabs(re), fabs(im));
    if (a != 0.0 && !(kLo <= a && a <= kHi) && isfinite(a)) {
        int32_t e = exponent_of(a);
        double sc = exp2i32(-e);
        re *= sc;
        im *= sc;
        exp += e;
    }
    double nr = re_ * re - im_ * im, ni = re_ * im + im_ * re;
    re_ = nr;
    im_ = ni;
    ex += exp;
    a = fmax(fabs(re_), fabs(im_));
    if (a != 0.0 && !(kLo <= a && a <= kHi) && isfinite(a)) {
        int32_t e = exponent_of(a);
        double sc = exp2i32(-e);
        re_ *= sc;
        im_ *= sc;
        ex += e;
    }
}

// sample.rs:586-595, in 32-bit exponent arithmetic (exponents here stay far inside int32).
// CAP: the delta's form, max_e = 0, so k <= -e; the derivative's max_e (i64::MAX / 4)
// never binds.  The two half-power factors are the reference's (a single 2^-k can
// overflow for subnormal mantissas).
//
// The two scalings are ldexps: each rounds as the product with that (exact) power of two,
// so the result equals re * 2^-(k/2) * 2^-(k - k/2) bit for bit.  Returns true when the
// result is canonical: normalising it again would find k = 0 and change nothing, and the
// lazy rescale after a step would not fire (the exponent came from frexp, so the larger
// magnitude now lies in [1, 2) away from the log2 tie region; or the cap left the exponent
// at 0 with the magnitude above 2; or both parts are zero).
template <bool CAP>
__device__ inline bool normalise(double& re, double& im, int32_t& e) {
    double a = fmax(fabs(re), fabs(im));
    if (a == 0.0 || !isfinite(a)) return a == 0.0;  // (a non-finite delta stays unmarked)
    bool tie = __builtin_amdgcn_frexp_mant(a) > 1.0 - 0x1p-40;
    int32_t k = tie ? (int32_t)floor_log2(a) : __builtin_amdgcn_frexp_exp(a) - 1;
    if (CAP && -e < k) k = -e;
    re = ldexp(ldexp(re, -(k / 2)), -(k - k / 2));
    im = ldexp(ldexp(im, -(k / 2)), -(k - k / 2));
    e = e + k;
    return !tie;
}

// sample.rs:1094-1114 sum2 over two terms, for any input (non-finite terms, exponents near
// the i64 range: Rust's saturating float cast and wrapping adds)
__device__ inline void sum2_general(double t0r, double t0i, int64_t t0e, double t1r,
                                          double t1i, int64_t t1e, int64_t zero_e, double& ore,
                                          double& oim, int32_t& oe) {
    bool have = false;
    int64_t top = 0;
    // (a NaN term passes the != 0 test and takes exponent t.e, as in the reference)
    double a0 = fmax(fabs(t0r), fabs(t0i)), a1 = fmax(fabs(t1r), fabs(t1i));
    if (a0 != 0.0) {
        top = wrap_add(rust_as_i64(floor(log2(a0))), t0e);
        have = true;
    }
    if (a1 != 0.0) {
        int64_t g = wrap_add(rust_as_i64(floor(log2(a1))), t1e);
        if (!have || g > top) top = g;
        have = true;
    }
    if (!have) {
        ore = 0.0;
        oim = 0.0;
        oe = (int32_t)zero_e;
        return;
    }
    int64_t k0 = wrap_sub(t0e, top), k1 = wrap_sub(t1e, top);
    double r = 0.0, i = 0.0;
    r = r + (k0 < -2046 ? 0.0 : t0r * exp2i(k0 / 2) * exp2i(k0 - k0 / 2));
    i = i + (k0 < -2046 ? 0.0 : t0i * exp2i(k0 / 2) * exp2i(k0 - k0 / 2));
    r = r + (k1 < -2046 ? 0.0 : t1r * exp2i(k1 / 2) * exp2i(k1 - k1 / 2));
    i = i + (k1 < -2046 ? 0.0 : t1i * exp2i(k1 / 2) * exp2i(k1 - k1 / 2));
    ore = r;
    oim = i;
    oe = (int32_t)top;
}

// sum2 as the scaled tier's rebases call it (three times per rebase onto Z_0 far below the
// f64 range: the sum, then the step from Z_0 for dz/dc and for the delta).  For finite
// terms with exponents inside +-2^29 the reference's i64 arithmetic never wraps or
// saturates, so it runs in 32 bits, and floor(log2(a)) comes from the frexp exponent
// (floor_log2_fast: equal to the libm expression for every finite a > 0); anything else
// takes sum2_general.  x * 2^(k/2) * 2^(k - k/2) is one ldexp per factor for an int32 k
// (exp2i's clamp at +-4096 only matters beyond int32).
__device__ inline void sum2(double t0r, double t0i, int64_t t0e, double t1r, double t1i,
                            int64_t t1e, int64_t zero_e, double& ore, double& oim, int32_t& oe) {
    double a0 = fmax(fabs(t0r), fabs(t0i)), a1 = fmax(fabs(t1r), fabs(t1i));
    constexpr int64_t kLim = int64_t(1) << 29;
    bool fast = isfinite(a0) && isfinite(a1) && t0e > -kLim && t0e < kLim && t1e > -kLim &&
                t1e < kLim;
    if (!fast) {
        sum2_general(t0r, t0i, t0e, t1r, t1i, t1e, zero_e, ore, oim, oe);
        return;
    }
    if (a0 == 0.0 && a1 == 0.0) {
        ore = 0.0;
        oim = 0.0;
        oe = (int32_t)zero_e;
        return;
    }
    const int32_t e0 = (int32_t)t0e, e1 = (int32_t)t1e;
    int32_t g0 = a0 != 0.0 ? floor_log2_fast(a0) + e0 : INT32_MIN;
    int32_t g1 = a1 != 0.0 ? floor_log2_fast(a1) + e1 : INT32_MIN;
    int32_t top = g0 > g1 ? g0 : g1;
    int32_t k0 = e0 - top, k1 = e1 - top;
    double r = 0.0, i = 0.0;
    r = r + (k0 < -2046 ? 0.0 : t0r * exp2i32(k0 / 2) * exp2i32(k0 - k0 / 2));
    i = i + (k0 < -2046 ? 0.0 : t0i * exp2i32(k0 / 2) * exp2i32(k0 - k0 / 2));
    r = r + (k1 < -2046 ? 0.0 : t1r * exp2i32(k1 / 2) * exp2i32(k1 - k1 / 2));
    i = i + (k1 < -2046 ? 0.0 : t1i * exp2i32(k1 / 2) * exp2i32(k1 - k1 / 2));
    ore = r;
    oim = i;
    oe = top;
}

// Per-lane state that only the periodicity confirmation and the nucleus heuristic use
// (sample.rs:216-230): kept in LDS so the hot loop's registers hold just the iterate.
constexpr int kBlock = 256;

// Structure of arrays in LDS (one slot per lane: no bank conflicts between lanes).
struct RareLds {
    double qr[kBlock], qi[kBlock], qstep[kBlock], mur[kBlock], mui[kBlock], conr[kBlock],
        coni[kBlock];
    int32_t mue[kBlock], cone[kBlock];
};

// One lane's view of its LDS slot.
struct Rare {
    double &qr, &qi, &qstep, &mur, &mui, &conr, &coni;
    int32_t &mue, &cone;
    __device__ Rare(RareLds& l, int t)
        : qr(l.qr[t]), qi(l.qi[t]), qstep(l.qstep[t]), mur(l.mur[t]), mui(l.mui[t]),
          conr(l.conr[t]), coni(l.coni[t]), mue(l.mue[t]), cone(l.cone[t]) {}
};

// One lane's sample state (sample.rs:190-236) in registers.  (Zr, Zi) is Z_m, carried
// from the previous escape test, so a step loads the orbit once (its landing point).
struct Lane {
    double xr, xi, dr, di, ar, ai, sr, si;
    double Zr, Zi, Z1r, Z1i;            // Z_m and Z_{m+1}
    double R0;                          // level-0 r2 of the block starting at m (when any)
    double dc_r, dc_i;                  // dc in the delta's units (scaled tier: ar 2^(oe - e))
    int32_t delta_exp, derivative_exp;  // scaled tier only
    uint32_t m, n, sn, hold, chk, nucleus, blocks, skipped, plain;
    uint32_t flags;                     // kPending | kPredicted | kConStarted | returns << 4
    uint32_t bend;                      // where a lookup's step budget ends (budget_end)
    uint32_t ev;                        // ordinary tiers: the next Brent event (event_at)
    uint32_t tb;                        // batches: first block of this frame's BLA table
    int32_t oe;                         // batches: this frame's offset exponent h_e
};

constexpr uint32_t kPending = 1, kPredicted = 2, kConStarted = 4;
// Scaled tier: the delta (kCanonX) or dz/dc (kCanonD) is as the last block's normalisation
// left it and canonical, so the next block's first normalisation of it is the identity
// (bits 8 and 9, above the close-return count in bits 4-7).
constexpr uint32_t kCanonX = 256, kCanonD = 512;
// series_carry's mark of a sample running from the series start
constexpr uint32_t kSeriesTry = 0xffffffffu;

// The iterate where Bla::find's step budget ends (sample.rs:253-261): the hold's end while
// a near-return is awaited (ordinary tiers), no bound while a deferred save is pending, else
// the next scheduled save; at most max_iter and the nucleus test iterate p + 1.  (The
// reference's budget of 1 at n = 0 never meets a lookup: n = 0 has m = 0, where none is
// made.)  It reads only state that changes at a near-return, a save, a hold's expiry, a
// deferred save and the nucleus test, where it is recomputed, so a lookup subtracts n.
template <bool EXT>
__device__ inline uint32_t budget_end(const Lane& s, uint32_t max_iter) {
    uint32_t b = EXT ? s.chk
                     : (s.hold > 0 ? s.sn + s.hold : ((s.flags & kPending) ? 0xffffffffu : s.chk));
    b = b < max_iter ? b : max_iter;
    uint32_t nb = s.nucleus ? s.nucleus + 1 : 0xffffffffu;
    return b < nb ? b : nb;
}

// Ordinary tiers: the first iterate at which a hold can expire (sn + hold), a save be
// deferred or made (chk; 0 while a deferred save is pending): before it the updates after
// a step that met no near-return change nothing.  Recomputed with budget_end.
__device__ inline uint32_t event_at(const Lane& s) {
    return s.hold > 0 ? s.sn + s.hold : ((s.flags & kPending) ? 0u : s.chk);
}

__device__ inline uint32_t np2_above(uint32_t n) {
    // next_power_of_two(n + 1) for 1 <= n < 2^31
    return 1u << (32 - __builtin_clz(n));
}

// 2^delta_exp and 2^-derivative_exp, the scaled tier's unit of the delta and inverse unit
// of the derivative (1 on the ordinary tiers, whose exponents stay 0): recomputed where
// used (one v_ldexp each) rather than carried in registers.
template <bool EXT>
__device__ inline double scale_of(const Lane& s) {
    return EXT ? exp2i32(s.delta_exp) : 1.0;
}
template <bool EXT>
__device__ inline double inverse_of(const Lane& s) {
    return EXT ? exp2i32(-s.derivative_exp) : 1.0;
}

// Pointers into device (global) memory.  Pointers read from a struct in memory are
// generic to the compiler, which then emits flat accesses (slower to issue, and counted
// against the LDS wait counter too); these casts keep the hot accesses global.
#define FDG_GLOBAL __attribute__((address_space(1)))
template <class T>
__device__ inline FDG_GLOBAL T* global_ptr(T* p) {
    return (FDG_GLOBAL T*)p;
}
// Native vector types load from global memory as one 16- / 32-byte access (HIP's double2
// and double4 classes cannot be copied out of an address-space-qualified lvalue).
typedef double v2d __attribute__((ext_vector_type(2)));
typedef double v4d __attribute__((ext_vector_type(4)));
// (32-bit element indices.  The orbit stays below 4 GiB (max_iter < 2^27, checked by the
// host), so its byte offsets are formed in 32 bits: load32.  So do a single frame's
// tables'; a batch's tables, one per frame in one array, may exceed 4 GiB: table_load.)
__device__ inline double2 ld2(const FDG_GLOBAL v2d* p, uint32_t i) {
    v2d t = load32(p, i);
    return make_double2(t.x, t.y);
}
template <bool WIDE>
__device__ inline double4 ld4(const FDG_GLOBAL v4d* p, uint32_t i) {
    v4d t = table_load<WIDE>(p, i);
    return make_double4(t.x, t.y, t.z, t.w);
}

// Frame constants the loop reads, copied once into registers / SGPRs.
struct Hot {
    const FDG_GLOBAL v2d* orbit;
    const FDG_GLOBAL double* r2;
    ConstR2 r2c;                        // the same radii, for scalar loads
    const FDG_GLOBAL v4d* ab;
    uint32_t last, max_iter, nlevels, steps;
    uint64_t stride;
    uint32_t stride32;                  // the same (tables stay below 2^32 blocks)
    double esc2;
    int32_t oe;
    uint32_t len, count0;               // points; level-0 blocks
    double Z1r, Z1i, Z2r, Z2i, R00;     // Z_1, Z_2 and the level-0 r2 at m = 1
};

template <bool EXT>
__device__ inline void lane_start(Lane& s, Rare& r, const Hot& h, uint32_t nucleus, double ar,
                                  double ai, int32_t oe) {
    s.ar = ar;
    s.ai = ai;
    s.xr = 0.0;
    s.xi = 0.0;
    s.Zr = 0.0;  // Z_0 = 0
    s.Zi = 0.0;
    s.Z1r = h.Z1r;
    s.Z1i = h.Z1i;
    s.R0 = 0.0;  // m = 0: no block starts here
    s.delta_exp = EXT ? oe : 0;
    s.dc_r = ar;
    s.dc_i = ai;
    s.dr = 0.0;
    s.di = 0.0;
    s.derivative_exp = 0;
    s.m = 0;
    s.n = 0;
    s.sr = 0.0;
    s.si = 0.0;
    s.chk = 16;
    s.sn = EXT ? 0 : 8;
    s.hold = 0;
    s.nucleus = nucleus;
    s.flags = 0;
    s.blocks = 0;
    s.skipped = 0;
    s.plain = 0;
    r.qr = r.qi = r.qstep = 0.0;
    r.mur = 1.0;
    r.mui = 0.0;
    r.mue = 0;
    r.conr = 0.0;
    r.coni = 0.0;
    r.cone = 0;
    s.bend = budget_end<EXT>(s, h.max_iter);
    s.ev = event_at(s);
}

// Re-derive dc in the delta's units after its exponent changed (sample.rs:300-302).
template <bool EXT>
__device__ inline void refresh_dc(Lane& s, int32_t oe) {
    if (EXT) {
        double factor = exp2i32(oe - s.delta_exp);
        s.dc_r = s.ar * factor;
        s.dc_i = s.ai * factor;
    }
}


// A supplied start on the ordinary tiers (sample.rs:186-228 with Start from grid.rs:663-673):
// the series state at the sample's offset (Series::eval, series.rs:236-257: Horner on
// x = dc / rho for p = sum b[k] x^(k-1) and q = p', then S = x p and dz/dc = (p + x q) / rho),
// continued from iterate n with the state as the first Brent save, saves resuming at the
// next power of two, and no nucleus contraction (it would miss the skipped factors).
// Called after lane_start.
template <bool D, bool L>
__device__ inline void series_start(Lane& s, const Hot& h, const FrameDev* __restrict__ fp) {
    const FDG_GLOBAL v2d* b = global_ptr(reinterpret_cast<const v2d*>(fp->series_b));
    const double sc = fp->series_inv_rho;
    const double xr = s.ar * sc, xi = s.ai * sc;
    double pr = 0.0, pi = 0.0, qr = 0.0, qi = 0.0;
    for (uint32_t k = fp->series_len; k-- > 0;) {
        double2 c = ld2(b, k);
        if (D) {
            double tr = qr * xr - qi * xi + pr;
            qi = qr * xi + qi * xr + pi;
            qr = tr;
        }
        double tr = pr * xr - pi * xi + c.x;
        pi = pr * xi + pi * xr + c.y;
        pr = tr;
    }
    s.xr = pr * xr - pi * xi;
    s.xi = pr * xi + pi * xr;
    if (D) {
        double dr = pr + qr * xr - qi * xi, di = pi + qr * xi + qi * xr;
        s.dr = dr * sc;
        s.di = di * sc;
    }
    const uint32_t m = fp->series_n;  // 1 <= m <= len - 2
    double2 zm = ld2(h.orbit, m), z1 = ld2(h.orbit, m + 1);
    s.m = m;
    s.n = m;
    s.Zr = zm.x;
    s.Zi = zm.y;
    s.Z1r = z1.x;
    s.Z1i = z1.y;
    // the level-0 radius of the block starting at m, when one does (k = m - 1 even)
    uint32_t k0 = m - 1;
    s.R0 = (L && (k0 & 1u) == 0 && (k0 >> 1) < h.count0) ? h.r2[k0 >> 1] : 0.0;
    s.sr = zm.x + s.xr;
    s.si = zm.y + s.xi;
    s.sn = m;
    uint32_t chk = np2_above(m);
    s.chk = chk > 16u ? chk : 16u;
    s.nucleus = 0;
    s.bend = budget_end<false>(s, h.max_iter);
    s.ev = event_at(s);
}

enum : int { RUNNING = -1 };

// Rare path of a near-return on the ordinary tiers (sample.rs:524-558): Newton's cycle
// point prediction from the return multiplier; returns true when two consecutive
// predictions agree (Interior).
__device__ bool near_return(Lane& s, Rare& r, double fr, double fi, double size) {
    uint32_t period = s.n - s.sn;
    if (s.hold == period && scaled_below(r.mur, r.mui, r.mue, 1.0)) {
        double sc = r.mue < -1000 ? 0.0 : exp2i(r.mue < 1000 ? r.mue : 1000);
        double mr = r.mur * sc, mi = r.mui * sc;
        double gr = fr - s.sr, gi = fi - s.si;
        double er = mr - 1.0, ei = mi;
        double den = er * er + ei * ei;
        double pr2 = s.sr - (gr * er + gi * ei) / den;
        double pi2 = s.si - (gi * er - gr * ei) / den;
        double tol = fmin(0.1 * sqrt(den), 1e-2) * r.qstep + 16.0 * kU * size;
        if ((s.flags & kPredicted) && fabs(pr2 - r.qr) + fabs(pi2 - r.qi) <= tol) return true;
        r.qr = pr2;
        r.qi = pi2;
        r.qstep = (fabs(gr) + fabs(gi)) / sqrt(den);
        s.flags |= kPredicted;
    } else {
        s.flags &= ~kPredicted;
    }
    s.hold = period;
    s.sr = fr;
    s.si = fi;
    s.sn = s.n;
    s.flags &= ~kPending;
    r.mur = 1.0;
    r.mui = 0.0;
    r.mue = 0;
    return false;
}

// The scaled tier's rebase (sample.rs:422-449): exact two-term sums below 1e-290 and the
// exponent-safe step from Z_0 when the delta is below 1e-270.
template <bool D>
__device__ void ext_rebase(Lane& s, const Hot& h, int32_t oe, double R00, double fr, double fi,
                           double f2, double Zmr, double Zmi) {
    if (f2 > kSquareOk) {
        s.xr = fr;
        s.xi = fi;
        s.delta_exp = 0;
    } else {
        sum2(Zmr, Zmi, 0, s.xr, s.xi, s.delta_exp, s.delta_exp, s.xr, s.xi, s.delta_exp);
    }
    s.m = 0;
    s.Zr = 0.0;
    s.Zi = 0.0;
    s.Z1r = h.Z1r;
    s.Z1i = h.Z1i;
    if (scale_of<true>(s) < kTiny) {
        s.nucleus = 0;
        s.bend = budget_end<true>(s, h.max_iter);
        if (D) {
            double qr = 2.0 * (s.xr * s.dr - s.xi * s.di);
            double qi = 2.0 * (s.xr * s.di + s.xi * s.dr);
            sum2(qr, qi, (int64_t)s.delta_exp + s.derivative_exp, 1.0, 0.0, 0, s.derivative_exp,
                 s.dr, s.di, s.derivative_exp);
        }
        double wr = s.xr, wi = s.xi;
        sum2(wr * wr - wi * wi, 2.0 * wr * wi, 2 * (int64_t)s.delta_exp, s.ar, s.ai, oe,
             s.delta_exp, s.xr, s.xi, s.delta_exp);
        s.m = 1;
        s.n = s.n + 1;
        s.Zr = h.Z1r;
        s.Zi = h.Z1i;
        s.Z1r = h.Z2r;
        s.Z1i = h.Z2i;
        s.R0 = R00;
    }
    refresh_dc<true>(s, oe);
}

// One block or one plain step plus the checks after it (sample.rs:237-575).  Returns
// RUNNING, or the finished outcome kind (0 escaped with z in ofr/ofi, 1 interior,
// 2 unresolved).
template <bool EXT, bool D, bool L, bool B>
__device__ inline int lane_advance(Lane& s, Rare& r, const Hot& h, double& ofr, double& ofi) {
    // sample.rs:237 `while n < limit`: the budget ran out.
    if (s.n >= h.max_iter) return 2;
    // A batch's frames each have their own table (at block tb) and offset exponent.
    const uint32_t tb = B ? s.tb : 0u;
    const int32_t oe = B ? s.oe : h.oe;
    // Gathers for the NEXT iteration, assuming this one is a plain step that does not
    // rebase (the common case): Z_{m+2} and the level-0 r2 of a block starting at m+1.
    // Their L2 latency then overlaps this whole iteration.  (s.m < last here, and the
    // orbit buffer is padded past its end.)
    double2 pz2 = ld2(h.orbit, s.m + 2);
    double pr0 = 0.0;
    if (L && (s.m & 1u) == 0 && (s.m >> 1) < h.count0) pr0 = table_load<B>(h.r2, tb + (s.m >> 1));
    double2 Zn = make_double2(s.Z1r, s.Z1i);  // a plain step's landing point
    double2 Znext = pz2;
    double r0next = pr0;
    int lvl = -1;
    if (L) {
        // Bla::find (bla.rs:234-249): a block of 2^(i+1) steps starts at k = m - 1 divisible
        // by 2^(i+1), so an odd k (m = 0 included: k wraps) has none.
        uint32_t k = s.m - 1;
        if ((k & 1u) == 0 && (s.nucleus == 0 || s.m != 0)) {
            double d2 = EXT ? (s.xr * s.xr + s.xi * s.xi) * scale_of<EXT>(s) * scale_of<EXT>(s)
                            : s.xr * s.xr + s.xi * s.xi;
            // level 0 from the carried radius (0 when no such block exists) first: a wave
            // whose lanes all fail it skips the budget arithmetic
            if (d2 < s.R0) {
                // holds and deferred saves exist on the ordinary tiers only
                uint32_t b = s.bend > s.n ? s.bend - s.n : 0u;
                // Lanes in lockstep (the same table, start and budget in every searching
                // lane) take the cap and the radii through the scalar unit.
                const uint32_t k0 = __builtin_amdgcn_readfirstlane(k),
                               b0 = __builtin_amdgcn_readfirstlane(b),
                               t0 = B ? __builtin_amdgcn_readfirstlane(tb) : 0u;
                if (__ballot(k != k0 || b != b0 || (B && tb != t0)) == 0) {
                    int cap = find_cap(k0, h.nlevels, b0);
                    if (cap >= 0)
                        lvl = find_level_down_uniform(h.r2c, t0, h.stride32, k0, cap, d2);
                } else {
                    int cap = find_cap(k, h.nlevels, b);
                    if (cap >= 0) lvl = find_level_down<B>(h.r2, tb, h.stride32, k, cap, d2);
                }
            }
        }
    }
    double Zmr = s.Zr, Zmi = s.Zi;
    if (!EXT) {
        // A block (delta' = A delta + B dc, D' = A D + B, bla.rs) or a plain step
        // (delta' = 2 Z delta + delta^2 + dc, D' = 2 z D + 1, sample.rs:343-381), each in the
        // reference's operation order.  A wave whose lanes take both runs both branches.
        const bool blk = lvl >= 0;
        const double xr = s.xr, xi = s.xi;
        const double fr = Zmr + xr, fi = Zmi + xi;  // z = Z_m + delta (plain step)
        double mr, mi;                              // the step's multiplier
        const bool contract = s.nucleus && s.n >= 1;  // (n before the step)
        if (blk) {
            uint32_t at = bla_block_index(tb, h.stride32, s.m - 1, lvl);
            double4 ab = ld4<B>(h.ab, at);
            const double a_r = ab.x, a_i = ab.y, b_r = ab.z, b_i = ab.w;
            const uint32_t blen = 2u << lvl;
            Zn = ld2(h.orbit, s.m + blen);
            Znext = ld2(h.orbit, s.m + blen + 1);
            uint32_t i0 = (s.m + blen - 1) >> 1;
            r0next = i0 < h.count0 ? table_load<B>(h.r2, tb + i0) : 0.0;
            mr = a_r;
            mi = a_i;
            if (D) {
                double dr = s.dr, di = s.di;
                s.dr = a_r * dr - a_i * di + b_r;
                s.di = a_r * di + a_i * dr + b_i;
            }
            s.xr = a_r * xr - a_i * xi + b_r * s.dc_r - b_i * s.dc_i;
            s.xi = a_r * xi + a_i * xr + b_r * s.dc_i + b_i * s.dc_r;
            s.m += blen;
            s.n += blen;
            s.blocks += 1;
            s.skipped += blen;
        } else {
            mr = 2.0 * fr;
            mi = 2.0 * fi;
            if (D) {
                double dr = s.dr, di = s.di;
                s.dr = 2.0 * (fr * dr - fi * di) + 1.0;
                s.di = 2.0 * (fr * di + fi * dr);
            }
            s.xr = 2.0 * (Zmr * xr - Zmi * xi) + xr * xr - xi * xi + s.dc_r;
            s.xi = 2.0 * (Zmr * xi + Zmi * xr + xr * xi) + s.dc_i;
            s.m += 1;
            s.n += 1;
            s.plain += 1;
        }
        if (contract) {
            bool st = s.flags & kConStarted;
            con_multiply(r.conr, r.coni, r.cone, st, mr, mi);
            s.flags |= kConStarted;
        }
        if (s.hold > 0) mu_multiply(r.mur, r.mui, r.mue, mr, mi);
    } else if (lvl >= 0) {
        uint32_t at = bla_block_index(tb, h.stride32, s.m - 1, lvl);
        double4 ab = ld4<B>(h.ab, at);
        double a_r = ab.x, a_i = ab.y, b_r = ab.z, b_i = ab.w;
        uint32_t blen = 2u << lvl;
        Zn = ld2(h.orbit, s.m + blen);
        Znext = ld2(h.orbit, s.m + blen + 1);
        uint32_t i0 = (s.m + blen - 1) >> 1;
        r0next = L && i0 < h.count0 ? table_load<B>(h.r2, tb + i0) : 0.0;
        if (s.nucleus && s.n >= 1) {
            bool st = s.flags & kConStarted;
            con_multiply(r.conr, r.coni, r.cone, st, a_r, a_i);
            s.flags |= kConStarted;
        }
        // (skipped where it is the identity: the state is still the last block's
        // canonical result)
        if (!(s.flags & kCanonX)) {
            normalise<true>(s.xr, s.xi, s.delta_exp);
            refresh_dc<EXT>(s, oe);
        }
        if (D && !(s.flags & kCanonD)) normalise<false>(s.dr, s.di, s.derivative_exp);
        if (D) {
            double inv = inverse_of<EXT>(s);
            double t = a_r * s.dr - a_i * s.di + b_r * inv;
            s.di = a_r * s.di + a_i * s.dr + b_i * inv;
            s.dr = t;
        }
        double t = a_r * s.xr - a_i * s.xi + b_r * s.dc_r - b_i * s.dc_i;
        s.xi = a_r * s.xi + a_i * s.xr + b_r * s.dc_i + b_i * s.dc_r;
        s.xr = t;
        bool cx = normalise<true>(s.xr, s.xi, s.delta_exp);
        bool cd = D ? normalise<false>(s.dr, s.di, s.derivative_exp) : false;
        refresh_dc<EXT>(s, oe);
        s.flags = (s.flags & ~(kCanonX | kCanonD)) | (cx ? kCanonX : 0u) | (cd ? kCanonD : 0u);
        s.m += blen;
        s.n += blen;
        s.blocks += 1;
        s.skipped += blen;
    } else {
        const double scale = scale_of<EXT>(s);
        double fr = Zmr + scale * s.xr, fi = Zmi + scale * s.xi;
        if (s.nucleus && s.n >= 1) {
            bool st = s.flags & kConStarted;
            con_multiply(r.conr, r.coni, r.cone, st, 2.0 * fr, 2.0 * fi);
            s.flags |= kConStarted;
        }
        if (D) {
            double t = 2.0 * (fr * s.dr - fi * s.di) + inverse_of<EXT>(s);
            s.di = 2.0 * (fr * s.di + fi * s.dr);
            s.dr = t;
            if (fmax(fabs(s.dr), fabs(s.di)) > kHi) {
                s.dr = s.dr * kLo;
                s.di = s.di * kLo;
                s.derivative_exp += 32;
            }
        }
        double xr = s.xr, xi = s.xi;
        s.xr = 2.0 * (Zmr * xr - Zmi * xi) + scale * (xr * xr - xi * xi) + s.dc_r;
        s.xi = 2.0 * (Zmr * xi + Zmi * xr) + scale * (2.0 * xr * xi) + s.dc_i;
        s.flags &= ~(kCanonX | kCanonD);
        s.m += 1;
        s.n += 1;
        s.plain += 1;
    }
    double pr = scale_of<EXT>(s) * s.xr, pi = scale_of<EXT>(s) * s.xi;
    double fr = Zn.x + pr, fi = Zn.y + pi;
    double f2 = fr * fr + fi * fi;
    const bool escaped = f2 > h.esc2;
    // One exit: a lane that escaped (or is found interior) here runs the checks below
    // too, which change only state its outcome does not read (the orbit position, the
    // Brent/hold state and the LDS slot; n, the operation counts and dz/dc are final:
    // the scaled tier's rebase, which can change dz/dc, skips it), and keeps the first
    // outcome.
    int out = escaped ? 0 : RUNNING;
    ofr = fr;
    ofi = fi;
    if (s.nucleus && s.n == s.nucleus + 1) {
        if ((s.flags & kConStarted) && scaled_below(r.conr, r.coni, r.cone, 1e-6))
            out = escaped ? 0 : 1;
        s.nucleus = 0;
        s.bend = budget_end<EXT>(s, h.max_iter);
    }
    double delta2 = pr * pr + pi * pi;
    bool rebase = (EXT ? f2 <= delta2 : f2 < delta2) || s.m == h.last;
    if (EXT) {
        s.Zr = Zn.x;
        s.Zi = Zn.y;
        s.Z1r = Znext.x;
        s.Z1i = Znext.y;
        s.R0 = r0next;
        const bool rb = rebase && !escaped;
        if (rb) {
            double R00 = B ? (L && h.count0 > 0 ? h.r2[tb] : 0.0) : h.R00;
            ext_rebase<D>(s, h, oe, R00, fr, fi, f2, Zn.x, Zn.y);
            s.flags &= ~(kCanonX | kCanonD);
        }
        // (in sequence rather than if/else, so each branch updates the state in place)
        if (!rb && !(s.flags & kCanonX)) {
            // (a block's canonical delta has its larger part in [1, 2): nothing to rescale)
            double magnitude = fmax(fabs(s.xr), fabs(s.xi));
            int32_t k = 0;
            if (magnitude > kHi && s.delta_exp < 0)
                k = -s.delta_exp < 32 ? -s.delta_exp : 32;
            else if (magnitude < kLo && magnitude != 0.0)
                k = -32;
            if (k != 0) {
                double factor = exp2i32(-k);
                s.xr = s.xr * factor;
                s.xi = s.xi * factor;
                s.delta_exp = s.delta_exp + k;
                refresh_dc<EXT>(s, oe);
                s.flags &= ~kCanonX;
            }
        }
        // extended tier: judged on the physical delta before rebase/rescale
        uint32_t returns = (s.flags >> 4) & 15u;
        bool res = returns < 4 && pr * pr + pi * pi > kResolvable * kResolvable * f2;
        if (res) {
            double dist = fabs(fr - s.sr) + fabs(fi - s.si);
            double size = fabs(s.sr) + fabs(s.si);
            if (dist < 1e-13 * size + 1e-300) {
                if (contracting(reinterpret_cast<const double*>((const v2d*)h.orbit), h.len, s.m, pr, pi,
                                s.ar * exp2i(oe), s.ai * exp2i(oe), s.n - s.sn) &&
                    out == RUNNING)
                    out = 2;  // Unresolved
                s.flags += 16;
            }
        }
        bool save = res ? s.n >= s.chk : s.n >= s.chk;  // n == chk || n > chk
        if (save) {
            s.sr = fr;
            s.si = fi;
            s.sn = s.n;
            s.chk = s.chk * 2;
            s.bend = budget_end<true>(s, h.max_iter);
        }
        return out;
    }
    // ordinary tiers: rebase onto Z_0 = 0 (sample.rs:451)
    s.Zr = Zn.x;
    s.Zi = Zn.y;
    s.Z1r = Znext.x;
    s.Z1i = Znext.y;
    s.R0 = r0next;  // after a rebase m = 0 and no block is looked up
    if (rebase) {
        s.xr = fr;
        s.xi = fi;
        s.m = 0u;
        s.Zr = 0.0;
        s.Zi = 0.0;
        s.Z1r = h.Z1r;
        s.Z1i = h.Z1i;
    }
    // |delta|^2 after the rebase select: f2 when rebased, else delta2 (same expressions)
    bool res = (rebase ? f2 : delta2) > kResolvable * kResolvable * f2;
    double dist = fabs(fr - s.sr) + fabs(fi - s.si);
    double size = fabs(s.sr) + fabs(s.si);
    const bool near = res && dist < 1e-3 * size;
    if (near && near_return(s, r, fr, fi, size) && out == RUNNING) out = 1;  // Interior
    // Otherwise a hold's expiry and the saves happen only from n = ev on (a lane that met a
    // near-return holds from this iterate).  (Sequential rather than if/else, so each
    // branch updates the state in place.)
    const bool upd = !near && s.n >= s.ev;
    if (upd) {
        uint32_t h0 = s.hold;
        bool expired = h0 > 0 && s.n >= s.sn + h0;
        if (expired) {
            s.hold = 0;
            s.flags = (s.flags & ~kPredicted) | kPending;
        }
        // !res: a save that falls on an unjudgeable iterate is deferred (pending)
        bool defer = !res && h0 == 0 && s.n >= s.chk;
        s.flags |= defer ? kPending : 0u;
        // res: save at the deferred or scheduled point unless a return is being awaited
        bool save = res && (h0 == 0 || expired) && ((s.flags & kPending) || s.n >= s.chk);
        if (save) {
            uint32_t np = np2_above(s.n);
            s.chk = s.chk > np ? s.chk : np;
            s.sr = fr;
            s.si = fi;
            s.sn = s.n;
            s.flags &= ~kPending;
        }
    }
    if (near || upd) {
        s.bend = budget_end<false>(s, h.max_iter);
        s.ev = event_at(s);
    }
    return out;
}

__device__ inline uint32_t lanes_below(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// Persistent work-stealing kernel: every lane owns one sample at a time; lanes whose
// sample finished are refilled together from one atomic per wave, so the wave keeps
// 64 samples in flight however uneven their iteration counts are.  With B, the samples
// are those of a batch of frames sharing the orbit (frame-major), each frame with its own
// geometry and BLA table, so one launch drains every frame and no frame's tail idles
// the GPU.
template <bool EXT, bool D, bool L, bool B, bool LIST = false, bool SER = false>
__global__ void __launch_bounds__(kBlock, 4) perturb_f64(const FrameDev* __restrict__ fp) {
    __shared__ RareLds rare;
    // SER (ordinary tiers, single frames): per lane, kSeriesTry while the sample runs from
    // the series start, else the iterates its unresolved series attempt spent (grid.rs:
    // 623-649: that attempt's work counts, minus the skipped prefix, and it retries from 0)
    __shared__ uint32_t series_carry[SER ? kBlock : 1];
    // Statistics accumulate in LDS at each sample's end (rare next to the steps), not in
    // per-lane registers held across the loop: per-sample counts in the lane's own slots,
    // folded into the block's counters at the end.
    __shared__ unsigned long long lane_slots[kLaneStats * kBlock];
    LaneStats<kBlock> ls{lane_slots};
    ls.clear();
    __syncthreads();
    const uint32_t lane = __lane_id();
    Rare r(rare, threadIdx.x);
    FDG_GLOBAL uint8_t* o_kind = global_ptr(fp->kind);
    FDG_GLOBAL uint64_t* o_n = global_ptr(fp->n);
    FDG_GLOBAL double* o_zr = global_ptr(fp->zr);
    FDG_GLOBAL double* o_zi = global_ptr(fp->zi);
    FDG_GLOBAL double* o_dr = global_ptr(fp->dr);
    FDG_GLOBAL double* o_di = global_ptr(fp->di);
    FDG_GLOBAL int64_t* o_dexp = global_ptr(fp->dexp);
    FDG_GLOBAL uint32_t* o_ops = global_ptr(fp->ops);
    Hot h;
    h.orbit = global_ptr(reinterpret_cast<const v2d*>(fp->orbit));
    h.r2 = global_ptr(fp->bla.r2);
    h.r2c = (ConstR2)fp->bla.r2;
    h.ab = global_ptr(reinterpret_cast<const v4d*>(fp->bla.ab));
    h.last = (uint32_t)fp->len - 1;
    h.len = (uint32_t)fp->len;
    h.max_iter = (uint32_t)fp->max_iter;
    h.nlevels = fp->bla.nlevels;
    h.steps = (uint32_t)fp->bla.steps;
    h.stride = fp->bla.stride;
    h.stride32 = (uint32_t)fp->bla.stride;
    h.esc2 = fp->r2;
    h.oe = (int32_t)fp->h_e;
    h.count0 = (uint32_t)(fp->bla.steps >> 1);
    {
        double2 z1 = ld2(h.orbit, 1), z2 = ld2(h.orbit, 2);
        h.Z1r = z1.x;
        h.Z1i = z1.y;
        h.Z2r = z2.x;
        h.Z2i = z2.y;
        h.R00 = (L && h.count0 > 0) ? h.r2[0] : 0.0;
    }
    // LIST: an index list (the mixed form's refinement pass) replaces 0 .. count
    const uint32_t count = LIST ? *fp->list_count : fp->count, nx = fp->nx,
                   nucleus = (uint32_t)fp->nucleus;
    // An empty refinement list (a frame with no escaped sample): nothing to claim, and no
    // statistics to add (count is uniform over the launch, so every thread leaves here).
    if (LIST && count == 0) return;
    uint32_t* next = fp->next;
    Lane s;
    bool active = false;
    bool drained = false;
    uint32_t idx = 0;
    unsigned long long witers = 0, liters = 0;
    // Each wave claims samples in chunks, one global atomic per chunk rather than per
    // refill; the chunk shrinks with the work left (guided self-scheduling) so the last
    // chunks are small and the tail short.  [cb, ce) is the wave's unclaimed range.
    // finished: samples this wave has started (closed-form interiors included; a masked-out
    // index or an exhausted claim starts none)
    uint32_t cb = 0, ce = 0, seen = 0, finished = 0;
    const uint32_t waves = gridDim.x * (kBlock / 64);
    // claim policy (the launch's tuning; 0 = the defaults)
    const uint32_t cdiv = fp->claim_div ? fp->claim_div : 4u;
    const uint32_t ccap = fp->claim_cap ? fp->claim_cap : 1024u;
    // Refilling only once every lane is idle keeps a wave's 64 samples (a tile) in lockstep:
    // the same lookups (through the scalar unit) and the same branches, most of their way.
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
                // quarter of it.  (The mixed form's list holds escaped samples, long ones.)
                uint32_t div = !LIST && witers + 64 < finished ? 1u : cdiv;
                uint32_t size = max(min(left / (div * waves), ccap), want - avail);
                // Every claim is a whole number of 64-sample groups, so every claimed
                // range starts on a group boundary: no two waves write parts of one
                // 64-byte line of the 1-byte outcome kinds.
                size = (size + 63u) & ~63u;
                uint32_t leader = __ffsll((unsigned long long)idle) - 1;
                if (lane == leader) nb = atomicAdd(next, size);
                // readlane keeps the range wave-uniform (scalar registers)
                nb = __builtin_amdgcn_readlane(nb, leader);
                seen = nb + size;
                ne = nb < count ? min(nb + size, count) : nb;
            }
            uint32_t rank = lanes_below(idle);
            uint32_t my = rank < avail ? cb + rank : nb + (rank - avail);
            bool have = rank < avail || my < ne;
            if (avail >= want) {
                cb += want;
            } else {
                cb = min(nb + (want - avail), ne);
                ce = ne;
                if (nb >= count) drained = true;  // the global range is exhausted
            }
            bool took = false;
            if (!active && have) {
                // batches: the mask (when set) covers every frame's samples, frame-major
                uint32_t bf = B ? my / fp->per_frame : 0u;
                uint32_t loc = B ? sample_of(my - bf * fp->per_frame, nx, fp->ny) : 0u;
                bool bmask = B && fp->mask && !fp->mask[fp->frames[bf].out0 + loc];
                // single frames: the list entry, or the sample the claimed index stands for
                uint32_t one = !B ? (LIST ? fp->list[my] : sample_of(my, nx, fp->ny)) : 0u;
                if (B && !bmask) {
                    took = true;
                    uint32_t f = bf;
                    const BatchFrameDev& g = fp->frames[f];
                    idx = g.out0 + loc;
                    uint32_t i = loc % nx, j = loc / nx;
                    double x = (double)i + 0.5 - g.half_x;
                    double y = (double)j + 0.5 - g.half_y;
                    double ux = g.cos_r * x + g.sin_r * y;
                    double uy = g.sin_r * x - g.cos_r * y;
                    o_ops[idx] = 0;
                    double ar = EXT ? ux * g.h_m : ux * g.h, ai = EXT ? uy * g.h_m : uy * g.h;
                    if (!EXT && fp->tier == 0 && in_main_components(fp->c_re + ar, fp->c_im + ai)) {
                        // f64 tier: closed-form interior (grid.rs Job::one, interior.rs:6-11)
                        o_kind[idx] = 1;
                        o_n[idx] = 0;
                        o_zr[idx] = 0.0;
                        o_zi[idx] = 0.0;
                        o_dr[idx] = 0.0;
                        o_di[idx] = 0.0;
                        o_dexp[idx] = 0;
                        if (L) ls.add(ST_CLOSED, 1ull);
                    } else {
                        lane_start<EXT>(s, r, h, nucleus, ar, ai, EXT ? (int32_t)g.h_e : 0);
                        s.tb = g.tb;
                        s.oe = (int32_t)g.h_e;
                        active = true;
                    }
                } else if (!B && (LIST || !fp->mask || fp->mask[one])) {
                    took = true;
                    idx = one;
                    uint32_t i = idx % nx, j = idx / nx;
                    double x = (double)i + 0.5 - fp->half_x;
                    double y = (double)j + 0.5 - fp->half_y;
                    double ux = fp->cos_r * x + fp->sin_r * y;
                    double uy = fp->sin_r * x - fp->cos_r * y;
                    o_ops[idx] = 0;
                    if (fp->tier == 0) {
                        double ar = ux * fp->h, ai = uy * fp->h;
                        if (in_main_components(fp->c_re + ar, fp->c_im + ai)) {
                            o_kind[idx] = 1;
                            o_n[idx] = 0;
                            o_zr[idx] = 0.0;
                            o_zi[idx] = 0.0;
                            o_dr[idx] = 0.0;
                            o_di[idx] = 0.0;
                            o_dexp[idx] = 0;
                            if (L) ls.add(ST_CLOSED, 1ull);
                        } else {
                            lane_start<EXT>(s, r, h, nucleus, ar, ai, 0);
                            active = true;
                        }
                    } else if (fp->tier == 1) {
                        lane_start<EXT>(s, r, h, nucleus, ux * fp->h, uy * fp->h, 0);
                        active = true;
                    } else {
                        lane_start<EXT>(s, r, h, nucleus, ux * fp->h_m, uy * fp->h_m, h.oe);
                        active = true;
                    }
                    if (SER && active) {
                        // (the series kernels run the ordinary tiers only)
                        series_start<D, L>(s, h, fp);
                        series_carry[threadIdx.x] = kSeriesTry;
                    }
                }
            }
            if (!LIST) finished += __popcll(__ballot(took));
        }
        // Fill the wave before stepping: a masked-out index or a closed-form interior
        // leaves its lane idle, and the steps below refill only at `refill` idle lanes.
        if (__ballot(!active) && !drained) continue;
        if (!__ballot(active)) break;  // (drained)
        // Step until `refill` lanes are idle (then refill) or, once drained, all are done.
        do {
            witers += 1;
            liters += __popcll(__ballot(active));
            if (active) {
                double fr = 0.0, fi = 0.0;
                int k = lane_advance<EXT, D, L, B>(s, r, h, fr, fi);
                uint32_t carried = 0, prefix = 0;
                if (SER && k != RUNNING) {
                    carried = series_carry[threadIdx.x];
                    if (carried == kSeriesTry) {
                        carried = 0;
                        prefix = fp->series_n;
                        if (k == 2) {
                            // an unresolved series start retries from zero, where the
                            // nucleus heuristic the start declined applies again; the
                            // work counters run on (one Skip per sample)
                            series_carry[threadIdx.x] = h.max_iter - prefix;
                            uint32_t bl = s.blocks, sk = s.skipped, pl = s.plain;
                            lane_start<EXT>(s, r, h, nucleus, s.ar, s.ai, 0);
                            s.blocks = bl;
                            s.skipped = sk;
                            s.plain = pl;
                            k = RUNNING;
                        }
                    }
                }
                if (k != RUNNING) {
                    // Every field is written (zeros where the outcome has none), so the
                    // output buffers need no clearing between passes.
                    bool esc = k == 0;
                    o_kind[idx] = (uint8_t)k;
                    o_n[idx] = s.n;
                    o_ops[idx] = s.plain + s.blocks;
                    o_zr[idx] = esc ? fr : 0.0;
                    o_zi[idx] = esc ? fi : 0.0;
                    o_dr[idx] = esc ? s.dr : 0.0;
                    o_di[idx] = esc ? s.di : 0.0;
                    o_dexp[idx] = esc ? s.derivative_exp : 0;
                    unsigned long long spent = (k == 2 ? h.max_iter : s.n) - prefix + carried;
                    ls.add(ST_EQUIVALENT, spent);
                    ls.add(ST_ITERATIONS, L ? spent - s.skipped + s.blocks : spent);
                    if (L) {
                        ls.add(ST_BLOCKS, (unsigned long long)s.blocks);
                        ls.add(ST_SKIPPED, (unsigned long long)s.skipped);
                        if (s.blocks == 0) ls.add(ST_FALLBACK, 1ull);
                    }
                    ls.add(ST_PLAIN, (unsigned long long)s.plain);
                    ls.max(ST_MAXOPS, (unsigned long long)(s.plain + s.blocks));
                    active = false;
                }
            }
        } while (__popcll(__ballot(!active)) < refill || (drained && __ballot(active)));
    }
    // The block's counters take the LDS of the lanes' rare state, which no lane reads
    // after the barrier.
    __syncthreads();
    unsigned long long* stat = reinterpret_cast<unsigned long long*>(rare.qr);
    if (threadIdx.x < ST_COUNT) stat[threadIdx.x] = 0;
    __syncthreads();
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

}  // namespace

// ------------------------------------------------------------------ launchers

hipError_t launch_bla_build(const double* orbit, double* r, double* r2, double* ab, double* alpha,
                            double* beta, uint32_t* valid, uint64_t steps, uint64_t stride,
                            double eps, double dc_max, int64_t dc_exp, uint32_t max_levels,
                            hipStream_t stream) {
    BlaBuild t{r, r2, ab, alpha, beta, steps, stride, eps, dc_max, dc_exp, valid};
    uint64_t cnt = steps >> 1;
    if (cnt == 0) return hipSuccess;
    // Levels exist while steps >> (lv + 1) > 0; each launch builds up to ten of them.
    uint32_t top = 0;
    while (top < max_levels && (steps >> (top + 1)) > 0) top++;
    for (uint32_t lv0 = 0; lv0 < top; lv0 += kSpanLevels) {
        uint32_t lv_end = lv0 + kSpanLevels < top ? lv0 + kSpanLevels : top;
        uint64_t spans = ((stride >> (lv0 + 1)) + kSpan - 1) / kSpan;
        unsigned grid = (unsigned)(spans < 65536 ? spans : 65536);
        hipLaunchKernelGGL(bla_span, dim3(grid), dim3(kSpanThreads), 0, stream, orbit, t, lv0,
                           lv_end);
    }
    return hipGetLastError();
}

__global__ void zero_words(unsigned long long* p, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

// Everything a single frame's pass needs between its table build and its perturbation
// kernel, in one launch of one wave: the descriptor `d` (a kernel argument) stored at
// `dst` with the table's level count, which comes from the build's per-level flags
// `valid` (levels end before the first level without a valid block, bla.rs:442-445) or, with valid null, from *levels_in (0 when
// both are null); the first `nzero` (<= 64) counter words zeroed; the count also written
// to *levels_out (nullable) for the host's readback; and, with clear, the flags reset to 0
// for the next build.
__global__ void __launch_bounds__(64) pass_prologue(FrameDev d, FrameDev* dst, uint32_t* valid,
                                                    const uint32_t* levels_in, uint64_t steps,
                                                    uint32_t max_levels,
                                                    unsigned long long* counters,
                                                    uint32_t nzero, uint32_t* levels_out,
                                                    bool clear) {
    const uint32_t t = threadIdx.x;
    if (t < nzero) counters[t] = 0ull;
    uint32_t n = 0;
    if (valid) {
        while (n < max_levels && (steps >> (n + 1)) > 0 && valid[n]) n++;
    } else if (levels_in) {
        n = *levels_in;
    }
    const uint64_t* src = reinterpret_cast<const uint64_t*>(&d);
    uint64_t* out = reinterpret_cast<uint64_t*>(dst);
    for (uint32_t i = t; i < sizeof(FrameDev) / 8; i += 64) out[i] = src[i];
    // every lane has read the flags and stored its words before lane 0 patches the count
    // and the flags are cleared
    __syncthreads();
    if (t == 0) {
        dst->bla.nlevels = n;
        if (levels_out) *levels_out = n;
    }
    if (valid && clear)
        for (uint32_t i = t; i < max_levels; i += 64) valid[i] = 0u;
}

static_assert(sizeof(FrameDev) % 8 == 0, "descriptor copied in 8-byte words");

hipError_t launch_pass_prologue(const FrameDev& d, FrameDev* dst, uint32_t* valid,
                                const uint32_t* levels_in, uint64_t steps, uint32_t max_levels,
                                unsigned long long* counters, uint32_t nzero,
                                uint32_t* levels_out, bool clear, hipStream_t stream) {
    if (nzero > 64) return hipErrorInvalidValue;
    hipLaunchKernelGGL(pass_prologue, dim3(1), dim3(64), 0, stream, d, dst, valid, levels_in,
                       steps, max_levels, counters, nzero, levels_out, clear);
    return hipGetLastError();
}

hipError_t launch_zero_words(unsigned long long* p, uint32_t n, hipStream_t stream) {
    hipLaunchKernelGGL(zero_words, dim3((n + 255) / 256), dim3(256), 0, stream, p, n);
    return hipGetLastError();
}

hipError_t launch_perturb_f64_series(const FrameDev* f, bool deriv, bool bla, int grid,
                                     hipStream_t stream) {
#define FDG_LAUNCH_SER(D, L)                                                              \
    hipLaunchKernelGGL((perturb_f64<false, D, L, false, false, true>), dim3(grid),          \
                       dim3(kBlock), 0, stream, f)
    switch ((deriv ? 2 : 0) | (bla ? 1 : 0)) {
        case 0: FDG_LAUNCH_SER(false, false); break;
        case 1: FDG_LAUNCH_SER(false, true); break;
        case 2: FDG_LAUNCH_SER(true, false); break;
        default: FDG_LAUNCH_SER(true, true); break;
    }
#undef FDG_LAUNCH_SER
    return hipGetLastError();
}

hipError_t launch_perturb_f64(const FrameDev* f, bool ext, bool deriv, bool bla, bool batch,
                              int grid, hipStream_t stream, bool list) {
    if (list) {
        // the mixed form's refinement pass: an index list, never a batch
#define FDG_LAUNCH_LIST(E, D, L) \
    hipLaunchKernelGGL((perturb_f64<E, D, L, false, true>), dim3(grid), dim3(kBlock), 0, stream, f)
        int k = (ext ? 4 : 0) | (deriv ? 2 : 0) | (bla ? 1 : 0);
        switch (k) {
            case 0: FDG_LAUNCH_LIST(false, false, false); break;
            case 1: FDG_LAUNCH_LIST(false, false, true); break;
            case 2: FDG_LAUNCH_LIST(false, true, false); break;
            case 3: FDG_LAUNCH_LIST(false, true, true); break;
            case 4: FDG_LAUNCH_LIST(true, false, false); break;
            case 5: FDG_LAUNCH_LIST(true, false, true); break;
            case 6: FDG_LAUNCH_LIST(true, true, false); break;
            default: FDG_LAUNCH_LIST(true, true, true); break;
        }
#undef FDG_LAUNCH_LIST
        return hipGetLastError();
    }
#define FDG_LAUNCH(E, D, L, B) \
    hipLaunchKernelGGL((perturb_f64<E, D, L, B>), dim3(grid), dim3(kBlock), 0, stream, f)
    int key = (batch ? 8 : 0) | (ext ? 4 : 0) | (deriv ? 2 : 0) | (bla ? 1 : 0);
    switch (key) {
        case 0: FDG_LAUNCH(false, false, false, false); break;
        case 1: FDG_LAUNCH(false, false, true, false); break;
        case 2: FDG_LAUNCH(false, true, false, false); break;
        case 3: FDG_LAUNCH(false, true, true, false); break;
        case 4: FDG_LAUNCH(true, false, false, false); break;
        case 5: FDG_LAUNCH(true, false, true, false); break;
        case 6: FDG_LAUNCH(true, true, false, false); break;
        case 7: FDG_LAUNCH(true, true, true, false); break;
        case 8: FDG_LAUNCH(false, false, false, true); break;
        case 9: FDG_LAUNCH(false, false, true, true); break;
        case 10: FDG_LAUNCH(false, true, false, true); break;
        case 11: FDG_LAUNCH(false, true, true, true); break;
        case 12: FDG_LAUNCH(true, false, false, true); break;
        case 13: FDG_LAUNCH(true, false, true, true); break;
        case 14: FDG_LAUNCH(true, true, false, true); break;
        default: FDG_LAUNCH(true, true, true, true); break;
    }
#undef FDG_LAUNCH
    return hipGetLastError();
}

}  // namespace fdg
