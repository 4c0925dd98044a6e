// The reference GPU route's certified error control on the host: the reference orbit's
// error radii (Reference::error_radius), the guarded BLA table with derivative envelopes
// (GuardedBla::build), the admission of an externally computed escaped state
// (external_escape_ok) and its overlay through the column encoder (overlay_escaped).  Each
// keeps the reference's floating-point operation order; hypot is glibc's algorithm
// (hypot_ref), the logarithms and arcsines the C library's, as the reference's f64 methods
// resolve to on Linux.
#include <cfloat>
#include <cmath>
#include <cstring>
#include <vector>

#include "abi_guard.hpp"
#include "fdg.h"
#include "host.hpp"

namespace fdg {
namespace {

constexpr double kU = DBL_EPSILON / 2.0;
constexpr double kLn2 = 0.693147180559945309417232121458176568;
constexpr double kTau = 6.283185307179586476925286766559005768;
constexpr double kPi = 3.141592653589793238462643383279502884;

struct GBlock {
    double ar = 0, ai = 0, br = 0, bi = 0, r = 0, alpha = 0, beta = 0, dalpha = 0, dbeta = 0;
};

// crates/fd-kernel/src/bla/guarded.rs:9-18
double l1(double re, double im) { return std::fabs(re) + std::fabs(im); }
double up(double x) { return x * (1.0 + 64.0 * DBL_EPSILON) + DBL_MIN; }

// guarded.rs:34-50 GuardedBlock::single over Block::single (bla.rs:72-85)
GBlock single(double zr, double zi, double q, double eps) {
    GBlock g;
    double az = hypot_ref(zr, zi);
    double r = 2.0 * eps * az;
    if (!(r > 0.0 && az <= 2.0)) return g;  // Block::single's invalid block
    if (!std::isfinite(q) || q < 0.0 || l1(zr, zi) + q > 2.0) return g;
    g.ar = 2.0 * zr + 0.0;
    g.ai = 2.0 * zi + 0.0;
    g.br = 1.0;
    g.bi = 0.0;
    g.r = r;
    g.beta = 0.0;
    g.dalpha = up(2.0 * r + 2.0 * q);
    g.alpha = up(r + 2.0 * q);
    g.dbeta = 0.0;
    return g;
}

// guarded.rs:51-97 GuardedBlock::merge: x then y, for |dc| <= c
GBlock merge(const GBlock& x, const GBlock& y, double c) {
    GBlock z;
    if (x.r == 0.0 || y.r == 0.0) return z;
    const double ax = l1(x.ar, x.ai), ay = l1(y.ar, y.ai), bx = l1(x.br, x.bi);
    const double r = std::fmin(x.r, (y.r - up((bx + x.beta) * c)) / up(ax + x.alpha)) *
                     (1.0 - 64.0 * DBL_EPSILON);
    GBlock m;
    // affine(x.a, y.a, 0) and affine(x.b, y.a, y.b) (bla.rs:124-129)
    m.ar = y.ar * x.ar - y.ai * x.ai + 0.0 + 0.0;
    m.ai = y.ar * x.ai + y.ai * x.ar + 0.0 + 0.0;
    m.br = y.ar * x.br - y.ai * x.bi + y.br + 0.0;
    m.bi = y.ar * x.bi + y.ai * x.br + y.bi + 0.0;
    const double ra = up(32.0 * DBL_EPSILON * ay * ax);
    const double rb = up(32.0 * DBL_EPSILON * (ay * bx + l1(y.br, y.bi)));
    // guarded.rs:13-18 remainder, for the state and the derivative envelopes
    const double alpha = up(ay * x.alpha + y.alpha * (ax + x.alpha));
    const double beta = up(ay * x.beta + y.alpha * (bx + x.beta) + y.beta);
    const double da = up(ay * x.dalpha + y.dalpha * (ax + x.dalpha));
    const double db = up(ay * x.dbeta + y.dalpha * (bx + x.dbeta) + y.dbeta);
    m.r = r;
    m.alpha = up(alpha + ra);
    m.beta = up(beta + rb);
    m.dalpha = up(da + ra);
    m.dbeta = up(db + rb);
    if (r > 0.0 && std::isfinite(m.ar) && std::isfinite(m.ai) && std::isfinite(m.br) &&
        std::isfinite(m.bi) && std::isfinite(m.r) && std::isfinite(m.alpha) &&
        std::isfinite(m.beta) && std::isfinite(m.dalpha) && std::isfinite(m.dbeta))
        return m;
    return z;
}

// store.rs:119-139 bounds: false unless the radii prove escape (|z| - ez > 2, ed < |D|/2)
bool bounds(uint64_t n, double az, double ez, double ad, double ed, double* lo_out,
            double* bound_out) {
    const double lo = az - ez;
    if (!(lo > 2.0 && ed < 0.5 * ad)) return false;
    const double lnz = std::log(az), llo = std::log(lo), lhi = std::log(az + ez);
    const double nu = std::log1p(-std::log1p(-ez / az) / llo) / kLn2 +
                      8.0 * kU * ((double)n + 8.0);
    const double g = lhi * std::exp2(-(double)n);
    const double de_lo = (lo * llo) / (az * lnz) * ad / (ad + ed) * std::exp(-g) *
                         (1.0 - 4.0 / (lo * lo)) * (1.0 - 1e-6);
    const double de_hi = ((az + ez) * lhi) / (az * lnz) * ad / (ad - ed) * (1.0 + 1e-6);
    *lo_out = de_lo;
    *bound_out = nu * de_hi / de_lo * (1.0 + 1e-6);
    return true;
}

// store.rs:141-150 de_px: 2 |z| ln|z| / |dz/dc| in output pixels, exponent-safe
float de_px(double az, double log2z, double dr, double di, int64_t dexp, double px_m,
            int64_t px_e) {
    const double lnz = log2z * kLn2;
    const double base = 2.0 * az * lnz / (hypot_ref(dr, di) * px_m);
    int64_t k = -(dexp + px_e);
    k = k < -4000 ? -4000 : (k > 4000 ? 4000 : k);
    return (float)(base * exp2i(k / 2) * exp2i(k - k / 2));
}

}  // namespace

// reference.rs:122-148 Reference::error_radius
void error_radius(const double* re, const double* im, uint64_t len, int64_t bits, double c_re,
                  double c_im, double* q) {
    const double c1 = std::fabs(c_re) + std::fabs(c_im);
    double fixed = 0.0;
    if (bits >= 0) fixed = std::fmax(exp2i(4 - bits), DBL_MIN);
    double a = 0.0;  // |A_m - F_m|, F the orbit before its rounding to f64
    for (uint64_t m = 0; m < len; m++) {
        const double s = std::sqrt(re[m] * re[m] + im[m] * im[m]) * (1.0 + 4.0 * kU);
        double round, defect;
        if (bits >= 0) {
            // three truncated squares per component and the centre's truncation
            round = 2.0 * kU * s + DBL_MIN;
            defect = fixed;
        } else {
            // rounding of x^2 - y^2 + c and 2xy + c, plus the centre's rounding
            round = 0.0;
            defect = 8.0 * kU * (s * s + c1) + 2.0 * kU * c1 + DBL_MIN;
        }
        q[m] = a + round;
        a = (a * (2.0 * (s + round) + a) + defect) * (1.0 + 8.0 * kU);
    }
}

// store.rs:155-198 external_escape_ok
bool external_escape_ok(uint64_t n, double zr, double zi, double dr, double di, int64_t dexp,
                        double ez, double ed, double px_m, int64_t px_e) {
    const double az = hypot_ref(zr, zi), ad = hypot_ref(dr, di);
    if (!(ez >= 0.0 && ed >= 0.0 && std::isfinite(ez) && std::isfinite(ed) &&
          std::isfinite(az) && std::isfinite(ad)))
        return false;
    double lo, bound;
    if (!bounds(n, az, ez, ad, ed, &lo, &bound)) return false;
    const double de = (double)de_px(az, std::log2(az), dr, di, dexp, px_m, px_e);
    const double displacement = bound * de * lo * kLn2 / 2.0;
    // relative uncertainty of |z| ln|z| / |dz/dc|, the encoded DE
    const double base = az * std::log(az);
    const double de_hi = (az + ez) * std::log(az + ez) / base * ad / (ad - ed);
    const double de_lo = (az - ez) * std::log(az - ez) / base * ad / (ad + ed);
    const double angle = std::asin(ez / az) + std::asin(ed / ad);
    return std::isfinite(displacement) && displacement <= 5e-4 && std::isfinite(de) &&
           de_hi - 1.0 + (double)FLT_EPSILON <= 1e-3 &&
           1.0 - de_lo + (double)FLT_EPSILON <= 1e-3 &&
           angle + kTau / 65536.0 <= 0.1 * (kPi / 180.0);
}

}  // namespace fdg

using namespace fdg;

extern "C" {

int fdg_error_radius(const double* re, const double* im, uint64_t len, int64_t bits,
                     double c_re, double c_im, double* q) {
    return guarded<int>(nullptr, [&]() -> int {
        if (!re || !im || !q) return FDG_EINVAL;
        error_radius(re, im, len, bits, c_re, c_im, q);
        return FDG_OK;
    });
}

int fdg_guarded_bla(const double* re, const double* im, const double* q, uint64_t len, double eps,
                    double dc_max, double* blocks, uint64_t cap_blocks, uint32_t* levels) {
    return guarded<int>(nullptr, [&]() -> int {
        if (!re || !im || !q) return FDG_EINVAL;
        // bla.rs:27-33 check, and guarded.rs:111-118: one finite nonnegative radius each
        if (!(eps > 0.0 && eps <= 1.0 / 16777216.0 && dc_max > 0.0 && std::isfinite(dc_max))) {
            g_create_error = "BLA needs 0 < eps <= 2^-24 and a finite dc_max > 0";
            return FDG_EINVAL;
        }
        for (uint64_t k = 0; k < len; k++)
            if (!std::isfinite(q[k]) || q[k] < 0.0) {
                g_create_error = "guarded BLA needs one finite nonnegative reference radius per "
                                 "point";
                return FDG_EINVAL;
            }
        uint64_t n = len >= 2 ? len - 2 : 0;
        std::vector<GBlock> next(n);
        for (uint64_t k = 1; k <= n; k++) next[k - 1] = single(re[k], im[k], q[k], eps);
        uint32_t nl = 0;
        uint64_t w = 0, need = 0;
        for (;;) {
            const uint64_t m = n / 2;
            bool any = false;
            for (uint64_t i = 0; i < m; i++) {
                next[i] = merge(next[2 * i], next[2 * i + 1], dc_max);
                any = any || next[i].r > 0.0;
            }
            n = m;
            if (!any) break;  // the table ends before the first level without a valid block
            need += n;
            for (uint64_t i = 0; i < n && blocks; i++, w++) {
                if (w >= cap_blocks) continue;
                const GBlock& g = next[i];
                double* o = blocks + 9 * w;
                o[0] = g.ar, o[1] = g.ai, o[2] = g.br, o[3] = g.bi, o[4] = g.r;
                o[5] = g.alpha, o[6] = g.beta, o[7] = g.dalpha, o[8] = g.dbeta;
            }
            nl++;
        }
        if (levels) *levels = nl;
        if (blocks && need > cap_blocks) {
            g_create_error = "block buffer too small";
            return FDG_EINVAL;
        }
        return FDG_OK;
    });
}

int fdg_external_escape_ok(const fdg_escaped_state* s, double ez, double ed, double px_m,
                           int64_t px_e) {
    return guarded<int>(nullptr, [&]() -> int {
        if (!s) return FDG_EINVAL;
        return external_escape_ok(s->n, s->zr, s->zi, s->dr, s->di, s->dexp, ez, ed, px_m, px_e)
                   ? 1
                   : 0;
    });
}

int fdg_overlay_escaped(const fdg_view* v, const fdg_params* p, fdg_columns* out,
                        const fdg_escaped_state* states, const uint8_t* present) {
    return guarded<int>(nullptr, [&]() -> int {
        // grid.rs:123-171 overlay_escaped
        if (!v || !p || !out || !out->class_ || !states || !present) return FDG_EINVAL;
        if (p->columns & (FDG_COL_BOUND | FDG_COL_STRIPE)) {
            g_create_error =
                "escaped overlay needs a matching sample grid without bound or stripe columns";
            return FDG_EINVAL;
        }
        Plane pl;
        std::string e;
        if (p->ss == 0 || p->nx % p->ss != 0 || p->ny % p->ss != 0 ||
            !Plane::make(v->center_re, v->center_im, v->width, v->rotation, p->nx, p->ny, &pl,
                         &e)) {
            g_create_error = e.empty() ? "grid must be a whole number of pixels" : e;
            return FDG_EINVAL;
        }
        const size_t count = (size_t)p->nx * p->ny;
        for (size_t i = 0; i < count; i++) {
            if (!present[i]) continue;
            const fdg_escaped_state& s = states[i];
            if (s.n > p->max_iter || !std::isfinite(s.zr) || !std::isfinite(s.zi) ||
                hypot_ref(s.zr, s.zi) <= p->escape_radius || !std::isfinite(s.dr) ||
                !std::isfinite(s.di) || hypot_ref(s.dr, s.di) == 0.0) {
                g_create_error = "escaped overlay received an invalid state";
                return FDG_EINVAL;
            }
        }
        Encoder enc(pl, p->ss);
        double* nu = (p->columns & FDG_COL_NU) ? out->nu : nullptr;
        float* de = (p->columns & FDG_COL_DE) ? out->de : nullptr;
        uint16_t* nm = (p->columns & FDG_COL_NORMAL) ? out->normal : nullptr;
        for (size_t i = 0; i < count; i++) {
            if (!present[i]) continue;
            const fdg_escaped_state& s = states[i];
            enc.put(i, KIND_ESCAPED, s.n, s.zr, s.zi, s.dr, s.di, s.dexp, out->class_, nu, de,
                    nm);
        }
        return FDG_OK;
    });
}

}  // extern "C"
