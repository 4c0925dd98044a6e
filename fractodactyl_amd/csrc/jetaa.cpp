// Jet-AA for ss 2 film frames (crates/fd-cli/src/jetaa.rs): pass 1 computes each pixel's
// top-left sample, the rebuild predicts the other three from the pass-1 columns, pass 2
// computes the predictions it rejected.
//
// F = 2^-n log z_n is holomorphic in the screen coordinate s = dx - i dy (output pixels),
// and a sample's columns give F up to one scale: Re F ~ 2^-nu, |F'| = 2 Re F / de, arg F'
// the normal.  The four neighbouring pixels' F' give F'' and F''' (a stencil whose h^3
// and h^4 error terms cancel); from F, F', F'', F''' at a pixel's sample, nu, de and the
// normal follow at its other three positions.  A prediction is kept only when the one
// from the neighbouring pixel's jet on its other side agrees (|dnu| < 3e-3, |d ln de| <
// 3e-2, normals within 9e-3 rad), and then is their mean.  Every expression keeps the
// reference's operation order, so the rebuilt columns are its bytes.
#include <cmath>
#include <cstring>
#include <memory>
#include <thread>
#include <vector>

#include "host.hpp"

namespace fdg {
namespace {

constexpr double kM = 3.0;       // consistency tolerance multiplier (jetaa.rs:32)
constexpr float kDeFloor = 0.3f;  // no prediction nearer the set than this (jetaa.rs:36)

struct Cx {
    double re, im;
};
inline Cx add(Cx a, Cx b) { return {a.re + b.re, a.im + b.im}; }
inline Cx sub(Cx a, Cx b) { return {a.re - b.re, a.im - b.im}; }
inline Cx mul(Cx a, Cx b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
inline Cx scale(Cx a, double k) { return {a.re * k, a.im * k}; }

// Powers (d, d^2/2, d^3/6) of the screen offset d = dx - i dy (jetaa.rs:63-76).
struct Off {
    Cx d, d2h, d3s;
    Off(double dx, double dy) {
        d = {dx, -dy};
        Cx d2 = mul(d, d);
        d2h = scale(d2, 0.5);
        d3s = scale(mul(d2, d), 1.0 / 6.0);
    }
};

// A pixel's jet: its sample's nu and F', F'', F''' with Re F scaled to 1 (jetaa.rs:78-103).
struct Jet {
    bool ok;
    double nu;
    Cx f1, f2, f3;
    // (nu, de, normal x, normal y) at the offset o; false where the expansion fails
    bool at(const Off& o, double* nu_o, float* de_o, double* nx, double* ny) const {
        double re = 1.0 + mul(f1, o.d).re + mul(f2, o.d2h).re + mul(f3, o.d3s).re;
        Cx fp = add(add(f1, mul(f2, o.d)), mul(f3, o.d2h));
        double de = 2.0 * re / std::sqrt(fp.re * fp.re + fp.im * fp.im);
        double nu2 = nu - std::log2(re);
        if (!(re > 0.0 && std::isfinite(nu2) && std::isfinite(de) && de > 0.0)) return false;
        *nu_o = nu2;
        *de_o = (float)de;
        *nx = fp.re;
        *ny = fp.im;
        return true;
    }
};

// Samples::unit (crates/fd-samples/src/samples.rs:53-58): the stored angle as an f32
// unit vector.
inline void unit(uint16_t angle, float* c, float* s) {
    float a = (float)angle * (6.28318530717958647692f / 65536.0f);
    *s = std::sin(a);
    *c = std::cos(a);
}

// Samples::angle (samples.rs:60-65)
inline uint16_t angle_of(double x, double y) {
    double t = std::round(std::atan2(y, x) * (65536.0 / 6.283185307179586));
    if (std::isnan(t)) return 0;
    int64_t v = t >= 9.2233720368547758e18 ? INT64_MAX
                : t <= -9.2233720368547758e18 ? INT64_MIN : (int64_t)t;
    return (uint16_t)(uint64_t)v;
}

// f(i) for i in [0, n) on all cores; the result does not depend on the split.
template <class F>
void par_for(size_t n, F f) {
    unsigned threads = std::thread::hardware_concurrency();
    if (threads < 1) threads = 1;
    size_t chunk = (n + threads - 1) / threads;
    if (n < 4096 || threads == 1) {
        for (size_t i = 0; i < n; i++) f(i);
        return;
    }
    run_parts(threads, [&f, chunk, n](unsigned t) {
        size_t b = t * chunk, e = std::min(n, b + chunk);
        for (size_t i = b; i < e; i++) f(i);
    });
}

// A predicted sample (jetaa.rs:119-121): none (compute it), the pixel's own non-escaped
// kind, or an escaped sample's columns.
struct Pred {
    uint8_t what = 0;  // 0 compute, 1 own kind, 2 escaped
    double nu = 0;
    float de = 0;
    uint16_t normal = 0;
};

}  // namespace

void jetaa_first(size_t w, size_t h, uint8_t* mask) {
    for (size_t i = 0; i < 4 * w * h; i++)
        mask[i] = (i % (2 * w)) % 2 == 0 && (i / (2 * w)) % 2 == 0;
}

void jetaa_rebuild(uint8_t* cls, double* nu, float* de, uint16_t* normal, size_t w, size_t h,
                   uint8_t* todo) {
    const size_t nx = 2 * w, n = w * h;
    auto idx = [nx](size_t px, size_t py, size_t k) {
        return (2 * py + k / 2) * nx + 2 * px + k % 2;
    };
    auto real = [&](size_t q) { return idx(q % w, q / w, 0); };
    auto kind = [&](size_t i) { return (uint8_t)(cls[i] & 3); };  // 3: invalid encoding
    // per pixel sample: (nu, 2 F' / Re F) = (nu, 2 e^(i normal) / de), escaped only
    // Per-pixel scratch that each pixel's own iteration writes before any read (the host
    // frames are large: 8.3M pixels at 3840x2160, so nothing is zero-filled on one thread)
    std::unique_ptr<uint8_t[]> g_ok(new uint8_t[n]);
    std::unique_ptr<double[]> g_nu(new double[n]);
    std::unique_ptr<Cx[]> g_c(new Cx[n]);
    par_for(n, [&](size_t q) {
        size_t i = real(q);
        g_ok[q] = kind(i) == KIND_ESCAPED;
        if (!g_ok[q]) return;
        float x, y;
        unit(normal[i], &x, &y);
        g_nu[q] = nu[i];
        g_c[q] = scale(Cx{(double)x, (double)y}, 2.0 / (double)de[i]);
    });
    std::unique_ptr<Jet[]> jets(new Jet[n]);
    par_for(n, [&](size_t q) {
        size_t px = q % w, py = q / w;
        jets[q].ok = false;
        if (px == 0 || py == 0 || px + 1 >= w || py + 1 >= h || !g_ok[q]) return;
        const double nu0 = g_nu[q];
        // F' of a neighbour relative to Re F here: Re F scales as 2^-nu
        auto at = [&](size_t r, Cx* out) {
            if (!g_ok[r]) return false;
            *out = scale(g_c[r], std::exp2(nu0 - g_nu[r]));
            return true;
        };
        Cx rt, lf, up, dn;
        if (!at(q + 1, &rt) || !at(q - 1, &lf) || !at(q - w, &up) || !at(q + w, &dn)) return;
        Cx a = sub(rt, lf), b = sub(up, dn);
        Jet j;
        j.nu = nu0;
        j.f1 = g_c[q];
        j.f2 = scale(Cx{a.re + b.im, a.im - b.re}, 0.25);
        j.f3 = scale(sub(sub(add(rt, lf), up), dn), 0.5);
        j.ok = std::isfinite(j.f2.re) && std::isfinite(j.f2.im) && std::isfinite(j.f3.re) &&
               std::isfinite(j.f3.im);
        jets[q] = j;
    });
    // offsets of samples 1..3 from the pixel's own sample and from the right, lower and
    // diagonal neighbour's
    const Off own[3] = {Off(0.5, 0.0), Off(0.0, 0.5), Off(0.5, 0.5)};
    const Off other[3] = {Off(-0.5, 0.0), Off(0.0, -0.5), Off(-0.5, -0.5)};
    const double max_dnu = kM * 1e-3, max_dlnde = kM * 1e-2, max_cross = std::sin(kM * 3e-3);
    // A pixel's predictions for its samples 1..3 (jetaa.rs:123-245): none (compute it),
    // the pixel's own non-escaped kind, or an escaped sample's columns.
    auto predict = [&](size_t q, size_t px, size_t py, Pred* pred) {
        if (px == 0 || py == 0 || px + 2 >= w || py + 2 >= h) return;
        size_t r0 = real(q);
        if (!g_ok[q]) {
            // inside: the sample's own kind, only where every neighbour has it too
            uint8_t k0 = kind(r0);
            const size_t nbs[5] = {q + 1, q + w, q + w + 1, q - 1, q - w};
            for (size_t r : nbs)
                if (kind(real(r)) != k0) return;
            for (int k = 0; k < 3; k++) pred[k].what = 1;
            return;
        }
        if (de[r0] < kDeFloor) return;
        const Jet& j = jets[q];
        if (!j.ok) return;
        const size_t nb[3] = {q + 1, q + w, q + w + 1};
        for (int k = 0; k < 3; k++) {
            double anu, ax, ay, bnu, bx, by;
            float ade, bde;
            if (!j.at(own[k], &anu, &ade, &ax, &ay)) continue;
            const Jet& jn = jets[nb[k]];
            if (!jn.ok || !jn.at(other[k], &bnu, &bde, &bx, &by)) continue;
            if (std::fabs(anu - bnu) >= max_dnu ||
                std::fabs(std::log((double)(ade / bde))) >= max_dlnde)
                continue;
            double na = std::sqrt(ax * ax + ay * ay), nb_ = std::sqrt(bx * bx + by * by);
            Cx ua{ax / na, ay / na}, ub{bx / nb_, by / nb_};
            double cross = ua.re * ub.im - ua.im * ub.re, dot = ua.re * ub.re + ua.im * ub.im;
            if (dot > 0.0 && std::fabs(cross) < max_cross) {
                pred[k].what = 2;
                pred[k].nu = 0.5 * (anu + bnu);
                pred[k].de = 0.5f * (ade + bde);
                pred[k].normal = angle_of(ua.re + ub.re, ua.im + ub.im);
            }
        }
    };
    // Each pixel predicts its samples 1..3 and writes them and their todo entries itself:
    // the reads are of pixels' sample 0 and of the jets, the writes of samples 1..3 (and
    // todo), so no pixel reads what another writes and the order does not matter.
    par_for(n, [&](size_t q) {
        size_t px = q % w, py = q / w;
        Pred pred[3] = {};
        predict(q, px, py, pred);
        const uint8_t inside = cls[real(q)];
        todo[idx(px, py, 0)] = 0;
        for (int k = 0; k < 3; k++) {
            size_t i = idx(px, py, k + 1);
            const Pred& p = pred[k];
            todo[i] = p.what == 0;
            if (p.what == 1) {
                // Class::new(kind, Heuristic), or the invalid class as it is
                cls[i] = (inside & 3) != 3 ? (uint8_t)(inside & 3) : inside;
            } else if (p.what == 2) {
                cls[i] = KIND_ESCAPED;  // escaped, heuristic evidence
                if (nu) nu[i] = p.nu;
                if (de) de[i] = p.de;
                if (normal) normal[i] = p.normal;
            }
        }
    });
}

void jetaa_merge(uint8_t* cls, double* nu, float* de, uint16_t* normal, const uint8_t* fcls,
                 const double* fnu, const float* fde, const uint16_t* fnormal,
                 const uint8_t* mask, size_t n) {
    for (size_t i = 0; i < n; i++) {
        if (!mask[i]) continue;
        cls[i] = fcls[i];
        if (nu && fnu) nu[i] = fnu[i];
        if (de && fde) de[i] = fde[i];
        if (normal && fnormal) normal[i] = fnormal[i];
    }
}

}  // namespace fdg
