// Host side of the MI355X deep-zoom path: exact decimal views, the arbitrary-precision
// reference orbit, sample-grid geometry, and the column encoder / .fds writer.
//
// These restate the reference's host-side pieces (crates/fd-fixed, crates/fd-kernel/src/
// {view,reference,store}.rs, crates/fd-samples) with the same floating-point operation
// order, so a frame computed here and on the GPU encodes to the reference's bytes.
#pragma once
#include <cstdint>
#include <string>
#include <thread>
#include <vector>

namespace fdg {

// part(w) for every w in [0, n): parts 1.. on threads of their own, part 0 on the caller.
// A thread that cannot be started (std::system_error, or no memory for its handle) runs
// its part on the caller instead, so every part runs once and every started thread is
// joined — a joinable std::thread is never destroyed.  Parts must not throw.
template <class F>
void run_parts(unsigned n, F part) {
    std::vector<std::thread> th;
    unsigned w = 1;
    try {
        th.reserve(n);
        for (; w < n; w++) th.emplace_back(part, w);
    } catch (...) {
    }
    for (unsigned v = w; v < n; v++) part(v);
    part(0);
    for (auto& t : th) t.join();
}

// 2^k as a double: exact where representable, 0 below the subnormals, inf above
// (crates/fd-fixed/src/lib.rs:30-42).
double exp2i(int64_t k);

// glibc's hypot (An Improved Algorithm for hypot(a,b), C. F. Borges 2019, unfused
// correction), which the reference's f64::hypot resolves to on Linux.  Restated here so
// host and device produce the same bits.
double hypot_ref(double x, double y);

enum Tier : int32_t { TIER_F64 = 0, TIER_FIXED = 1, TIER_SCALED = 2 };

// An exact decimal +-digits * 10^exp10 (crates/fd-fixed/src/decimal.rs:6-50).
struct Decimal {
    bool neg = false;
    std::vector<uint8_t> digits;
    int64_t exp10 = 0;
    static bool parse(const std::string& s, Decimal* out);
    double log2_abs() const;
    // Truncated toward zero to `limbs` little-endian u64 limbs (one integer limb).
    bool to_fixed(size_t limbs, bool* neg_out, std::vector<uint64_t>* mag) const;
};

// Sign-magnitude fixed point, one integer limb (crates/fd-fixed/src/lib.rs:16-23).
struct Fixed {
    bool neg = false;
    std::vector<uint64_t> mag;
    static size_t limbs_for(uint64_t bits) { return 1 + (size_t)((bits + 63) / 64); }
    static bool parse(const std::string& s, size_t limbs, Fixed* out);
    bool frexp(double* m, int64_t* e) const;
    double to_f64() const;
};

// Sample-grid geometry and tier choice (crates/fd-kernel/src/view.rs:18-135).
struct Plane {
    double c_re = 0, c_im = 0, h_m = 1;
    int64_t h_e = 0;
    Tier tier = TIER_F64;
    uint64_t bits = 128;
    double cos_r = 1, sin_r = 0, half_x = 0, half_y = 0;
    static bool make(const std::string& re, const std::string& im, const std::string& width,
                     double rotation, uint32_t nx, uint32_t ny, Plane* out, std::string* err);
    double h() const { return h_m * exp2i(h_e); }
    void unit_offset(uint32_t i, uint32_t j, double* ux, double* uy) const;
    std::string kernel(Tier t) const;
};

// Reference orbit Z_0 = 0 .. Z_{len-1}, structure of arrays in f64
// (crates/fd-kernel/src/reference.rs:5-96).
struct Orbit {
    std::vector<double> re, im;
    uint32_t bits = 53;
    // The fixed-point orbit's state at point cycle_at equalled the one cycle_period points
    // earlier (an exact repeat: a nucleus centre to the working precision, or an attracting
    // cycle the truncated orbit settled on), so the points after it are copies (0: none).
    uint64_t cycle_at = 0, cycle_period = 0;
    static constexpr uint64_t kMaxLen = 1ull << 20;
    // f64 orbit at an f64 centre, stopping at the escape radius.
    static Orbit f64(double c_re, double c_im, uint64_t max_iter, double radius,
                     uint64_t max_len = kMaxLen);
    // Exact-centre fixed-point orbit (3 truncated squarings per step), stopping at
    // |Z|^2 > 2^32.
    static bool fixed(const std::string& re, const std::string& im, uint64_t bits,
                      uint64_t max_iter, Orbit* out, uint64_t max_len = kMaxLen);
    uint64_t len() const { return re.size(); }
    // Appends the copies of the points after an exact repeat, up to point `steps`.
    void tile(uint64_t steps);
    // Reference::nucleus_period (reference.rs:102-120); 0 for None.
    uint64_t nucleus_period() const { return nucleus_period_of(re.data(), im.data(), len()); }
    static uint64_t nucleus_period_of(const double* re, const double* im, uint64_t len);
};

// Newton on the period-P nucleus equation z_P(c) = 0 from c (decimal strings) at `bits`
// fraction bits (zone_build.rs:215-233's iteration): z in the orbit's truncated fixed point,
// the derivative z_P' in floating point with a carried exponent, c -= z_P / z_P' until
// the step is below 2^-(bits - 8) (converged) or max_iters steps have run.
struct Nucleus {
    std::string re, im;      // the last c, as decimals to the working precision
    unsigned iterations = 0;
    bool converged = false;
    double log2_step = 0;    // log2 |z_P / z_P'| of the last step
};
bool nucleus_newton(const std::string& re, const std::string& im, uint64_t period, uint64_t bits,
                    unsigned max_iters, Nucleus* out, std::string* err);

// Reference::error_radius (reference.rs:122-148): q[m] >= |A_m - Z_m|; bits < 0 for an f64
// orbit from (c_re, c_im), else the fixed-point fraction bits.
void error_radius(const double* re, const double* im, uint64_t len, int64_t bits, double c_re,
                  double c_im, double* q);

// Upper bound m 2^e on |dc| over a view's samples (grid.rs:423-444).
void dc_bound(const Plane& pl, Tier tier, uint32_t nx, uint32_t ny, double* dc, int64_t* e);

// Guard thresholds and order schedule of the parameter series (series.rs:28-55 SeriesConfig,
// its Default).
struct SeriesConfig {
    uint32_t start_order = 8, max_order = 64;
    double grow_at = 1e-4, shift_px = 1e-8, derivative_rel = 1e-6;
};

// Why a series build stopped (series.rs:57-82 `stop`).
enum SeriesStop : int32_t { STOP_REFERENCE = 0, STOP_RADIUS, STOP_NONFINITE, STOP_ESCAPE,
                            STOP_BUDGET, STOP_STALL };

// An admitted series start (series.rs:57-82 Series): every sample with |dc| <= rho starts
// at iterate n with delta S_n(dc / rho) = sum b[k] (dc / rho)^k.
struct Series {
    uint64_t n = 0;
    double rho = 0;
    uint64_t points = 0;
    std::vector<double> b;  // b[1..] as (re, im) pairs
    double state_bound = 0, derivative_bound = 0, shift_px = 0;
    uint64_t built_steps = 0, coefficient_pairs = 0, growth_steps = 0;
    int32_t stop = STOP_REFERENCE;
    // Series::build_stalled (series.rs:99-234) over an orbit of `len` points, for samples
    // with |dc| <= rho and output pixel `pixel`; stalled_steps 32 is Series::build.
    static Series build(const double* re, const double* im, uint64_t len, double rho,
                        double pixel, uint64_t max_iter, const SeriesConfig& cfg,
                        uint64_t stalled_steps = 32);
};

// Raw per-sample result of the kernel (sample.rs:21-40).
enum Kind : uint8_t { KIND_ESCAPED = 0, KIND_INTERIOR = 1, KIND_UNRESOLVED = 2 };

// Column bits (crates/fd-samples/src/column.rs:9-23).
enum : uint32_t { COL_CLASS = 1, COL_NU = 2, COL_DE = 4, COL_NORMAL = 8, COL_BOUND = 16,
                  COL_STRIPE = 32 };

// Store::put for heuristic evidence (store.rs:66-112): one escaped/interior/unresolved
// outcome into the class/nu/de/normal columns.
struct Encoder {
    Plane pl;
    double px_m;
    int64_t px_e;
    Encoder(const Plane& p, uint32_t ss) : pl(p), px_m(p.h_m * (double)ss), px_e(p.h_e) {}
    void put(size_t idx, uint8_t kind, uint64_t n, double zr, double zi, double dr, double di,
             int64_t dexp, uint8_t* cls, double* nu, float* de, uint16_t* normal) const;
};

// Jet-AA (crates/fd-cli/src/jetaa.rs) on an ss 2 sample grid of w x h pixels (2w x 2h
// samples, row-major).  jetaa_first: the pass-1 mask, each pixel's top-left sample
// (jetaa.rs:38-44).  jetaa_rebuild: predicts the other three samples of every pixel from
// the pass-1 columns, writing the accepted ones into the columns, and sets todo (4wh
// bytes) for the rejected ones (jetaa.rs:123-245).  jetaa_merge: copies the samples
// selected by mask from the second pass's columns (jetaa.rs:248-261).
void jetaa_first(size_t w, size_t h, uint8_t* mask);
void jetaa_rebuild(uint8_t* cls, double* nu, float* de, uint16_t* normal, size_t w, size_t h,
                   uint8_t* todo);
void jetaa_merge(uint8_t* cls, double* nu, float* de, uint16_t* normal, const uint8_t* fcls,
                 const double* fnu, const float* fde, const uint16_t* fnormal,
                 const uint8_t* mask, size_t n);

// fd_samples::write (write.rs:8-57, header.rs:311-330): the canonical .fds bytes.
std::vector<uint8_t> write_fds(const std::string& re, const std::string& im,
                               const std::string& width, double rotation, uint32_t nx,
                               uint32_t ny, uint32_t ss, uint64_t max_iter, double escape_radius,
                               uint32_t columns, const std::string& kernel, const uint8_t* cls,
                               const double* nu, const float* de, const uint16_t* normal);

}  // namespace fdg
