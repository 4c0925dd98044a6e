// Device helpers whose bits must equal the host's (and the reference CPU kernel's).
// Compiled with -ffp-contract=off; gfx950's f64 add/mul/div/sqrt are IEEE
// round-to-nearest, so these reproduce the host results exactly.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace fdg {
namespace dm {

// 2^k: exact where representable, 0 below the subnormals, inf above
// (crates/fd-fixed/src/lib.rs:30-42).  ldexp(1, k) is exactly that: it rounds 2^k to
// nearest, which is 2^k itself down to 2^-1074, 0 from 2^-1075 (a tie, to even) down,
// and inf above 2^1023; one v_ldexp_f64 once k is clamped into int range.
__device__ inline double exp2i(int64_t k) {
    int32_t c = k > 4096 ? 4096 : (k < -4096 ? -4096 : (int32_t)k);
    return ldexp(1.0, c);
}

// The same for an int32 exponent: ldexp's exponent operand is 32-bit and it rounds 2^k
// to nearest like exp2i for every k (0 below the subnormals, inf above the range), so
// no clamp is needed; one v_ldexp_f64.
__device__ inline double exp2i32(int32_t k) { return ldexp(1.0, k); }

// floor(log2(a)) for finite a > 0, as the reference computes it (a.log2().floor()).
// Only ever used to pick a power-of-two rescaling, which is exact whichever of two
// adjacent exponents is chosen.
__device__ inline int64_t floor_log2(double a) { return (int64_t)floor(log2(a)); }

// Rust's `x as i64` for a float: saturating, and 0 for NaN, where a C++ cast is undefined.
// sum2 (sample.rs:1097) takes a NaN term's exponent this way: the term then sits at its
// own exponent and drops out of the sum when far below the other term.
__device__ inline int64_t rust_as_i64(double x) {
    if (__builtin_isnan(x)) return 0;
    if (x >= 0x1p63) return INT64_MAX;
    if (x < -0x1p63) return INT64_MIN;
    return (int64_t)x;
}

// i64 + and - as the reference's release build computes them (wrapping, no overflow checks)
__device__ inline int64_t wrap_add(int64_t a, int64_t b) {
    return (int64_t)((uint64_t)a + (uint64_t)b);
}
__device__ inline int64_t wrap_sub(int64_t a, int64_t b) {
    return (int64_t)((uint64_t)a - (uint64_t)b);
}

// The same value without the libm call on all but a sliver of inputs.  a = f 2^e with f in
// [0.5, 1) (subnormals included) has floor(log2(a)) = e - 1 exactly; the rounded log2 can
// only differ where log2(a) = e + log2(f) lies within an ulp of the integer e, i.e. for f
// within about 2^-43 of 1 (|log2(a)| <= 1075).  Those inputs (f > 1 - 2^-40) take the
// reference's expression itself.
__device__ inline int32_t floor_log2_fast(double a) {
    if (__builtin_amdgcn_frexp_mant(a) > 1.0 - 0x1p-40) return (int32_t)floor_log2(a);
    return __builtin_amdgcn_frexp_exp(a) - 1;
}

__device__ inline uint64_t next_pow2(uint64_t x) {
    if (x <= 1) return 1;
    return 1ull << (64 - __builtin_clzll(x - 1));
}

// glibc's hypot (C. F. Borges, "An Improved Algorithm for hypot(a,b)", 2019, unfused
// correction), which Rust's f64::hypot calls on Linux: the BLA radii depend on its bits.
__device__ inline double hypot_kernel(double ax, double ay) {
    double h = sqrt(ax * ax + ay * ay);
    double t1, t2;
    if (h <= 2.0 * ay) {
        double delta = h - ay;
        t1 = ax * (2.0 * delta - ax);
        t2 = (delta - 2.0 * (ax - ay)) * delta;
    } else {
        double delta = h - ax;
        t1 = 2.0 * delta * (ax - 2.0 * ay);
        t2 = (4.0 * delta - ay) * ay + delta * delta;
    }
    h -= (t1 + t2) / (2.0 * h);
    return h;
}

__device__ inline double hypot_ref(double x, double y) {
    if (!isfinite(x) || !isfinite(y)) {
        if (isinf(x) || isinf(y)) return __longlong_as_double(0x7ff0000000000000ll);
        return x + y;
    }
    x = fabs(x);
    y = fabs(y);
    double ax = x < y ? y : x, ay = x < y ? x : y;
    const double kScale = 0x1p-600, kLarge = 0x1p+511, kTiny = 0x1p-511, kEps = 0x1p-54;
    if (ax > kLarge) {
        if (ay <= ax * kEps) return ax + ay;
        return hypot_kernel(ax * kScale, ay * kScale) / kScale;
    }
    if (ay < kTiny) {
        if (ax >= ay / kEps) return ax + ay;
        return hypot_kernel(ax / kScale, ay / kScale) * kScale;
    }
    if (ay <= ax * kEps) return ax + ay;
    return hypot_kernel(ax, ay);
}

}  // namespace dm
}  // namespace fdg
