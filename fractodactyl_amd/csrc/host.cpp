// Host side of the deep-zoom path; see host.hpp.  Built with -ffp-contract=off: every
// floating-point expression keeps the reference's operation order.
#include "host.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <thread>
#include <utility>

namespace fdg {

using u128 = unsigned __int128;

double exp2i(int64_t k) {
    uint64_t b;
    if (k > 1023) return INFINITY;
    if (k >= -1022)
        b = (uint64_t)(k + 1023) << 52;
    else if (k >= -1074)
        b = 1ull << (k + 1074);
    else
        return 0.0;
    double d;
    std::memcpy(&d, &b, 8);
    return d;
}

static int64_t biased_exponent(double v) {
    uint64_t b;
    std::memcpy(&b, &v, 8);
    return (int64_t)((b >> 52) & 0x7ff) - 1023;
}

// Borges' unfused correction step on ax >= ay > 0 in range.
static double hypot_kernel(double ax, double ay) {
    double h = std::sqrt(ax * ax + ay * ay);
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

double hypot_ref(double x, double y) {
    if (!std::isfinite(x) || !std::isfinite(y)) {
        if (std::isinf(x) || std::isinf(y)) return INFINITY;
        return x + y;
    }
    x = std::fabs(x);
    y = std::fabs(y);
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

// ------------------------------------------------------------------ limbs

namespace limbs {
static int cmp(const uint64_t* a, const uint64_t* b, size_t l) {
    for (size_t i = l; i-- > 0;)
        if (a[i] != b[i]) return a[i] < b[i] ? -1 : 1;
    return 0;
}
static void add(const uint64_t* a, const uint64_t* b, uint64_t* o, size_t l) {
    uint64_t carry = 0;
    for (size_t i = 0; i < l; i++) {
        u128 s = (u128)a[i] + b[i] + carry;
        o[i] = (uint64_t)s;
        carry = (uint64_t)(s >> 64);
    }
}
static void sub(const uint64_t* a, const uint64_t* b, uint64_t* o, size_t l) {
    uint64_t borrow = 0;
    for (size_t i = 0; i < l; i++) {
        u128 d = (u128)a[i] - b[i] - borrow;
        o[i] = (uint64_t)d;
        borrow = (uint64_t)(d >> 64) & 1;
    }
}
static bool nonzero(const uint64_t* a, size_t l) {
    for (size_t i = 0; i < l; i++)
        if (a[i]) return true;
    return false;
}
// Signed sum; returns true when the result is negative (zero is never negative).
static bool add_signed(bool an, const uint64_t* a, bool bn, const uint64_t* b, uint64_t* o,
                       size_t l) {
    if (an == bn) {
        add(a, b, o, l);
        return an && nonzero(o, l);
    }
    if (cmp(a, b, l) < 0) {
        sub(b, a, o, l);
        return bn;
    }
    sub(a, b, o, l);
    return an && nonzero(o, l);
}
// rp[0 .. n) += up[0 .. n) * v, returning the carry limb.  Four limbs per block with two
// independent carry chains (adcx: the previous product's high limb, adox: rp) and mulx,
// which leaves the flags alone (BMI2 + ADX, every x86-64 server CPU since 2014 and the
// ones this library runs on); the plain loop finishes the tail and serves CPUs without.
static bool g_adx = __builtin_cpu_supports("bmi2") && __builtin_cpu_supports("adx");

static uint64_t addmul_1(uint64_t* rp, const uint64_t* up, size_t n, uint64_t v) {
    uint64_t carry = 0;
    size_t i = 0;
    if (g_adx) {
        for (; i + 4 <= n; i += 4) {
            asm("xorl %%r10d, %%r10d\n\t"
                "mulx 0(%[up]), %%r8, %%r9\n\t"
                "adcx %[c], %%r8\n\t"
                "adox 0(%[rp]), %%r8\n\t"
                "movq %%r8, 0(%[rp])\n\t"
                "mulx 8(%[up]), %%r8, %[c]\n\t"
                "adcx %%r9, %%r8\n\t"
                "adox 8(%[rp]), %%r8\n\t"
                "movq %%r8, 8(%[rp])\n\t"
                "mulx 16(%[up]), %%r8, %%r9\n\t"
                "adcx %[c], %%r8\n\t"
                "adox 16(%[rp]), %%r8\n\t"
                "movq %%r8, 16(%[rp])\n\t"
                "mulx 24(%[up]), %%r8, %[c]\n\t"
                "adcx %%r9, %%r8\n\t"
                "adox 24(%[rp]), %%r8\n\t"
                "movq %%r8, 24(%[rp])\n\t"
                "adcx %%r10, %[c]\n\t"
                "adox %%r10, %[c]"
                : [c] "+&r"(carry)
                : [up] "r"(up + i), [rp] "r"(rp + i), "d"(v)
                : "r8", "r9", "r10", "cc", "memory");
        }
    }
    for (; i < n; i++) {
        u128 t = (u128)up[i] * v + rp[i] + carry;
        rp[i] = (uint64_t)t;
        carry = (uint64_t)(t >> 64);
    }
    return carry;
}

// Exact schoolbook square into prod[0 .. 2l): the off-diagonal products once (a row per
// limb), doubled, then the diagonal.
static void sqr_base(const uint64_t* a, size_t l, uint64_t* prod) {
    std::memset(prod, 0, 2 * l * sizeof(uint64_t));
    for (size_t i = 0; i + 1 < l; i++) {
        if (!a[i]) continue;
        prod[i + l] = addmul_1(prod + 2 * i + 1, a + i + 1, l - i - 1, a[i]);
    }
    uint64_t top = 0;
    for (size_t i = 0; i < 2 * l; i++) {
        uint64_t nt = prod[i] >> 63;
        prod[i] = (prod[i] << 1) | top;
        top = nt;
    }
    uint64_t carry = 0;
    for (size_t i = 0; i < l; i++) {
        u128 sq = (u128)a[i] * a[i];
        u128 t = (u128)prod[2 * i] + (uint64_t)sq + carry;
        prod[2 * i] = (uint64_t)t;
        u128 t2 = (u128)prod[2 * i + 1] + (uint64_t)(sq >> 64) + (uint64_t)(t >> 64);
        prod[2 * i + 1] = (uint64_t)t2;
        carry = (uint64_t)(t2 >> 64);
    }
}

// Below this many limbs the schoolbook square is faster than splitting.
constexpr size_t kKaratsuba = 40;

// Orbits of at least this many limbs (about 5,000 bits) square on three threads: below,
// a square (0.4 us at 56 limbs on the EPYC hosts of this GPU) costs less than the
// barrier and the exchange of the squares between cores.
constexpr size_t kOrbitThreadLimbs = 80;

// Scratch limbs sqr_full needs for an l-limb square.
static size_t sqr_scratch(size_t l) {
    if (l < kKaratsuba) return 0;
    size_t hh = l - l / 2;
    return 5 * hh + 2 + sqr_scratch(hh);
}

// Exact square into prod[0 .. 2l) by Karatsuba: with a = a1 B^h + a0,
// a^2 = a1^2 B^2h + (a0^2 + a1^2 - (a1 - a0)^2) B^h + a0^2 — three half-size squares.
// The full product is exact, so its truncation is the schoolbook one bit for bit.
static void sqr_full(const uint64_t* a, size_t l, uint64_t* prod, uint64_t* scratch) {
    if (l < kKaratsuba) {
        sqr_base(a, l, prod);
        return;
    }
    const size_t h = l / 2, hh = l - h;  // a0: h limbs, a1: hh limbs (hh >= h)
    uint64_t* d = scratch;               // |a1 - a0|, hh limbs
    uint64_t* t = d + hh;                // d^2, 2 hh limbs
    uint64_t* m = t + 2 * hh;            // a0^2 + a1^2 - d^2, 2 hh + 1 limbs
    uint64_t* deeper = m + 2 * hh + 1;
    sqr_full(a, h, prod, deeper);            // a0^2 -> prod[0 .. 2h)
    sqr_full(a + h, hh, prod + 2 * h, deeper);  // a1^2 -> prod[2h .. 2l)
    // d = |a1 - a0| with a0 zero-extended to hh limbs
    int c = 0;
    for (size_t i = hh; i-- > 0;) {
        uint64_t x = a[h + i], y = i < h ? a[i] : 0;
        if (x != y) {
            c = x > y ? 1 : -1;
            break;
        }
    }
    const uint64_t* big = c >= 0 ? a + h : a;
    const uint64_t* small = c >= 0 ? a : a + h;
    size_t big_n = c >= 0 ? hh : h, small_n = c >= 0 ? h : hh;
    uint64_t borrow = 0;
    for (size_t i = 0; i < hh; i++) {
        uint64_t x = i < big_n ? big[i] : 0, y = i < small_n ? small[i] : 0;
        u128 r = (u128)x - y - borrow;
        d[i] = (uint64_t)r;
        borrow = (uint64_t)(r >> 64) & 1;
    }
    sqr_full(d, hh, t, deeper);
    // m = a0^2 + a1^2 - d^2 (= 2 a0 a1 >= 0)
    uint64_t carry = 0;
    for (size_t i = 0; i < 2 * hh; i++) {
        uint64_t x = i < 2 * h ? prod[i] : 0;
        u128 r = (u128)x + prod[2 * h + i] + carry;
        m[i] = (uint64_t)r;
        carry = (uint64_t)(r >> 64);
    }
    m[2 * hh] = carry;
    borrow = 0;
    for (size_t i = 0; i <= 2 * hh; i++) {
        uint64_t y = i < 2 * hh ? t[i] : 0;
        u128 r = (u128)m[i] - y - borrow;
        m[i] = (uint64_t)r;
        borrow = (uint64_t)(r >> 64) & 1;
    }
    // prod += m B^h
    carry = 0;
    for (size_t i = 0; h + i < 2 * l; i++) {
        uint64_t y = i <= 2 * hh ? m[i] : 0;
        if (i > 2 * hh && !carry) break;
        u128 r = (u128)prod[h + i] + y + carry;
        prod[h + i] = (uint64_t)r;
        carry = (uint64_t)(r >> 64);
    }
}

// The high half of a square, exactly: floor(a^2 / B^(l-1)) (B = 2^64) from the products
// a[i] a[j] with i + j >= m = l - 3 only — about half of them.  The omitted products sum
// to L < 2m B^(m+1) (column k holds k + 1 ordered products, each below B^(k+2)), so with
// H the sum of the kept ones, floor((H + L) / B^(l-1)) = floor(H / B^(l-1)) whenever
// (H mod B^(l-1)) + 2m B^(m+1) < B^(l-1), i.e. H's limb l-2 is below B - 2m - 1 (its
// limbs below m are zero).  Returns false, having computed nothing usable, in the other
// case (probability ~2m / 2^64), where the caller takes the full square.
static bool sqr_high(const uint64_t* a, size_t l, uint64_t* prod) {
    if (l < 8) return false;
    const size_t m = l - 3;
    std::memset(prod + m, 0, (2 * l - m) * sizeof(uint64_t));
    for (size_t i = 0; i + 1 < l; i++) {
        size_t j0 = std::max(i + 1, m > i ? m - i : (size_t)0);
        if (j0 >= l || !a[i]) continue;
        prod[i + l] = addmul_1(prod + i + j0, a + j0, l - j0, a[i]);
    }
    // double the off-diagonal sum (zero below column m), then add a[i]^2 for 2i >= m
    uint64_t top = 0;
    for (size_t k = m; k < 2 * l; k++) {
        uint64_t nt = prod[k] >> 63;
        prod[k] = (prod[k] << 1) | top;
        top = nt;
    }
    uint64_t carry = 0;
    for (size_t i = (m + 1) / 2; i < l; i++) {
        u128 sq = (u128)a[i] * a[i];
        u128 t = (u128)prod[2 * i] + (uint64_t)sq + carry;
        prod[2 * i] = (uint64_t)t;
        u128 t2 = (u128)prod[2 * i + 1] + (uint64_t)(sq >> 64) + (uint64_t)(t >> 64);
        prod[2 * i + 1] = (uint64_t)t2;
        carry = (uint64_t)(t2 >> 64);
    }
    return prod[l - 2] < ~(uint64_t)0 - 2 * (uint64_t)m - 1;
}

// o = floor(a^2) at 64(l-1) fraction bits: the exact 2l-limb square truncated to the l
// limbs from l-1 up (crates/fd-fixed/src/limbs.rs:66-101), from its high half when that
// is provably the same.  prod: 2l limbs; scratch: sqr_scratch(l) limbs.
static void sqr(const uint64_t* a, uint64_t* prod, uint64_t* o, size_t l, uint64_t* scratch) {
    if (!sqr_high(a, l, prod)) sqr_full(a, l, prod, scratch);
    std::memcpy(o, prod + (l - 1), l * sizeof(uint64_t));
}
}  // namespace limbs

static bool frexp_limbs(bool neg, const uint64_t* mag, size_t l, double* m, int64_t* e) {
    size_t k = l;
    while (k > 0 && mag[k - 1] == 0) k--;
    if (k == 0) return false;
    k--;
    double below1 = k >= 1 ? (double)mag[k - 1] : 0.0;
    double below2 = k >= 2 ? (double)mag[k - 2] : 0.0;
    double v = (double)mag[k] + below1 * exp2i(-64) + below2 * exp2i(-128);
    int64_t ve = biased_exponent(v);
    double mm = v * exp2i(-ve);
    *m = neg ? -mm : mm;
    *e = ve + 64 * ((int64_t)k - ((int64_t)l - 1));
    return true;
}

static double to_f64_limbs(bool neg, const uint64_t* mag, size_t l) {
    double m;
    int64_t e;
    if (!frexp_limbs(neg, mag, l, &m, &e)) return 0.0;
    if (e < -1000) return m * exp2i(e + 100) * exp2i(-100);
    return m * exp2i(e);
}

bool Fixed::frexp(double* m, int64_t* e) const { return frexp_limbs(neg, mag.data(), mag.size(), m, e); }
double Fixed::to_f64() const { return to_f64_limbs(neg, mag.data(), mag.size()); }

// ------------------------------------------------------------------ decimal

static bool is_ws(char c) { return c == ' ' || (c >= '\t' && c <= '\r'); }

bool Decimal::parse(const std::string& s, Decimal* out) {
    size_t b = 0, e = s.size();
    while (b < e && is_ws(s[b])) b++;
    while (e > b && is_ws(s[e - 1])) e--;
    std::string t = s.substr(b, e - b);
    bool neg = false;
    if (!t.empty() && (t[0] == '-' || t[0] == '+')) {
        neg = t[0] == '-';
        t.erase(0, 1);
    }
    size_t epos = t.find_first_of("eE");
    int64_t exp = 0;
    std::string mant = t;
    if (epos != std::string::npos) {
        std::string es = t.substr(epos + 1);
        mant = t.substr(0, epos);
        // Rust's i64 parse: optional sign, at least one digit, no overflow.
        size_t i = 0;
        bool eneg = false;
        if (!es.empty() && (es[0] == '+' || es[0] == '-')) {
            eneg = es[0] == '-';
            i = 1;
        }
        if (i == es.size()) return false;
        u128 v = 0;
        for (; i < es.size(); i++) {
            if (es[i] < '0' || es[i] > '9') return false;
            v = v * 10 + (unsigned)(es[i] - '0');
            if (v > ((u128)1 << 63)) return false;
        }
        if (!eneg && v > (u128)INT64_MAX) return false;
        exp = eneg ? (int64_t)(0 - (uint64_t)v) : (int64_t)v;
    }
    size_t dot = mant.find('.');
    std::string ip = dot == std::string::npos ? mant : mant.substr(0, dot);
    std::string fp = dot == std::string::npos ? std::string() : mant.substr(dot + 1);
    Decimal d;
    for (char c : ip + fp) {
        uint8_t v = (uint8_t)(c - '0');
        if (v > 9) return false;
        d.digits.push_back(v);
    }
    if (d.digits.empty()) return false;
    bool zero = true;
    for (uint8_t v : d.digits) zero = zero && v == 0;
    if (exp < INT64_MIN + (int64_t)fp.size()) return false;
    d.neg = neg && !zero;
    d.exp10 = exp - (int64_t)fp.size();
    *out = std::move(d);
    return true;
}

double Decimal::log2_abs() const {
    size_t first = 0;
    while (first < digits.size() && digits[first] == 0) first++;
    if (first == digits.size()) return -INFINITY;
    size_t sig = digits.size() - first;
    double head = 0.0;
    for (size_t i = 0; i < sig && i < 17; i++) head = head * 10.0 + (double)digits[first + i];
    double rest = sig > 17 ? (double)(sig - 17) : 0.0;
    return (std::log10(head) + rest + (double)exp10) * 3.32192809488736234787031942948939018;
}

static void mul_small(std::vector<uint64_t>& v, uint64_t m) {
    uint64_t c = 0;
    for (auto& x : v) {
        u128 t = (u128)x * m + c;
        x = (uint64_t)t;
        c = (uint64_t)(t >> 64);
    }
    if (c) v.push_back(c);
}
static void add_small(std::vector<uint64_t>& v, uint64_t a) {
    uint64_t c = a;
    for (auto& x : v) {
        uint64_t s = x + c;
        c = s < x;
        x = s;
        if (!c) return;
    }
    if (c) v.push_back(c);
}
static void div_small(std::vector<uint64_t>& v, uint64_t d) {
    u128 r = 0;
    for (size_t i = v.size(); i-- > 0;) {
        u128 cur = (r << 64) | v[i];
        v[i] = (uint64_t)(cur / d);
        r = cur % d;
    }
}
static uint64_t pow10(unsigned k) {
    uint64_t p = 1;
    while (k--) p *= 10;
    return p;
}

bool Decimal::to_fixed(size_t limbs, bool* neg_out, std::vector<uint64_t>* mag) const {
    std::vector<uint64_t> n{0};
    for (size_t c = 0; c < digits.size(); c += 19) {
        size_t len = std::min<size_t>(19, digits.size() - c);
        uint64_t val = 0;
        for (size_t i = 0; i < len; i++) val = val * 10 + digits[c + i];
        mul_small(n, pow10((unsigned)len));
        add_small(n, val);
    }
    int64_t e = exp10;
    while (e > 0) {
        unsigned k = (unsigned)std::min<int64_t>(e, 19);
        mul_small(n, pow10(k));
        e -= k;
        if (n.size() > 1) return false;
    }
    std::vector<uint64_t> v(limbs - 1, 0);
    v.insert(v.end(), n.begin(), n.end());
    while (e < 0) {
        unsigned k = (unsigned)std::min<int64_t>(-e, 19);
        div_small(v, pow10(k));
        e += k;
        if (!limbs::nonzero(v.data(), v.size())) break;
    }
    if (v.size() > limbs && limbs::nonzero(v.data() + limbs, v.size() - limbs)) return false;
    v.resize(limbs);
    *neg_out = neg && limbs::nonzero(v.data(), limbs);
    *mag = std::move(v);
    return true;
}

bool Fixed::parse(const std::string& s, size_t limbs, Fixed* out) {
    Decimal d;
    if (!Decimal::parse(s, &d)) return false;
    return d.to_fixed(limbs, &out->neg, &out->mag);
}

// ------------------------------------------------------------------ plane

bool Plane::make(const std::string& re, const std::string& im, const std::string& width,
                 double rotation, uint32_t nx, uint32_t ny, Plane* p, std::string* err) {
    if (nx == 0 || ny == 0) {
        *err = "empty sample grid";
        return false;
    }
    Decimal w;
    if (!Decimal::parse(width, &w)) {
        *err = "not a decimal number: \"" + width + "\"";
        return false;
    }
    double lw = w.log2_abs();
    if (w.neg || !std::isfinite(lw)) {
        *err = "view width must be positive";
        return false;
    }
    Fixed wf;
    if (!w.to_fixed(Fixed::limbs_for((uint64_t)std::ceil(std::fmax(64.0 - lw, 64.0))), &wf.neg,
                    &wf.mag)) {
        *err = "value too large for fixed point";
        return false;
    }
    double wm;
    int64_t we;
    if (!wf.frexp(&wm, &we)) {
        *err = "view width underflowed";
        return false;
    }
    double q = wm / (double)nx;
    int64_t qe = biased_exponent(q);
    p->h_m = q * exp2i(-qe);
    p->h_e = we + qe;
    Decimal tmp;
    if (!Decimal::parse(re, &tmp) || !Decimal::parse(im, &tmp)) {
        *err = "not a decimal number";
        return false;
    }
    p->c_re = std::strtod(re.c_str(), nullptr);
    p->c_im = std::strtod(im.c_str(), nullptr);
    if (!(std::isfinite(p->c_re) && std::isfinite(p->c_im))) {
        *err = "view centre out of range";
        return false;
    }
    double c_mag = std::fmax(std::fmax(std::fabs(p->c_re), std::fabs(p->c_im)), exp2i(-1022));
    bool f64_ok = std::log2(c_mag) - 52.0 <= (double)p->h_e - 10.0 && p->h_e > -1000;
    p->tier = f64_ok ? TIER_F64 : (p->h_e >= -900 ? TIER_FIXED : TIER_SCALED);
    p->bits = (uint64_t)std::max<int64_t>(128 - p->h_e, 128);
    p->sin_r = std::sin(rotation);
    p->cos_r = std::cos(rotation);
    p->half_x = (double)nx / 2.0;
    p->half_y = (double)ny / 2.0;
    return true;
}

void Plane::unit_offset(uint32_t i, uint32_t j, double* ux, double* uy) const {
    double x = (double)i + 0.5 - half_x;
    double y = (double)j + 0.5 - half_y;
    *ux = cos_r * x + sin_r * y;
    *uy = sin_r * x - cos_r * y;
}

std::string Plane::kernel(Tier t) const {
    switch (t) {
        case TIER_F64: return "pert-f64/1";
        case TIER_FIXED: return "pert-fx/1 bits=" + std::to_string(bits);
        default: return "pert-fx-scaled/1 bits=" + std::to_string(bits);
    }
}

// ------------------------------------------------------------------ orbit

Orbit Orbit::f64(double c_re, double c_im, uint64_t max_iter, double radius, uint64_t max_len) {
    Orbit o;
    uint64_t steps = std::min<uint64_t>(max_iter, max_len - 1);
    o.re.reserve(std::min<uint64_t>(steps, 1 << 16) + 1);
    o.im.reserve(std::min<uint64_t>(steps, 1 << 16) + 1);
    o.re.push_back(0.0);
    o.im.push_back(0.0);
    double x = 0.0, y = 0.0;
    for (uint64_t k = 0; k < steps; k++) {
        double t = x * x - y * y + c_re;
        y = 2.0 * x * y + c_im;
        x = t;
        o.re.push_back(x);
        o.im.push_back(y);
        if (x * x + y * y > radius * radius) break;
    }
    o.bits = 53;
    return o;
}

namespace {

// A reusable barrier for the orbit's three threads: the last to arrive advances the
// phase, the others spin on it (a step is a few microseconds, far below a futex wake).
struct SpinBarrier {
    std::atomic<uint32_t> arrived{0}, phase{0};
    const uint32_t n;
    explicit SpinBarrier(uint32_t count) : n(count) {}
    void wait() {
        uint32_t ph = phase.load(std::memory_order_acquire);
        if (arrived.fetch_add(1, std::memory_order_acq_rel) + 1 == n) {
            arrived.store(0, std::memory_order_relaxed);
            phase.store(ph + 1, std::memory_order_release);
            return;
        }
        while (phase.load(std::memory_order_acquire) == ph) __builtin_ia32_pause();
    }
};

// Exact repeats of the orbit's fixed-point state (x, y).  The truncated fixed-point map is
// deterministic, so once a state recurs every later point repeats with the same period, and
// the rest of the orbit is a copy of earlier points — bit for bit what the squarings would
// give.  Two watches: Z_i == Z_1 = c, which a nucleus centre reaches at i = P + 1 once Z_P
// squares to zero at the working precision, and Brent's state saved at each power-of-two
// index, which finds any other cycle (an attracting cycle the truncated orbit settles on)
// by the index max(tail, period) rounded up to a power of two, plus the period.  Signs are
// canonical (zero is non-negative), so equal values have equal limbs.
struct CycleWatch {
    size_t l;
    const uint64_t *cr, *ci;
    bool crn, cin;
    uint64_t* saved;                   // 2l limbs: x, then y
    bool sxn = false, syn = false;
    uint64_t save_at = 0, next_save = 1;
    CycleWatch(size_t limbs, const Fixed& re, const Fixed& im, uint64_t* buf)
        : l(limbs), cr(re.mag.data()), ci(im.mag.data()), saved(buf) {
        crn = re.neg && std::any_of(cr, cr + l, [](uint64_t v) { return v != 0; });
        cin = im.neg && std::any_of(ci, ci + l, [](uint64_t v) { return v != 0; });
        std::fill(saved, saved + 2 * l, 0);   // Z_0 = 0
    }
    // Z_i's state (i >= 1); the period when Z_i equals an earlier point, else 0.
    uint64_t check(uint64_t i, bool xn, const uint64_t* x, bool yn, const uint64_t* y) {
        if (i >= 2 && xn == crn && yn == cin && std::equal(x, x + l, cr) &&
            std::equal(y, y + l, ci))
            return i - 1;
        if (xn == sxn && yn == syn && std::equal(x, x + l, saved) &&
            std::equal(y, y + l, saved + l))
            return i - save_at;
        if (i == next_save) {
            std::copy(x, x + l, saved);
            std::copy(y, y + l, saved + l);
            sxn = xn;
            syn = yn;
            save_at = i;
            next_save = 2 * i;
        }
        return 0;
    }
};

// One thread's share of the fixed-point orbit (orbit.rs:14-43: x' = x^2 - y^2 + cr,
// y' = (x + y)^2 - (x^2 + y^2) + ci).  Thread `role` squares x (0), y (1) or x + y (2)
// into its slot of the step's buffer; after the barrier every thread combines the three
// squares into its own copy of (x, y), so one barrier per step suffices (the squares are
// double-buffered by step parity, and a thread can only write a buffer again after all
// have passed the next barrier, by which time they have read it).  With roles {0, 1, 2}
// on one thread this is the serial loop.
// Every buffer is allocated by the calling thread before a worker starts, so no worker
// allocates (an exception in a worker thread would end the process), and the workers
// wait at a start gate: when a second thread cannot be started the first is told to
// leave before it reaches the step barrier, and the caller runs the serial loop.
struct OrbitShared {
    size_t l;
    uint64_t steps;
    const Fixed *cr, *ci;
    std::vector<uint64_t> sq;       // [parity][role] squares, l limbs each
    std::vector<uint64_t> buf[3];   // each worker's registers and scratch, by first role
    std::atomic<int> gate{0};       // 0 wait, 1 run, 2 leave
    SpinBarrier bar;
    Orbit* out;
    uint64_t cycle_at = 0, cycle_period = 0;  // set by the emitting worker
    OrbitShared(size_t limbs, uint32_t threads) : l(limbs), sq(6 * limbs), bar(threads) {
        for (uint32_t w = 0; w < threads; w++)
            buf[w].assign(8 * limbs + limbs::sqr_scratch(limbs), 0);
    }
};

void orbit_worker(OrbitShared* sh, int first_role, int last_role, bool emit) {
    const size_t l = sh->l;
    const bool threaded = !(first_role == 0 && last_role == 2);
    if (threaded) {
        int g;
        while ((g = sh->gate.load(std::memory_order_acquire)) == 0) __builtin_ia32_pause();
        if (g == 2) return;
    }
    std::vector<uint64_t>& buf = sh->buf[first_role];
    uint64_t *x = &buf[0], *y = &buf[l], *s = &buf[2 * l], *t = &buf[3 * l],
             *prod = &buf[4 * l], *scratch = &buf[8 * l];
    CycleWatch watch(l, *sh->cr, *sh->ci, &buf[6 * l]);
    bool xn = false, yn = false;
    for (uint64_t k = 0; k < sh->steps; k++) {
        uint64_t* sq = &sh->sq[(k & 1) * 3 * l];
        uint64_t *x2 = sq, *y2 = sq + l, *s2 = sq + 2 * l;
        for (int role = first_role; role <= last_role; role++) {
            if (role == 0) {
                limbs::sqr(x, prod, x2, l, scratch);
            } else if (role == 1) {
                limbs::sqr(y, prod, y2, l, scratch);
            } else {
                limbs::add_signed(xn, x, yn, y, s, l);  // its sign is squared away
                limbs::sqr(s, prod, s2, l, scratch);
            }
        }
        if (threaded) sh->bar.wait();
        bool tn = limbs::add_signed(false, x2, true, y2, t, l);
        xn = limbs::add_signed(tn, t, sh->cr->neg, sh->cr->mag.data(), x, l);
        limbs::add(x2, y2, s, l);
        tn = limbs::add_signed(false, s2, true, s, t, l);
        yn = limbs::add_signed(tn, t, sh->ci->neg, sh->ci->mag.data(), y, l);
        double fx = to_f64_limbs(xn, x, l), fy = to_f64_limbs(yn, y, l);
        if (emit) {
            sh->out->re.push_back(fx);
            sh->out->im.push_back(fy);
        }
        // every thread holds the same (x, y), so all stop at the same step
        if (fx * fx + fy * fy > 4294967296.0) break;
        if (uint64_t p = watch.check(k + 1, xn, x, yn, y)) {
            if (emit) {
                sh->cycle_at = k + 1;
                sh->cycle_period = p;
            }
            break;
        }
    }
}

// The same loop with the limb count a compile-time constant, for the shallow tiers (up to
// 16 limbs, ~960 bits: the f64-to-1e-280 frames), where the per-step overhead of the
// general loop (calls, length tests, the high-half check) outweighs the squares
// themselves; the fixed loops unroll.  Same operations, same bits.
template <size_t L>
struct Small {
    static void sqr(const uint64_t* a, uint64_t* o) {
        uint64_t prod[2 * L] = {};
        for (size_t i = 0; i + 1 < L; i++) {
            uint64_t carry = 0;
            for (size_t j = i + 1; j < L; j++) {
                u128 t = (u128)a[i] * a[j] + prod[i + j] + carry;
                prod[i + j] = (uint64_t)t;
                carry = (uint64_t)(t >> 64);
            }
            prod[i + L] = carry;
        }
        uint64_t top = 0;
        for (size_t i = 0; i < 2 * L; i++) {
            uint64_t nt = prod[i] >> 63;
            prod[i] = (prod[i] << 1) | top;
            top = nt;
        }
        uint64_t carry = 0;
        for (size_t i = 0; i < L; i++) {
            u128 sq = (u128)a[i] * a[i];
            u128 t = (u128)prod[2 * i] + (uint64_t)sq + carry;
            prod[2 * i] = (uint64_t)t;
            u128 t2 = (u128)prod[2 * i + 1] + (uint64_t)(sq >> 64) + (uint64_t)(t >> 64);
            prod[2 * i + 1] = (uint64_t)t2;
            carry = (uint64_t)(t2 >> 64);
        }
        for (size_t i = 0; i < L; i++) o[i] = prod[L - 1 + i];
    }
    static void add(const uint64_t* a, const uint64_t* b, uint64_t* o) {
        uint64_t carry = 0;
        for (size_t i = 0; i < L; i++) {
            u128 t = (u128)a[i] + b[i] + carry;
            o[i] = (uint64_t)t;
            carry = (uint64_t)(t >> 64);
        }
    }
    static void sub(const uint64_t* a, const uint64_t* b, uint64_t* o) {
        uint64_t borrow = 0;
        for (size_t i = 0; i < L; i++) {
            u128 d = (u128)a[i] - b[i] - borrow;
            o[i] = (uint64_t)d;
            borrow = (uint64_t)(d >> 64) & 1;
        }
    }
    static bool nonzero(const uint64_t* a) {
        uint64_t any = 0;
        for (size_t i = 0; i < L; i++) any |= a[i];
        return any != 0;
    }
    static bool less(const uint64_t* a, const uint64_t* b) {
        for (size_t i = L; i-- > 0;)
            if (a[i] != b[i]) return a[i] < b[i];
        return false;
    }
    static bool add_signed(bool an, const uint64_t* a, bool bn, const uint64_t* b, uint64_t* o) {
        if (an == bn) {
            add(a, b, o);
            return an && nonzero(o);
        }
        if (less(a, b)) {
            sub(b, a, o);
            return bn;
        }
        sub(a, b, o);
        return an && nonzero(o);
    }
    static void orbit(const Fixed& cr, const Fixed& ci, uint64_t steps, Orbit* out) {
        uint64_t x[L] = {}, y[L] = {}, x2[L], y2[L], s[L], s2[L], t[L], saved[2 * L];
        CycleWatch watch(L, cr, ci, saved);
        bool xn = false, yn = false;
        for (uint64_t k = 0; k < steps; k++) {
            sqr(x, x2);
            sqr(y, y2);
            add_signed(xn, x, yn, y, s);
            sqr(s, s2);
            bool tn = add_signed(false, x2, true, y2, t);
            xn = add_signed(tn, t, cr.neg, cr.mag.data(), x);
            add(x2, y2, s);
            tn = add_signed(false, s2, true, s, t);
            yn = add_signed(tn, t, ci.neg, ci.mag.data(), y);
            double fx = to_f64_limbs(xn, x, L), fy = to_f64_limbs(yn, y, L);
            out->re.push_back(fx);
            out->im.push_back(fy);
            if (fx * fx + fy * fy > 4294967296.0) break;
            if (uint64_t p = watch.check(k + 1, xn, x, yn, y)) {
                out->cycle_at = k + 1;
                out->cycle_period = p;
                break;
            }
        }
    }
};

template <size_t... Ls>
bool orbit_small(size_t l, const Fixed& cr, const Fixed& ci, uint64_t steps, Orbit* out,
                 std::index_sequence<Ls...>) {
    bool done = false;
    ((l == Ls + 3 ? (Small<Ls + 3>::orbit(cr, ci, steps, out), done = true) : false), ...);
    return done;
}

}  // namespace

bool Orbit::fixed(const std::string& re, const std::string& im, uint64_t bits, uint64_t max_iter,
                  Orbit* out, uint64_t max_len) {
    size_t l = Fixed::limbs_for(bits);
    Fixed cr, ci;
    if (!Fixed::parse(re, l, &cr) || !Fixed::parse(im, l, &ci)) return false;
    uint64_t steps = std::min<uint64_t>(max_iter, max_len - 1);
    Orbit o;
    o.re.reserve(steps + 1);
    o.im.reserve(steps + 1);
    o.re.push_back(0.0);
    o.im.push_back(0.0);
    if (orbit_small(l, cr, ci, steps, &o, std::make_index_sequence<14>())) {  // 3..16 limbs
        o.bits = (uint32_t)bits;
        o.tile(steps);
        *out = std::move(o);
        return true;
    }
    // Three threads from kOrbitThreadLimbs limbs up, where a square outweighs a barrier.
    bool threaded = l >= limbs::kOrbitThreadLimbs && std::thread::hardware_concurrency() >= 3;
    if (threaded) {
        OrbitShared sh(l, 3);
        sh.steps = steps;
        sh.cr = &cr;
        sh.ci = &ci;
        sh.out = &o;
        std::thread t1, t2;
        try {
            t1 = std::thread(orbit_worker, &sh, 1, 1, false);
            t2 = std::thread(orbit_worker, &sh, 2, 2, false);
        } catch (...) {
            // (std::system_error: no thread for it) the started one leaves at the gate
            sh.gate.store(2, std::memory_order_release);
            if (t1.joinable()) t1.join();
            threaded = false;
        }
        if (threaded) {
            sh.gate.store(1, std::memory_order_release);
            orbit_worker(&sh, 0, 0, true);
            t1.join();
            t2.join();
            o.cycle_at = sh.cycle_at;
            o.cycle_period = sh.cycle_period;
        }
    }
    if (!threaded) {
        OrbitShared sh(l, 1);
        sh.steps = steps;
        sh.cr = &cr;
        sh.ci = &ci;
        sh.out = &o;
        orbit_worker(&sh, 0, 2, true);
        o.cycle_at = sh.cycle_at;
        o.cycle_period = sh.cycle_period;
    }
    o.bits = (uint32_t)bits;
    o.tile(steps);
    *out = std::move(o);
    return true;
}

void Orbit::tile(uint64_t steps) {
    if (!cycle_period) return;
    // Z_j = Z_{j - p} for every j past the repeat (cycle_at): the copy the squarings would
    // have produced, up to the same last point
    for (uint64_t j = cycle_at + 1; j <= steps; j++) {
        re.push_back(re[j - cycle_period]);
        im.push_back(im[j - cycle_period]);
    }
}

// ------------------------------------------------------------------ nucleus

namespace {

// A complex value (re, im) 2^e with the larger part near 1: the Newton derivative, which
// outgrows f64 within a few hundred steps of a deep period.
struct CxExp {
    double re, im;
    int64_t e;
    void renorm() {
        double a = std::fmax(std::fabs(re), std::fabs(im));
        if (!(a > 0.0) || !std::isfinite(a)) return;
        int k;
        std::frexp(a, &k);
        re = std::ldexp(re, -k);
        im = std::ldexp(im, -k);
        e += k;
    }
};

// v 2^ve (v finite) as a fixed-point magnitude of l limbs (64 (l - 1) fraction bits): v's
// 53-bit integer mantissa placed at its bit position, the bits below the last limb dropped;
// false when the value reaches 2^64.  Zero (also after truncation) is non-negative.
bool fixed_from_scaled(double v, int64_t ve, size_t l, bool* neg, std::vector<uint64_t>* mag) {
    mag->assign(l, 0);
    *neg = false;
    if (v == 0.0) return true;
    int k;
    double f = std::frexp(std::fabs(v), &k);                  // |v| = f 2^k, f in [0.5, 1)
    uint64_t m = (uint64_t)std::ldexp(f, 53);                  // exact 53-bit integer
    int64_t pos = ve + k - 53 + 64 * ((int64_t)l - 1);         // bit position of m's lsb
    if (pos + 53 > 64 * (int64_t)l) return false;
    if (pos <= -64) return true;
    if (pos < 0) {
        m >>= -pos;
        pos = 0;
    }
    size_t limb = (size_t)(pos / 64), bit = (size_t)(pos % 64);
    (*mag)[limb] |= m << bit;
    if (bit && limb + 1 < l) (*mag)[limb + 1] |= m >> (64 - bit);
    *neg = v < 0.0 && limbs::nonzero(mag->data(), l);
    return true;
}

// A fixed-point value as a decimal with `digits` fraction digits (truncated).
std::string fixed_to_decimal(bool neg, const uint64_t* mag, size_t l, size_t digits) {
    std::string out = neg ? "-" : "";
    out += std::to_string(mag[l - 1]);
    out += '.';
    std::vector<uint64_t> frac(mag, mag + (l - 1));
    while (digits > 0) {
        unsigned k = (unsigned)std::min<size_t>(digits, 19);
        uint64_t p = pow10(k), carry = 0;
        for (auto& x : frac) {
            u128 t = (u128)x * p + carry;
            x = (uint64_t)t;
            carry = (uint64_t)(t >> 64);
        }
        std::string chunk = std::to_string(carry);
        out += std::string(k - chunk.size(), '0') + chunk;
        digits -= k;
    }
    return out;
}

}  // namespace

bool nucleus_newton(const std::string& re, const std::string& im, uint64_t period, uint64_t bits,
                    unsigned max_iters, Nucleus* out, std::string* err) {
    if (period < 1 || bits < 64) {
        *err = "the nucleus search needs a period >= 1 and at least 64 bits";
        return false;
    }
    const size_t l = Fixed::limbs_for(bits);
    Fixed c[2];
    if (!Fixed::parse(re, l, &c[0]) || !Fixed::parse(im, l, &c[1])) {
        *err = "bad decimal for the nucleus search's start";
        return false;
    }
    std::vector<uint64_t> buf(8 * l + limbs::sqr_scratch(l));
    uint64_t *x = &buf[0], *y = &buf[l], *s = &buf[2 * l], *t = &buf[3 * l],
             *prod = &buf[4 * l], *x2 = &buf[6 * l], *y2 = &buf[7 * l],
             *scratch = &buf[8 * l];
    std::vector<uint64_t> s2(l);
    const double limit = -((double)bits - 8.0);  // converged: |step| < 2^-(bits - 8)
    out->iterations = 0;
    out->converged = false;
    out->log2_step = INFINITY;
    for (unsigned it = 0; it < max_iters; it++) {
        // P steps of z' = z^2 + c in the orbit's truncated fixed point (orbit.rs:14-43) and
        // dz' = 2 z dz + 1 in floating point with a carried exponent, from z = dz = 0
        std::fill(x, x + l, 0);
        std::fill(y, y + l, 0);
        bool xn = false, yn = false;
        CxExp dz{0.0, 0.0, 0};
        for (uint64_t k = 0; k < period; k++) {
            double zr = to_f64_limbs(xn, x, l), zi = to_f64_limbs(yn, y, l);
            if (zr * zr + zi * zi > 4294967296.0) {
                *err = "the nucleus search's orbit escaped: the start is too far from a "
                       "nucleus of this period";
                return false;
            }
            CxExp p{2.0 * (zr * dz.re - zi * dz.im), 2.0 * (zr * dz.im + zi * dz.re), dz.e};
            if (p.e > 100) {
                dz = p;  // (the + 1 lies below the mantissa)
            } else if (p.e < -100) {
                dz = CxExp{1.0, 0.0, 0};
            } else {
                dz = CxExp{std::ldexp(p.re, (int)p.e) + 1.0, std::ldexp(p.im, (int)p.e), 0};
            }
            dz.renorm();
            limbs::sqr(x, prod, x2, l, scratch);
            limbs::sqr(y, prod, y2, l, scratch);
            limbs::add_signed(xn, x, yn, y, s, l);
            limbs::sqr(s, prod, s2.data(), l, scratch);
            bool tn = limbs::add_signed(false, x2, true, y2, t, l);
            xn = limbs::add_signed(tn, t, c[0].neg, c[0].mag.data(), x, l);
            limbs::add(x2, y2, s, l);
            tn = limbs::add_signed(false, s2.data(), true, s, t, l);
            yn = limbs::add_signed(tn, t, c[1].neg, c[1].mag.data(), y, l);
        }
        // step = z_P / z_P', with z_P's parts brought to one exponent
        double mr = 0, mi = 0;
        int64_t er = INT64_MIN, ei = INT64_MIN;
        bool hr = frexp_limbs(xn, x, l, &mr, &er), hi = frexp_limbs(yn, y, l, &mi, &ei);
        out->iterations = it + 1;
        if (!hr && !hi) {  // z_P is exactly 0: c is the nucleus to the last bit
            out->converged = true;
            out->log2_step = -INFINITY;
            break;
        }
        int64_t E = std::max(hr ? er : INT64_MIN, hi ? ei : INT64_MIN);
        double ar = hr ? std::ldexp(mr, (int)std::max<int64_t>(er - E, -2000)) : 0.0;
        double ai = hi ? std::ldexp(mi, (int)std::max<int64_t>(ei - E, -2000)) : 0.0;
        double den = dz.re * dz.re + dz.im * dz.im;
        double qr = (ar * dz.re + ai * dz.im) / den, qi = (ai * dz.re - ar * dz.im) / den;
        int64_t qe = E - dz.e;
        if (!std::isfinite(qr) || !std::isfinite(qi)) {
            *err = "the nucleus search's derivative vanished";
            return false;
        }
        double mag = std::sqrt(qr * qr + qi * qi);
        out->log2_step = std::log2(mag) + (double)qe;
        // c -= step, in the fixed point (the part of the step below the last limb drops)
        for (int part = 0; part < 2; part++) {
            bool sn;
            std::vector<uint64_t> sm, nc(l);
            if (!fixed_from_scaled(part == 0 ? qr : qi, qe, l, &sn, &sm)) {
                *err = "the nucleus search diverged";
                return false;
            }
            c[part].neg = limbs::add_signed(c[part].neg, c[part].mag.data(), !sn, sm.data(),
                                            nc.data(), l);
            c[part].mag = std::move(nc);
        }
        if (out->log2_step < limit) {
            out->converged = true;
            break;
        }
    }
    const size_t digits = (size_t)std::ceil((double)bits * 0.30102999566398120) + 4;
    out->re = fixed_to_decimal(c[0].neg, c[0].mag.data(), l, digits);
    out->im = fixed_to_decimal(c[1].neg, c[1].mag.data(), l, digits);
    if (!out->converged) *err = "the nucleus search did not converge";
    return out->converged;
}

uint64_t Orbit::nucleus_period_of(const double* re, const double* im, uint64_t len) {
    if (len < 2) return 0;
    double record = hypot_ref(re[1], im[1]);
    double log_beta = 0.0;
    for (uint64_t p = 2; p < len / 2; p++) {
        double previous = hypot_ref(re[p - 1], im[p - 1]);
        log_beta += 1.0 + std::log2(previous);
        double radius = hypot_ref(re[p], im[p]);
        if (radius > 1e-140 && radius < 1e-24 && radius < record * 1e-12 &&
            1.0 + log_beta + std::log2(radius) < -40.0 &&
            hypot_ref(re[2 * p], im[2 * p]) <= 2.0 * radius)
            return p;
        record = std::fmin(record, radius);
    }
    return 0;
}

void dc_bound(const Plane& pl, Tier tier, uint32_t nx, uint32_t ny, double* dc, int64_t* e) {
    double half = hypot_ref((double)nx / 2.0, (double)ny / 2.0);
    if (tier == TIER_SCALED) {
        *dc = pl.h_m * half;
        *e = pl.h_e;
    } else {
        *dc = pl.h() * half;
        *e = 0;
    }
}

// ------------------------------------------------------------------ series

// The series recurrence on the coefficients' own arrays: with dc = rho x the delta after
// n steps is S_n(x) = sum b[k] x^k, b[1]' = 2 Z b[1] + rho, b[k]' = 2 Z b[k] + (S^2)[k],
// keeping `active` orders; E and F bound the state and d/dx error of what is dropped
// (series.rs:8-24).  Each sum runs in the reference's order, so the coefficients and
// bounds are its bits.
Series Series::build(const double* re, const double* im, uint64_t len, double rho,
                     double pixel, uint64_t max_iter, const SeriesConfig& cfg,
                     uint64_t stalled_steps) {
    const uint32_t k_max = std::max<uint32_t>(cfg.max_order, 1);
    uint32_t active = std::min(std::max<uint32_t>(cfg.start_order, 1), k_max);
    std::vector<double> br(k_max + 1, 0.0), bi(k_max + 1, 0.0);
    std::vector<double> sq_r(2 * k_max + 1, 0.0), sq_i(2 * k_max + 1, 0.0);
    Series out;
    out.rho = rho;
    out.points = len;
    const uint64_t limit = std::min(len >= 2 ? len - 2 : 0, max_iter >= 1 ? max_iter - 1 : 0);
    // the last admitted step: n, coefficients, E, F, their budgets and Dmin
    bool have = false;
    uint64_t best_n = 0;
    std::vector<double> best;
    double best_e = 0, best_f = 0, budget_e = 0, budget_f = 0, best_dmin = 0;
    if (!(rho > 0.0 && std::isfinite(rho) && pixel > 0.0)) {
        out.stop = STOP_RADIUS;
        return out;
    }
    const double cap_e = cfg.shift_px * pixel * 2.0 / rho, cap_f = cfg.derivative_rel * 2.0;
    double e = 0.0, f = 0.0;
    for (uint64_t n = 0; n < limit; n++) {
        double zr = re[n], zi = im[n];
        const double az = hypot_ref(zr, zi);
        std::fill(sq_r.begin(), sq_r.begin() + 2 * active + 1, 0.0);
        std::fill(sq_i.begin(), sq_i.begin() + 2 * active + 1, 0.0);
        for (uint32_t i = 1; i <= active; i++) {
            const double pr = br[i], pi = bi[i];
            if (pr == 0.0 && pi == 0.0) continue;
            sq_r[2 * i] += pr * pr - pi * pi;
            sq_i[2 * i] += 2.0 * pr * pi;
            const double tr = 2.0 * pr, ti = 2.0 * pi;
            for (uint32_t j = i + 1; j <= active; j++) {
                sq_r[i + j] += tr * br[j] - ti * bi[j];
                sq_i[i + j] += tr * bi[j] + ti * br[j];
            }
        }
        out.coefficient_pairs += (uint64_t)active * (active + 1) / 2;
        double m = 0.0, d = 0.0, t = 0.0, td = 0.0;
        for (uint32_t k = 1; k <= active; k++) {
            const double a = hypot_ref(br[k], bi[k]);
            m += a;
            d += (double)k * a;
        }
        for (uint32_t k = active + 1; k <= 2 * active; k++) {
            const double a = hypot_ref(sq_r[k], sq_i[k]);
            t += a;
            td += (double)k * a;
        }
        const double e_next = 2.0 * (az + m) * e + e * e + t;
        f = 2.0 * (az + m + e) * f + 2.0 * e * d + td;
        e = e_next;
        for (uint32_t k = 1; k <= active; k++) {
            const double xr = br[k], xi = bi[k];
            br[k] = 2.0 * (zr * xr - zi * xi) + sq_r[k];
            bi[k] = 2.0 * (zr * xi + zi * xr) + sq_i[k];
        }
        br[1] += rho;
        const uint64_t step = n + 1;
        out.built_steps = step;
        zr = re[step];
        zi = im[step];
        double tail_d = 0.0;
        m = 0.0;
        for (uint32_t k = 1; k <= active; k++) {
            const double a = hypot_ref(br[k], bi[k]);
            m += a;
            if (k >= 2) tail_d += (double)k * a;
        }
        if (!(std::isfinite(e) && std::isfinite(f))) {
            out.stop = STOP_NONFINITE;
            break;
        }
        if (hypot_ref(zr, zi) + m + e >= 2.0) {  // no sample may leave the escape disk
            out.stop = STOP_ESCAPE;
            break;
        }
        const double dmin = hypot_ref(br[1], bi[1]) - tail_d - f;
        const double budget = dmin > 0.0 ? cfg.shift_px * pixel * dmin / rho : 0.0;
        if (dmin > 0.0 && e <= budget && f <= cfg.derivative_rel * dmin) {
            have = true;
            best_n = step;
            best.resize(2 * active);
            for (uint32_t k = 1; k <= active; k++) {
                best[2 * (k - 1)] = br[k];
                best[2 * (k - 1) + 1] = bi[k];
            }
            best_e = e;
            best_f = f;
            budget_e = budget;
            budget_f = cfg.derivative_rel * dmin;
            best_dmin = dmin;
        } else if (e > cap_e || f > cap_f) {
            // |b1| <= M < 2 caps every later budget: nothing past here can pass
            out.stop = STOP_BUDGET;
            break;
        }
        if (active < k_max && e > cfg.grow_at * budget) {
            active = std::min(2 * active, k_max);
            out.growth_steps++;
        }
        if (stalled_steps > 0 && have && step - best_n >= stalled_steps) {
            out.stop = STOP_STALL;
            break;
        }
    }
    if (have) {
        // drop high orders whose norms fit in half the remaining slack of each bound
        const double slack_e = budget_e - best_e, slack_f = budget_f - best_f;
        double te = 0.0, tf = 0.0;
        size_t kept = best.size() / 2;
        while (kept > 1) {
            const double a = hypot_ref(best[2 * (kept - 1)], best[2 * (kept - 1) + 1]);
            if (te + a > 0.5 * slack_e || tf + (double)kept * a > 0.5 * slack_f) break;
            te += a;
            tf += (double)kept * a;
            kept--;
        }
        best.resize(2 * kept);
        out.n = best_n;
        out.b = std::move(best);
        out.state_bound = best_e + te;
        out.derivative_bound = best_f + tf;
        out.shift_px = out.state_bound * rho / (best_dmin * pixel);
    }
    return out;
}

// ------------------------------------------------------------------ encoder

static uint16_t quantised_angle(double x, double y) {
    double t = std::atan2(y, x) * (65536.0 / 6.283185307179586);
    double r = std::round(t);
    if (std::isnan(r)) return 0;
    int64_t v = r >= 9.2233720368547758e18 ? INT64_MAX
                : r <= -9.2233720368547758e18 ? INT64_MIN : (int64_t)r;
    return (uint16_t)(uint64_t)v;
}

void Encoder::put(size_t idx, uint8_t kind, uint64_t n, double zr, double zi, double dr,
                  double di, int64_t dexp, uint8_t* cls, double* nu, float* de,
                  uint16_t* normal) const {
    if (kind != KIND_ESCAPED) {
        if (cls) cls[idx] = kind == KIND_UNRESOLVED ? 2 : 1;
        return;
    }
    double z2 = zr * zr + zi * zi;
    double log2z = 0.5 * std::log2(z2);
    if (cls) cls[idx] = 0;
    if (nu) nu[idx] = (double)n + 1.0 - std::log2(log2z);
    if (de) {
        // store.rs:141-150: 2|z| ln|z| / |dz/dc| in output pixels, exponent-safe.
        double az = std::sqrt(z2);
        double lnz = log2z * 0.693147180559945309417232121458176568;
        double base = 2.0 * az * lnz / (hypot_ref(dr, di) * px_m);
        int64_t k = std::min<int64_t>(std::max<int64_t>(-(dexp + px_e), -4000), 4000);
        de[idx] = (float)(base * exp2i(k / 2) * exp2i(k - k / 2));
    }
    if (normal) {
        double a = zr * dr + zi * di, b = zi * dr - zr * di;
        normal[idx] = quantised_angle(pl.cos_r * a + pl.sin_r * b, pl.sin_r * a - pl.cos_r * b);
    }
}

// ------------------------------------------------------------------ .fds

namespace {
struct Out {
    std::vector<uint8_t> b;
    template <class T>
    void put(T v) {
        const uint8_t* p = reinterpret_cast<const uint8_t*>(&v);
        b.insert(b.end(), p, p + sizeof(T));
    }
    void bytes(const void* p, size_t n) {
        const uint8_t* q = static_cast<const uint8_t*>(p);
        b.insert(b.end(), q, q + n);
    }
    void str(const std::string& s) {
        put<uint16_t>((uint16_t)s.size());
        bytes(s.data(), s.size());
    }
    void pad8() { b.resize((b.size() + 7) & ~size_t(7), 0); }
};
}  // namespace

std::vector<uint8_t> write_fds(const std::string& re, const std::string& im,
                               const std::string& width, double rotation, uint32_t nx,
                               uint32_t ny, uint32_t ss, uint64_t max_iter, double escape_radius,
                               uint32_t columns, const std::string& kernel, const uint8_t* cls,
                               const double* nu, const float* de, const uint16_t* normal) {
    Out o;
    uint32_t cols = columns | COL_CLASS;
    o.bytes("FDSAMPLE", 8);
    o.put<uint16_t>(1);  // MAJOR
    o.put<uint16_t>(1);  // MINOR
    o.put<uint32_t>(cols);
    o.put<uint32_t>(nx);
    o.put<uint32_t>(ny);
    o.put<uint32_t>(ss);
    o.put<uint32_t>(0);
    o.put<uint64_t>(max_iter);
    o.put<double>(escape_radius);
    o.put<double>(rotation);
    o.str(re);
    o.str(im);
    o.str(width);
    o.str(kernel);
    o.pad8();
    size_t n = (size_t)nx * ny;
    o.bytes(cls, n);
    o.pad8();
    if (cols & COL_NU) {
        o.bytes(nu, 8 * n);
        o.pad8();
    }
    if (cols & COL_DE) {
        o.bytes(de, 4 * n);
        o.pad8();
    }
    if (cols & COL_NORMAL) {
        o.bytes(normal, 2 * n);
        o.pad8();
    }
    return o.b;
}

}  // namespace fdg
