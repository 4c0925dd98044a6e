/*
 * fdo.c - CPU oracle for fractodactyl's deep-zoom perturbation path.
 *
 * TEST INFRASTRUCTURE ONLY.  This restates the reference's CPU renderer in
 * plain C so the HIP product can be checked against it; it is never linked or
 * called by the product library.  It must be built with -ffp-contract=off and
 * without fast-math: every floating-point expression below keeps the
 * reference's left-to-right operation order so results are bit-identical.
 *
 * Reference files restated (paths relative to the reference repository):
 *   crates/fd-fixed/src/{lib.rs,limbs.rs,orbit.rs,decimal.rs}
 *   crates/fd-kernel/src/{view.rs,reference.rs,bla.rs,number.rs,interior.rs,
 *                         sample.rs,store.rs,grid.rs}
 *   crates/fd-samples/src/{samples.rs,header.rs,wire.rs,write.rs}
 * Not restated: the stripe column (stripe.rs), the series initializer
 * (series.rs), refine.rs, zones, GuardedBla.
 */
#include "fdo.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

typedef unsigned __int128 u128;

static void set_err(char *err, size_t cap, const char *fmt, ...) {
    if (!err || cap == 0) return;
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(err, cap, fmt, ap);
    va_end(ap);
}

static double now_s(void) {
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

/* ---------------------------------------------------------------- fd-fixed */

/* crates/fd-fixed/src/lib.rs:30-42 exp2i */
double fdo_exp2i(int64_t k) {
    uint64_t b;
    double d;
    if (k > 1023) return INFINITY;
    if (k >= -1022) {
        b = (uint64_t)(k + 1023) << 52;
    } else if (k >= -1074) {
        b = 1ull << (k + 1074);
    } else {
        return 0.0;
    }
    memcpy(&d, &b, 8);
    return d;
}
#define exp2i fdo_exp2i

/* Rust's `x as i64` for a float: saturating, NaN to 0 (C's cast is undefined there).  It
 * matters where a term is NaN: sum2 (sample.rs:1097) then takes the term's own exponent,
 * and a NaN term far below the other one drops out of the sum. */
static inline int64_t rs_i64(double x) {
    if (x != x) return 0;
    if (x >= 9223372036854775808.0) return INT64_MAX;
    if (x < -9223372036854775808.0) return INT64_MIN;
    return (int64_t)x;
}

/* i64 + and - as the reference's release build computes them (wrapping: the workspace's
 * release profile leaves overflow checks off, Cargo.toml:11-14) */
static inline int64_t wadd(int64_t a, int64_t b) { return (int64_t)((uint64_t)a + (uint64_t)b); }
static inline int64_t wsub(int64_t a, int64_t b) { return (int64_t)((uint64_t)a - (uint64_t)b); }

static int64_t f64_exponent(double v) {
    uint64_t b;
    memcpy(&b, &v, 8);
    return (int64_t)((b >> 52) & 0x7ff) - 1023;
}

/* lib.rs:25-28 limbs_for */
static size_t limbs_for(uint64_t bits) { return 1 + (size_t)((bits + 63) / 64); }

/* limbs.rs:5-14 cmp */
static int limbs_cmp(const uint64_t *a, const uint64_t *b, size_t l) {
    for (size_t i = l; i-- > 0;) {
        if (a[i] < b[i]) return -1;
        if (a[i] > b[i]) return 1;
    }
    return 0;
}

/* limbs.rs:17-27 add */
static void limbs_add(const uint64_t *a, const uint64_t *b, uint64_t *out, size_t l) {
    unsigned c = 0;
    for (size_t i = 0; i < l; i++) {
        uint64_t s = a[i] + b[i];
        unsigned c1 = s < a[i];
        uint64_t s2 = s + c;
        unsigned c2 = s2 < s;
        out[i] = s2;
        c = c1 | c2;
    }
}

/* limbs.rs:30-40 sub (a >= b) */
static void limbs_sub(const uint64_t *a, const uint64_t *b, uint64_t *out, size_t l) {
    unsigned br = 0;
    for (size_t i = 0; i < l; i++) {
        uint64_t d = a[i] - b[i];
        unsigned b1 = a[i] < b[i];
        uint64_t d2 = d - br;
        unsigned b2 = d < (uint64_t)br;
        out[i] = d2;
        br = b1 | b2;
    }
}

static int any_nonzero(const uint64_t *v, size_t l) {
    for (size_t i = 0; i < l; i++)
        if (v[i]) return 1;
    return 0;
}

/* limbs.rs:43-60 add_signed: returns the sign (1 = negative) */
static int limbs_add_signed(int an, const uint64_t *a, int bn, const uint64_t *b, uint64_t *out,
                            size_t l) {
    if (an == bn) {
        limbs_add(a, b, out, l);
        return an && any_nonzero(out, l);
    }
    if (limbs_cmp(a, b, l) < 0) {
        limbs_sub(b, a, out, l);
        return bn;
    }
    limbs_sub(a, b, out, l);
    return an && any_nonzero(out, l);
}

/* limbs.rs:62-100 sqr: out = a^2 truncated to l limbs; prod scratch 2l */
static void limbs_sqr(const uint64_t *a, uint64_t *prod, uint64_t *out, size_t l) {
    memset(prod, 0, 2 * l * sizeof(uint64_t));
    for (size_t i = 0; i < l; i++) {
        u128 ai = a[i];
        if (ai == 0) continue;
        uint64_t c = 0;
        for (size_t j = i + 1; j < l; j++) {
            u128 t = ai * (u128)a[j] + (u128)prod[i + j] + (u128)c;
            prod[i + j] = (uint64_t)t;
            c = (uint64_t)(t >> 64);
        }
        prod[i + l] = c;
    }
    uint64_t top = 0;
    for (size_t i = 0; i < 2 * l; i++) {
        uint64_t next = prod[i] >> 63;
        prod[i] = (prod[i] << 1) | top;
        top = next;
    }
    uint64_t c = 0;
    for (size_t i = 0; i < l; i++) {
        u128 sq = (u128)a[i] * (u128)a[i];
        u128 t = (u128)prod[2 * i] + (u128)(uint64_t)sq + (u128)c;
        prod[2 * i] = (uint64_t)t;
        u128 t2 = (u128)prod[2 * i + 1] + (sq >> 64) + (t >> 64);
        prod[2 * i + 1] = (uint64_t)t2;
        c = (uint64_t)(t2 >> 64);
    }
    memcpy(out, prod + (l - 1), l * sizeof(uint64_t));
}

/* lib.rs:74-83 frexp on raw sign and limbs; returns 0 for zero */
static int fixed_frexp(int neg, const uint64_t *mag, size_t l, double *m_out, int64_t *e_out) {
    size_t k = l;
    while (k > 0 && mag[k - 1] == 0) k--;
    if (k == 0) return 0;
    k -= 1;
    double b1 = k >= 1 ? (double)mag[k - 1] : 0.0;
    double b2 = k >= 2 ? (double)mag[k - 2] : 0.0;
    double v = (double)mag[k] + b1 * exp2i(-64) + b2 * exp2i(-128);
    int64_t e = f64_exponent(v);
    double m = v * exp2i(-e);
    int64_t scale = 64 * ((int64_t)k - ((int64_t)l - 1));
    *m_out = neg ? -m : m;
    *e_out = e + scale;
    return 1;
}

/* lib.rs:86-93 to_f64 */
static double fixed_to_f64(int neg, const uint64_t *mag, size_t l) {
    double m;
    int64_t e;
    if (!fixed_frexp(neg, mag, l, &m, &e)) return 0.0;
    if (e < -1000) return m * exp2i(e + 100) * exp2i(-100);
    return m * exp2i(e);
}

/* decimal.rs:6-15 Decimal */
typedef struct {
    int neg;
    uint8_t *digits;
    size_t nd;
    int64_t exp10;
} decimal;

static int is_space(char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == '\f' || c == '\v'; }

/* Rust i64::from_str: optional sign, >= 1 digit, overflow is an error. */
static int parse_i64(const char *s, size_t n, int64_t *out) {
    size_t i = 0;
    int neg = 0;
    if (n > 0 && (s[0] == '+' || s[0] == '-')) {
        neg = s[0] == '-';
        i = 1;
    }
    if (i == n) return -1;
    u128 v = 0;
    for (; i < n; i++) {
        if (s[i] < '0' || s[i] > '9') return -1;
        v = v * 10 + (unsigned)(s[i] - '0');
        if (v > ((u128)1 << 63)) return -1;
    }
    if (!neg && v > (u128)INT64_MAX) return -1;
    *out = neg ? (int64_t)(0 - (uint64_t)v) : (int64_t)v;
    return 0;
}

/* decimal.rs:19-50 Decimal::parse */
static int decimal_parse(const char *s, decimal *d) {
    size_t n = strlen(s), b = 0, e = n;
    while (b < n && is_space(s[b])) b++;
    while (e > b && is_space(s[e - 1])) e--;
    const char *t = s + b;
    size_t tn = e - b;
    int neg = 0;
    if (tn > 0 && (t[0] == '-' || t[0] == '+')) {
        neg = t[0] == '-';
        t++;
        tn--;
    }
    size_t epos = tn;
    for (size_t i = 0; i < tn; i++)
        if (t[i] == 'e' || t[i] == 'E') {
            epos = i;
            break;
        }
    int64_t exp = 0;
    if (epos < tn && parse_i64(t + epos + 1, tn - epos - 1, &exp) != 0) return -1;
    size_t mant_n = epos, dot = mant_n;
    for (size_t i = 0; i < mant_n; i++)
        if (t[i] == '.') {
            dot = i;
            break;
        }
    size_t int_n = dot, frac_n = dot < mant_n ? mant_n - dot - 1 : 0;
    size_t nd = int_n + frac_n;
    if (nd == 0) return -1;
    uint8_t *digits = (uint8_t *)malloc(nd);
    for (size_t i = 0; i < int_n; i++) digits[i] = (uint8_t)(t[i] - '0');
    for (size_t i = 0; i < frac_n; i++) digits[int_n + i] = (uint8_t)(t[dot + 1 + i] - '0');
    int zero = 1;
    for (size_t i = 0; i < nd; i++) {
        if (digits[i] > 9) {
            free(digits);
            return -1;
        }
        if (digits[i]) zero = 0;
    }
    /* exp.checked_sub(frac.len()) */
    if ((int64_t)frac_n < 0 || (exp < 0 && exp < INT64_MIN + (int64_t)frac_n)) {
        free(digits);
        return -1;
    }
    d->neg = neg && !zero;
    d->digits = digits;
    d->nd = nd;
    d->exp10 = exp - (int64_t)frac_n;
    return 0;
}

/* decimal.rs:52-61 log2_abs */
static double decimal_log2_abs(const decimal *d) {
    size_t first = 0;
    while (first < d->nd && d->digits[first] == 0) first++;
    if (first == d->nd) return -INFINITY;
    size_t sn = d->nd - first;
    double head = 0.0;
    for (size_t i = 0; i < sn && i < 17; i++) head = head * 10.0 + (double)d->digits[first + i];
    double rest = sn > 17 ? (double)(sn - 17) : 0.0;
    return (log10(head) + rest + (double)d->exp10) * 3.32192809488736234787031942948939018;
}

typedef struct {
    uint64_t *v;
    size_t n, cap;
} vec64;

static void vec_push(vec64 *v, uint64_t x) {
    if (v->n == v->cap) {
        v->cap = v->cap ? 2 * v->cap : 8;
        v->v = (uint64_t *)realloc(v->v, v->cap * sizeof(uint64_t));
    }
    v->v[v->n++] = x;
}

/* limbs.rs:103-140 mul_small / add_small / div_small */
static void mul_small(vec64 *v, uint64_t m) {
    uint64_t c = 0;
    for (size_t i = 0; i < v->n; i++) {
        u128 t = (u128)v->v[i] * m + c;
        v->v[i] = (uint64_t)t;
        c = (uint64_t)(t >> 64);
    }
    if (c) vec_push(v, c);
}
static void add_small(vec64 *v, uint64_t a) {
    uint64_t c = a;
    for (size_t i = 0; i < v->n; i++) {
        uint64_t s = v->v[i] + c;
        c = s < v->v[i];
        v->v[i] = s;
        if (!c) return;
    }
    if (c) vec_push(v, c);
}
static void div_small(uint64_t *v, size_t n, uint64_t d) {
    u128 r = 0;
    for (size_t i = n; i-- > 0;) {
        u128 cur = (r << 64) | v[i];
        v[i] = (uint64_t)(cur / d);
        r = cur % d;
    }
}

static uint64_t pow10u(unsigned k) {
    uint64_t p = 1;
    while (k--) p *= 10;
    return p;
}

/* decimal.rs:63-96 to_fixed: truncated toward zero to `limbs` limbs. out has limbs. */
static int decimal_to_fixed(const decimal *d, size_t limbs, int *neg, uint64_t *out) {
    vec64 n = {0};
    vec_push(&n, 0);
    for (size_t c = 0; c < d->nd; c += 19) {
        size_t len = d->nd - c < 19 ? d->nd - c : 19;
        uint64_t val = 0;
        for (size_t i = 0; i < len; i++) val = val * 10 + d->digits[c + i];
        mul_small(&n, pow10u((unsigned)len));
        add_small(&n, val);
    }
    int64_t e = d->exp10;
    while (e > 0) {
        unsigned k = e < 19 ? (unsigned)e : 19;
        mul_small(&n, pow10u(k));
        e -= k;
        if (n.n > 1) {
            free(n.v);
            return -1;
        }
    }
    size_t vn = limbs - 1 + n.n;
    uint64_t *v = (uint64_t *)calloc(vn, sizeof(uint64_t));
    memcpy(v + limbs - 1, n.v, n.n * sizeof(uint64_t));
    free(n.v);
    while (e < 0) {
        unsigned k = -e < 19 ? (unsigned)(-e) : 19;
        div_small(v, vn, k == 19 ? 10000000000000000000ull : pow10u(k));
        e += k;
        if (!any_nonzero(v, vn)) break;
    }
    if (vn > limbs && any_nonzero(v + limbs, vn - limbs)) {
        free(v);
        return -1;
    }
    memcpy(out, v, limbs * sizeof(uint64_t));
    *neg = d->neg && any_nonzero(out, limbs);
    free(v);
    return 0;
}

static int fixed_parse(const char *s, size_t limbs, int *neg, uint64_t *out) {
    decimal d;
    if (decimal_parse(s, &d) != 0) return -1;
    int r = decimal_to_fixed(&d, limbs, neg, out);
    free(d.digits);
    return r;
}

/* orbit.rs:14-43 fd_fixed::orbit: emits Z_1, Z_2, ... */
typedef void (*emit_fn)(void *ctx, double a, double b);
static void fixed_orbit(int crn, const uint64_t *cr, int cin, const uint64_t *ci, size_t l,
                        uint64_t steps, double escape2, emit_fn emit, void *ctx) {
    uint64_t *buf = (uint64_t *)calloc(9 * l, sizeof(uint64_t));
    uint64_t *x = buf, *y = buf + l, *x2 = buf + 2 * l, *y2 = buf + 3 * l, *s = buf + 4 * l,
             *s2 = buf + 5 * l, *t = buf + 6 * l, *prod = buf + 7 * l;
    int xn = 0, yn = 0;
    for (uint64_t it = 0; it < steps; it++) {
        limbs_sqr(x, prod, x2, l);
        limbs_sqr(y, prod, y2, l);
        limbs_add_signed(xn, x, yn, y, s, l);
        limbs_sqr(s, prod, s2, l);
        int tn = limbs_add_signed(0, x2, 1, y2, t, l);
        xn = limbs_add_signed(tn, t, crn, cr, x, l);
        limbs_add(x2, y2, s, l);
        tn = limbs_add_signed(0, s2, 1, s, t, l);
        yn = limbs_add_signed(tn, t, cin, ci, y, l);
        double fx = fixed_to_f64(xn, x, l), fy = fixed_to_f64(yn, y, l);
        emit(ctx, fx, fy);
        if (fx * fx + fy * fy > escape2) break;
    }
    free(buf);
}

/* ---------------------------------------------------------------- view.rs */

/* view.rs:42-87 Plane::new */
int fdo_plane_new(const char *re, const char *im, const char *width, double rotation,
                  uint32_t nx, uint32_t ny, fdo_plane *p, char *err, size_t errcap) {
    if (nx == 0 || ny == 0) {
        set_err(err, errcap, "empty sample grid");
        return -1;
    }
    decimal w;
    if (decimal_parse(width, &w) != 0) {
        set_err(err, errcap, "not a decimal number: \"%s\"", width);
        return -1;
    }
    double lw = decimal_log2_abs(&w);
    if (w.neg || !isfinite(lw)) {
        free(w.digits);
        set_err(err, errcap, "view width must be positive");
        return -1;
    }
    double fb = fmax(64.0 - lw, 64.0);
    size_t l = limbs_for((uint64_t)ceil(fb));
    uint64_t *wf = (uint64_t *)calloc(l, sizeof(uint64_t));
    int wneg;
    if (decimal_to_fixed(&w, l, &wneg, wf) != 0) {
        free(w.digits);
        free(wf);
        set_err(err, errcap, "value too large for fixed point");
        return -1;
    }
    free(w.digits);
    double wm;
    int64_t we;
    if (!fixed_frexp(wneg, wf, l, &wm, &we)) {
        free(wf);
        set_err(err, errcap, "view width underflowed");
        return -1;
    }
    free(wf);
    double q = wm / (double)nx;
    int64_t qe = f64_exponent(q);
    p->h_m = q * exp2i(-qe);
    p->h_e = we + qe;
    decimal tmp;
    if (decimal_parse(re, &tmp) != 0) {
        set_err(err, errcap, "not a decimal number: \"%s\"", re);
        return -1;
    }
    free(tmp.digits);
    if (decimal_parse(im, &tmp) != 0) {
        set_err(err, errcap, "not a decimal number: \"%s\"", im);
        return -1;
    }
    free(tmp.digits);
    p->c_re = strtod(re, NULL);
    p->c_im = strtod(im, NULL);
    if (!(isfinite(p->c_re) && isfinite(p->c_im))) {
        set_err(err, errcap, "view centre out of range");
        return -1;
    }
    double c_mag = fmax(fmax(fabs(p->c_re), fabs(p->c_im)), exp2i(-1022));
    int f64_ok = log2(c_mag) - 52.0 <= (double)p->h_e - 10.0 && p->h_e > -1000;
    p->tier = f64_ok ? FDO_TIER_F64 : (p->h_e >= -900 ? FDO_TIER_FIXED : FDO_TIER_SCALED);
    int64_t b = 128 - p->h_e;
    p->bits = (uint64_t)(b > 128 ? b : 128);
    p->sin_r = sin(rotation);
    p->cos_r = cos(rotation);
    p->half_x = (double)nx / 2.0;
    p->half_y = (double)ny / 2.0;
    return 0;
}

/* view.rs:117-122 unit_offset */
static inline void unit_offset(const fdo_plane *p, size_t i, size_t j, double *ux, double *uy) {
    double x = (double)i + 0.5 - p->half_x;
    double y = (double)j + 0.5 - p->half_y;
    *ux = p->cos_r * x + p->sin_r * y;
    *uy = p->sin_r * x - p->cos_r * y;
}

/* view.rs:126-128 h */
static inline double plane_h(const fdo_plane *p) { return p->h_m * exp2i(p->h_e); }

/* ---------------------------------------------------------------- reference.rs */

#define MAX_LEN (1u << 20)

typedef struct {
    double *re, *im;
    uint64_t len, cap;
} orbit_t;

static void orbit_push(void *ctx, double a, double b) {
    orbit_t *o = (orbit_t *)ctx;
    if (o->len == o->cap) {
        o->cap = o->cap ? 2 * o->cap : 1024;
        o->re = (double *)realloc(o->re, o->cap * sizeof(double));
        o->im = (double *)realloc(o->im, o->cap * sizeof(double));
    }
    o->re[o->len] = a;
    o->im[o->len] = b;
    o->len++;
}

/* reference.rs:21-48 Reference::new / collect */
static void reference_f64(orbit_t *o, double c_re, double c_im, uint64_t max_iter, double radius) {
    uint64_t steps = max_iter < MAX_LEN - 1 ? max_iter : MAX_LEN - 1;
    double x = 0.0, y = 0.0;
    orbit_push(o, 0.0, 0.0);
    for (uint64_t k = 0; k < steps; k++) {
        double t = x * x - y * y + c_re;
        y = 2.0 * x * y + c_im;
        x = t;
        orbit_push(o, x, y);
        if (x * x + y * y > radius * radius) break;
    }
}

/* reference.rs:85-96 Reference::from_fixed */
static int reference_fixed(orbit_t *o, const char *re, const char *im, uint64_t bits,
                           uint64_t max_iter) {
    size_t l = limbs_for(bits);
    uint64_t *cr = (uint64_t *)calloc(l, 8), *ci = (uint64_t *)calloc(l, 8);
    int crn, cin;
    if (fixed_parse(re, l, &crn, cr) != 0 || fixed_parse(im, l, &cin, ci) != 0) {
        free(cr);
        free(ci);
        return -1;
    }
    uint64_t steps = max_iter < MAX_LEN - 1 ? max_iter : MAX_LEN - 1;
    orbit_push(o, 0.0, 0.0);
    fixed_orbit(crn, cr, cin, ci, l, steps, 4294967296.0, orbit_push, o);
    free(cr);
    free(ci);
    return 0;
}

int64_t fdo_reference_fixed(const char *re, const char *im, uint64_t bits, uint64_t max_iter,
                            double *out_re, double *out_im, uint64_t cap) {
    orbit_t o = {0};
    if (reference_fixed(&o, re, im, bits, max_iter) != 0) return -1;
    uint64_t n = o.len < cap ? o.len : cap;
    memcpy(out_re, o.re, n * 8);
    memcpy(out_im, o.im, n * 8);
    free(o.re);
    free(o.im);
    return (int64_t)o.len;
}

/* dz' = 2 z dz + 1 with z = (ar + i ai) 2^ae and dz = (*dr + i *di) 2^*dx, the result kept
 * as a mantissa in [2^-512, 2^512] (or 0) times 2^*dx */
static void dz_step(double ar, double ai, int64_t ae, double *dr, double *di, int64_t *dx) {
    double pr = 2.0 * (ar * *dr - ai * *di), pi = 2.0 * (ar * *di + ai * *dr);
    int64_t pe = ae + *dx;
    double a = fmax(fabs(pr), fabs(pi));
    if (a == 0.0 || (double)pe + log2(a) < -600.0) {
        /* the product is below 2^-600 of the added 1 */
        *dr = 1.0 + (a == 0.0 ? 0.0 : pr * exp2i(pe));
        *di = a == 0.0 ? 0.0 : pi * exp2i(pe);
        *dx = 0;
        return;
    }
    if (pe < -1000) {
        /* (the product is at least 2^-600: its mantissa can take the difference) */
        pr *= exp2i(pe + 1000);
        pi *= exp2i(pe + 1000);
        pe = -1000;
    }
    pr += exp2i(-pe); /* 0 once pe > 1074: the 1 is far below the product */
    *dr = pr;
    *di = pi;
    *dx = pe;
    a = fmax(fabs(*dr), fabs(*di));
    while (a > 0x1p512) {
        *dr *= 0x1p-512, *di *= 0x1p-512, *dx += 512, a *= 0x1p-512;
    }
    while (a != 0.0 && a < 0x1p-512) {
        *dr *= 0x1p512, *di *= 0x1p512, *dx -= 512, a *= 0x1p512;
    }
}

/* Direct iteration of z <- z^2 + c for the decimal c at `bits` fraction bits (the
 * fd_fixed::orbit arithmetic above) with dz/dc carried alongside in scaled form (each z_k
 * taken to 53 bits with its own exponent, so passes near 0 far below f64's range keep
 * their weight), continued in f64 past |z|^2 > 2^32 up to |z|^2 > r2 (|z| >= 2^16 makes
 * those few steps well conditioned).  Test infrastructure for the direct high-precision
 * fixture (tools/oracle.py:56-67 iterates the same recurrences in mpmath).  Returns n, the
 * first iterate with fl(|z_n|^2) > r2 (-1 when none through max_iter, -2 on a bad decimal);
 * out[0..1] = z_n, out[2..3] = the mantissa of dz/dc at n with *de its exponent, out[4..5] =
 * fl(|z_{n-1}|^2), fl(|z_n|^2). */
int64_t fdo_direct(const char *re, const char *im, uint64_t bits, uint64_t max_iter, double r2,
                   double *out, int64_t *de) {
    size_t l = limbs_for(bits);
    /* (prod holds a full 2l-limb product, as in fixed_orbit) */
    uint64_t *buf = (uint64_t *)calloc(11 * l, sizeof(uint64_t));
    uint64_t *cr = buf, *ci = buf + l, *x = buf + 2 * l, *y = buf + 3 * l, *x2 = buf + 4 * l,
             *y2 = buf + 5 * l, *s = buf + 6 * l, *s2 = buf + 7 * l, *t = buf + 8 * l,
             *prod = buf + 9 * l;
    int crn, cin;
    if (fixed_parse(re, l, &crn, cr) != 0 || fixed_parse(im, l, &cin, ci) != 0) {
        free(buf);
        return -2;
    }
    double c_re = fixed_to_f64(crn, cr, l), c_im = fixed_to_f64(cin, ci, l);
    int xn = 0, yn = 0;
    double dr = 0.0, di = 0.0, prev2 = 0.0, fx = 0.0, fy = 0.0;
    int64_t dx = 0;
    uint64_t n = 0;
    int fixed = 1;
    while (n < max_iter) {
        if (fixed) {
            /* dz/dc from z_n, then z_{n+1} */
            double mx = 0.0, my = 0.0;
            int64_t ex = INT64_MIN / 4, ey = INT64_MIN / 4;
            int hx = fixed_frexp(xn, x, l, &mx, &ex), hy = fixed_frexp(yn, y, l, &my, &ey);
            int64_t ez = hx || hy ? (ex > ey ? ex : ey) : 0;
            double ar = hx ? mx * exp2i(ex - ez) : 0.0, ai = hy ? my * exp2i(ey - ez) : 0.0;
            dz_step(ar, ai, ez, &dr, &di, &dx);
            limbs_sqr(x, prod, x2, l);
            limbs_sqr(y, prod, y2, l);
            limbs_add_signed(xn, x, yn, y, s, l);
            limbs_sqr(s, prod, s2, l);
            int tn = limbs_add_signed(0, x2, 1, y2, t, l);
            xn = limbs_add_signed(tn, t, crn, cr, x, l);
            limbs_add(x2, y2, s, l);
            tn = limbs_add_signed(0, s2, 1, s, t, l);
            yn = limbs_add_signed(tn, t, cin, ci, y, l);
            fx = fixed_to_f64(xn, x, l);
            fy = fixed_to_f64(yn, y, l);
        } else {
            dz_step(fx, fy, 0, &dr, &di, &dx);
            double t2 = fx * fx - fy * fy + c_re;
            fy = 2.0 * fx * fy + c_im;
            fx = t2;
        }
        n++;
        double f2 = fx * fx + fy * fy;
        if (f2 > r2) {
            out[0] = fx, out[1] = fy, out[2] = dr, out[3] = di, out[4] = prev2, out[5] = f2;
            *de = dx;
            free(buf);
            return (int64_t)n;
        }
        if (f2 > 4294967296.0) fixed = 0;
        prev2 = f2;
    }
    free(buf);
    return -1;
}

/* reference.rs:102-120 nucleus_period */
uint64_t fdo_nucleus_period(const double *re, const double *im, uint64_t len) {
    if (len < 2) return 0;
    double record = hypot(re[1], im[1]);
    double log_beta = 0.0;
    for (uint64_t p = 2; p < len / 2; p++) {
        double previous = hypot(re[p - 1], im[p - 1]);
        log_beta += 1.0 + log2(previous);
        double radius = hypot(re[p], im[p]);
        if (radius > 1e-140 && radius < 1e-24 && radius < record * 1e-12 &&
            1.0 + log_beta + log2(radius) < -40.0 && hypot(re[2 * p], im[2 * p]) <= 2.0 * radius)
            return p;
        record = fmin(record, radius);
    }
    return 0;
}

/* ---------------------------------------------------------------- bla.rs */

typedef struct {
    double ar, ai, br, bi, r, alpha, beta;
} block_t;

typedef struct {
    double eps, dc_max;
    int64_t dc_exp;
    uint64_t points;
    int nlevels;
    block_t **levels;
    double **r2;
    uint64_t *count;
} bla_t;

static int bla_check(double eps, double dc_max) {
    return eps > 0.0 && eps <= 1.0 / (double)(1ull << 24) && dc_max > 0.0 && isfinite(dc_max);
}

/* bla.rs:72-85 Block::single */
static block_t block_single(double zr, double zi, double eps) {
    block_t b = {0};
    double az = hypot(zr, zi);
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

/* bla.rs:114-121 times_exp2 */
static double times_exp2(double v, int64_t e) {
    if (e >= -1022) return v * exp2i(e);
    return v * exp2i(-1022) * exp2i(e + 1022);
}

/* bla.rs:91-111 Block::merge */
static block_t block_merge(const block_t *x, const block_t *y, double c, int64_t ce) {
    block_t z = {0};
    if (x->r == 0.0 || y->r == 0.0) return z;
    double ax = hypot(x->ar, x->ai), ay = hypot(y->ar, y->ai);
    double ahat = ax + x->alpha;
    double bhat = hypot(x->br, x->bi) + x->beta;
    double r = fmin(x->r, (y->r - times_exp2(bhat * c, ce)) / ahat);
    /* affine(x.a, y.a, 0) and affine(x.b, y.a, y.b), bla.rs:124-129 */
    z.ar = y->ar * x->ar - y->ai * x->ai + 0.0 + 0.0;
    z.ai = y->ar * x->ai + y->ai * x->ar + 0.0 + 0.0;
    z.br = y->ar * x->br - y->ai * x->bi + y->br + 0.0;
    z.bi = y->ar * x->bi + y->ai * x->br + y->bi + 0.0;
    z.r = r;
    z.alpha = ay * x->alpha + y->alpha * ahat;
    z.beta = ay * x->beta + y->alpha * bhat + y->beta;
    if (r > 0.0 && isfinite(z.ar) && isfinite(z.ai) && isfinite(z.br) && isfinite(z.bi) &&
        isfinite(z.r) && isfinite(z.alpha) && isfinite(z.beta))
        return z;
    block_t zero = {0};
    return zero;
}

static void bla_free(bla_t *t) {
    for (int i = 0; i < t->nlevels; i++) {
        free(t->levels[i]);
        free(t->r2[i]);
    }
    free(t->levels);
    free(t->r2);
    free(t->count);
    memset(t, 0, sizeof(*t));
}

/* bla.rs:161-181 Bla::build_scaled (+ Bla::new's r2 cache, bla.rs:184-216) */
static int bla_build(bla_t *t, const double *re, const double *im, uint64_t len, double eps,
                     double dc_max, int64_t dc_exp) {
    memset(t, 0, sizeof(*t));
    if (!bla_check(eps, dc_max)) return -1;
    uint64_t steps = len >= 2 ? len - 2 : 0;
    block_t *singles = (block_t *)malloc((steps ? steps : 1) * sizeof(block_t));
    for (uint64_t k = 1; k <= steps; k++) singles[k - 1] = block_single(re[k], im[k], eps);
    uint64_t cur_n = steps / 2;
    block_t *next = (block_t *)malloc((cur_n ? cur_n : 1) * sizeof(block_t));
    for (uint64_t i = 0; i < cur_n; i++)
        next[i] = block_merge(&singles[2 * i], &singles[2 * i + 1], dc_max, dc_exp);
    free(singles);
    int cap = 0;
    for (;;) {
        int any = 0;
        for (uint64_t i = 0; i < cur_n; i++)
            if (next[i].r > 0.0) {
                any = 1;
                break;
            }
        if (!any) break;
        uint64_t up_n = cur_n / 2;
        block_t *up = (block_t *)malloc((up_n ? up_n : 1) * sizeof(block_t));
        for (uint64_t i = 0; i < up_n; i++)
            up[i] = block_merge(&next[2 * i], &next[2 * i + 1], dc_max, dc_exp);
        if (t->nlevels == cap) {
            cap = cap ? 2 * cap : 16;
            t->levels = (block_t **)realloc(t->levels, cap * sizeof(block_t *));
            t->count = (uint64_t *)realloc(t->count, cap * sizeof(uint64_t));
        }
        t->levels[t->nlevels] = next;
        t->count[t->nlevels] = cur_n;
        t->nlevels++;
        next = up;
        cur_n = up_n;
    }
    free(next);
    t->eps = eps;
    t->dc_max = dc_max;
    t->dc_exp = dc_exp;
    t->points = len;
    t->r2 = (double **)malloc((t->nlevels ? t->nlevels : 1) * sizeof(double *));
    for (int i = 0; i < t->nlevels; i++) {
        t->r2[i] = (double *)malloc((t->count[i] ? t->count[i] : 1) * sizeof(double));
        for (uint64_t k = 0; k < t->count[i]; k++) t->r2[i][k] = t->levels[i][k].r * t->levels[i][k].r;
    }
    return 0;
}

int fdo_bla_build(const double *re, const double *im, uint64_t len, double eps, double dc_max,
                  int64_t dc_exp, double *out, uint64_t cap_blocks) {
    bla_t t;
    if (bla_build(&t, re, im, len, eps, dc_max, dc_exp) != 0) return -1;
    uint64_t w = 0;
    for (int i = 0; i < t.nlevels; i++)
        for (uint64_t k = 0; k < t.count[i]; k++) {
            if (w < cap_blocks) {
                const block_t *b = &t.levels[i][k];
                double *o = out + 7 * w;
                o[0] = b->ar, o[1] = b->ai, o[2] = b->br, o[3] = b->bi, o[4] = b->r,
                o[5] = b->alpha, o[6] = b->beta;
            }
            w++;
        }
    int n = t.nlevels;
    bla_free(&t);
    return n;
}

/* bla.rs:234-249 Bla::find; returns the level index i (block length 2 << i) or -1 */
static inline int bla_find(const bla_t *t, uint64_t m, double d2, uint64_t budget) {
    uint64_t k = m - 1;
    int tz = k == 0 ? 64 : __builtin_ctzll(k);
    int top = tz < t->nlevels ? tz : t->nlevels;
#define FDO_OK(i)                                                                              \
    ((2ull << (i)) <= budget && (k >> ((i) + 1)) < t->count[(i)] && d2 < t->r2[(i)][k >> ((i) + 1)])
    if (top == 0 || !FDO_OK(0)) return -1;
    int i = 0;
    while (i + 1 < top && FDO_OK(i + 1)) i++;
#undef FDO_OK
    return i;
}

/* ---------------------------------------------------------------- sample.rs */

#define HI 4294967296.0
#define LO (1.0 / 4294967296.0)
#define SQUARE_OK 1e-290
#define TINY (1.0 / 1e270)
#define U_RND (2.220446049250313e-16 / 2.0)
#define RESOLVABLE 1e-12

/* interior.rs:6-11 in_main_components */
static inline int in_main_components(double cr, double ci) {
    double y2 = ci * ci;
    double q = (cr - 0.25) * (cr - 0.25) + y2;
    return q * (q + (cr - 0.25)) <= 0.25 * y2 || (cr + 1.0) * (cr + 1.0) + y2 <= 0.0625;
}

/* interior.rs:16-47 map, interior.rs:55-64 contracting */
static int contracting(const double *zr, const double *zi, uint64_t len, uint64_t m0, double x0r,
                       double x0i, double ar, double ai, uint64_t period) {
    uint64_t last = len - 1, m = m0;
    double xr = x0r, xi = x0i, dr = 1.0, di = 0.0;
    for (uint64_t s = 0; s < period; s++) {
        double fr = zr[m] + xr, fi = zi[m] + xi;
        double t = 2.0 * (fr * dr - fi * di);
        di = 2.0 * (fr * di + fi * dr);
        dr = t;
        t = 2.0 * (zr[m] * xr - zi[m] * xi) + xr * xr - xi * xi + ar;
        xi = 2.0 * (zr[m] * xi + zi[m] * xr + xr * xi) + ai;
        xr = t;
        m += 1;
        fr = zr[m] + xr;
        fi = zi[m] + xi;
        double f2 = fr * fr + fi * fi;
        if (f2 > 4.0) return 0;
        if (f2 < xr * xr + xi * xi || m == last) {
            xr = fr;
            xi = fi;
            m = 0;
        }
    }
    return dr * dr + di * di < 1.0;
}

/* sample.rs:597-652 Multiplier */
typedef struct {
    double re, im;
    int64_t exp;
} multiplier;

static inline void mul_multiply(multiplier *s, double re, double im) {
    double nr = s->re * re - s->im * im, ni = s->re * im + s->im * re;
    s->re = nr;
    s->im = ni;
    double a = fmax(fabs(s->re), fabs(s->im));
    if (!(a > 1e-120 && a < 1e120)) {
        if (a > 0.0 && isfinite(a)) {
            int64_t e = (int64_t)floor(log2(a));
            double scale = exp2i(-e);
            s->re *= scale;
            s->im *= scale;
            s->exp += e;
        }
    }
}
static inline void mul_value(const multiplier *s, double *r, double *i) {
    double scale = s->exp < -1000 ? 0.0 : exp2i(s->exp < 1000 ? s->exp : 1000);
    *r = s->re * scale;
    *i = s->im * scale;
}
static inline int mul_below_one(const multiplier *s) {
    return isfinite(s->re) && isfinite(s->im) &&
           (s->exp < -64 ||
            (s->exp <= 64 && (s->re * s->re + s->im * s->im) * exp2i(2 * s->exp) < 1.0));
}

/* sample.rs:657-698 Contraction */
typedef struct {
    double re, im;
    int64_t exp;
    int started;
} contraction;

static inline void con_multiply(contraction *s, double re, double im, int64_t exp) {
    if (!s->started) {
        s->re = 1.0;
        s->started = 1;
    }
    /* (a non-finite a reaches the casts too, as in the reference: NaN gives e = 0) */
    double a = fmax(fabs(re), fabs(im));
    if (a != 0.0 && !(LO <= a && a <= HI)) {
        int64_t e = rs_i64(floor(log2(a)));
        double scale = exp2i(wsub(0, e));
        re *= scale;
        im *= scale;
        exp = wadd(exp, e);
    }
    double nr = s->re * re - s->im * im, ni = s->re * im + s->im * re;
    s->re = nr;
    s->im = ni;
    s->exp = wadd(s->exp, exp);
    a = fmax(fabs(s->re), fabs(s->im));
    if (a != 0.0 && !(LO <= a && a <= HI)) {
        int64_t e = rs_i64(floor(log2(a)));
        double scale = exp2i(wsub(0, e));
        s->re *= scale;
        s->im *= scale;
        s->exp = wadd(s->exp, e);
    }
}
static inline int con_small(const contraction *s) {
    return s->started && isfinite(s->re) && isfinite(s->im) &&
           (s->exp < -64 ||
            (s->exp <= 64 && (s->re * s->re + s->im * s->im) * exp2i(2 * s->exp) < 1e-6));
}

/* sample.rs:706-708 resolvable */
static inline int resolvable(double xr, double xi, double f2) {
    return xr * xr + xi * xi > RESOLVABLE * RESOLVABLE * f2;
}

/* sample.rs:586-595 normalise */
static inline void normalise(double *re, double *im, int64_t *e, int64_t max_e) {
    double a = fmax(fabs(*re), fabs(*im));
    if (a == 0.0 || !isfinite(a)) return;
    int64_t k = (int64_t)floor(log2(a));
    int64_t lim = max_e - *e; /* saturating_sub: operands are far from overflow here */
    if (lim < k) k = lim;
    double h = exp2i(-(k / 2)), l = exp2i(-(k - k / 2));
    *re = *re * h * l;
    *im = *im * h * l;
    *e = *e + k;
}

/* sample.rs:1094-1114 sum2 (two terms) */
static void sum2(const double t[2][2], const int64_t te[2], int64_t zero_e, double *ore,
                 double *oim, int64_t *oe) {
    int have = 0;
    int64_t top = 0;
    for (int i = 0; i < 2; i++) {
        double a = fmax(fabs(t[i][0]), fabs(t[i][1]));
        if (a != 0.0) { /* (NaN included: its exponent is then te[i] alone) */
            int64_t g = wadd(rs_i64(floor(log2(a))), te[i]);
            if (!have || g > top) top = g;
            have = 1;
        }
    }
    if (!have) {
        *ore = 0.0;
        *oim = 0.0;
        *oe = zero_e;
        return;
    }
    double r = 0.0, im = 0.0;
    for (int i = 0; i < 2; i++) {
        int64_t k = wsub(te[i], top);
        double a = k < -2046 ? 0.0 : t[i][0] * exp2i(k / 2) * exp2i(k - k / 2);
        double b = k < -2046 ? 0.0 : t[i][1] * exp2i(k / 2) * exp2i(k - k / 2);
        r = r + a;
        im = im + b;
    }
    *ore = r;
    *oim = im;
    *oe = top;
}

/* sample.rs:1116-1122 zero_step */
static void zero_step(double wr, double wi, int64_t e, double ar, double ai, int64_t ae,
                      double *ore, double *oim, int64_t *oe) {
    double t[2][2] = {{wr * wr - wi * wi, 2.0 * wr * wi}, {ar, ai}};
    int64_t te[2] = {wadd(e, e), ae};
    sum2(t, te, e, ore, oim, oe);
}

/* zero_step on explicit operands, for the tests: out = {re, im}, *oe = exponent */
void fdo_zero_step(double wr, double wi, int64_t e, double ar, double ai, int64_t ae,
                   double *out, int64_t *oe) {
    zero_step(wr, wi, e, ar, ai, ae, &out[0], &out[1], oe);
}

/* sample.rs:1124-1134 zero_step_dz */
static void zero_step_dz(double wr, double wi, int64_t we, double vr, double vi, int64_t ve,
                         double *ore, double *oim, int64_t *oe) {
    double t[2][2] = {{2.0 * (wr * vr - wi * vi), 2.0 * (wr * vi + wi * vr)}, {1.0, 0.0}};
    int64_t te[2] = {wadd(we, ve), 0};
    sum2(t, te, ve, ore, oim, oe);
}

typedef struct {
    int kind;
    uint64_t n;
    double zr, zi, dr, di;
    int64_t dexp;
} outcome_t;

typedef struct {
    uint64_t blocks, skipped;
} skip_t;

#define U64MAX UINT64_MAX
static inline uint64_t min_u64(uint64_t a, uint64_t b) { return a < b ? a : b; }
static inline uint64_t next_pow2(uint64_t x) {
    if (x <= 1) return 1;
    return 1ull << (64 - __builtin_clzll(x - 1));
}

/* sample.rs:69-80 Start: a sample's state after n iterates (n = 0: the ordinary start). */
typedef struct {
    uint64_t n;
    double xr, xi, dr, di;
} start_t;

/*
 * sample.rs:170-581 iterate<A, D, B=false, L> (no stripe trace), continuing from `st`
 * (NULL: the ordinary start; a series start only on the ordinary tiers).
 * EXT selects Binary64Exponent (number.rs:112-124) over Binary64 (number.rs:95-107).
 */
static inline __attribute__((always_inline)) outcome_t
iterate(const int EXT, const int D, const int L, const double *zr, const double *zi, uint64_t len,
        const bla_t *bla, skip_t *skip, int have_c, double c_re, double c_im, double ar, double ai,
        int64_t offset_exp, uint64_t max_iter, double r2, uint64_t nucleus_in, const start_t *st) {
    outcome_t out = {FDO_UNRESOLVED, 0, 0, 0, 0, 0, 0};
    if (have_c && in_main_components(c_re, c_im)) {
        out.kind = FDO_INTERIOR;
        out.n = 0;
        return out;
    }
    const uint64_t last = len - 1;
    double xr = st ? st->xr : 0.0, xi = st ? st->xi : 0.0;
    int64_t delta_exp = EXT ? offset_exp : 0;
    double scale = EXT ? exp2i(delta_exp) : 1.0;
    double dc_r = ar, dc_i = ai;
    double dr = st ? st->dr : 0.0, di = st ? st->di : 0.0;
    int64_t derivative_exp = 0;
    double inverse = 1.0;
    uint64_t m = st ? st->n : 0, n = m;
    double sr = 0.0, si = 0.0;
    uint64_t chk = 16, sn = EXT ? 0 : 8;
    if (n > 0) {
        /* sample.rs:203-211: a supplied state is the first save; saves resume at the next
         * power of two */
        sr = zr[m] + xr;
        si = zi[m] + xi;
        sn = n;
        chk = next_pow2(n + 1);
        if (chk < 16) chk = 16;
    }
    int returns = 0, pending = 0;
    multiplier mu = {1.0, 0.0, 0};
    uint64_t hold = 0;
    int predicted = 0;
    double qr = 0.0, qi = 0.0, qstep = 0.0;
    double adc = L ? sqrt(ar * ar + ai * ai) : 0.0;
    (void)adc;
    /* sample.rs:226-228: a supplied state skipped the contraction's first factors */
    uint64_t nucleus = n > 0 ? 0 : nucleus_in; /* 0 = None */
    contraction con = {0.0, 0.0, 0, 0};
    while (n < max_iter) {
        int lvl = -1;
        if (L && bla && (nucleus == 0 || m != 0)) {
            double d2 = EXT ? (xr * xr + xi * xi) * scale * scale : xr * xr + xi * xi;
            uint64_t b = min_u64(max_iter, hold > 0 ? sn + hold : (pending ? U64MAX : chk));
            b = min_u64(b, nucleus ? (n == 0 ? 1 : nucleus + 1) : U64MAX);
            b = b > n ? b - n : 0;
            lvl = bla_find(bla, m, d2, b);
        }
        if (lvl >= 0) {
            uint64_t k = m - 1, blen = 2ull << lvl;
            const block_t *bk = &bla->levels[lvl][k >> (lvl + 1)];
            double a_r = bk->ar, a_i = bk->ai, b_r = bk->br, b_i = bk->bi;
            if (nucleus && n >= 1) con_multiply(&con, a_r, a_i, 0);
            if (!EXT && hold > 0) mul_multiply(&mu, a_r, a_i);
            if (EXT) {
                normalise(&xr, &xi, &delta_exp, 0);
                if (D) {
                    normalise(&dr, &di, &derivative_exp, INT64_MAX / 4);
                    inverse = exp2i(-derivative_exp);
                }
                scale = exp2i(delta_exp);
                double factor = exp2i(offset_exp - delta_exp);
                dc_r = ar * factor;
                dc_i = ai * factor;
            }
            if (D) {
                double t = a_r * dr - a_i * di + b_r * inverse;
                di = a_r * di + a_i * dr + b_i * inverse;
                dr = t;
            }
            double t = a_r * xr - a_i * xi + b_r * dc_r - b_i * dc_i;
            xi = a_r * xi + a_i * xr + b_r * dc_i + b_i * dc_r;
            xr = t;
            if (EXT) {
                normalise(&xr, &xi, &delta_exp, 0);
                if (D) {
                    normalise(&dr, &di, &derivative_exp, INT64_MAX / 4);
                    inverse = exp2i(-derivative_exp);
                }
                scale = exp2i(delta_exp);
                double factor = exp2i(offset_exp - delta_exp);
                dc_r = ar * factor;
                dc_i = ai * factor;
            }
            m += blen;
            n += blen;
            skip->blocks += 1;
            skip->skipped += blen;
        } else {
            double fr = zr[m] + scale * xr, fi = zi[m] + scale * xi;
            if (nucleus && n >= 1) con_multiply(&con, 2.0 * fr, 2.0 * fi, 0);
            if (!EXT && hold > 0) mul_multiply(&mu, 2.0 * fr, 2.0 * fi);
            if (D) {
                double t = 2.0 * (fr * dr - fi * di) + inverse;
                di = 2.0 * (fr * di + fi * dr);
                dr = t;
                if (EXT && fmax(fabs(dr), fabs(di)) > HI) {
                    dr = dr * LO;
                    di = di * LO;
                    derivative_exp += 32;
                    inverse = exp2i(-derivative_exp);
                }
            }
            double Zr = zr[m], Zi = zi[m];
            if (EXT) {
                double t = 2.0 * (Zr * xr - Zi * xi) + scale * (xr * xr - xi * xi) + dc_r;
                xi = 2.0 * (Zr * xi + Zi * xr) + scale * (2.0 * xr * xi) + dc_i;
                xr = t;
            } else {
                double t = 2.0 * (Zr * xr - Zi * xi) + xr * xr - xi * xi + dc_r;
                xi = 2.0 * (Zr * xi + Zi * xr + xr * xi) + dc_i;
                xr = t;
            }
            m += 1;
            n += 1;
        }
        double physical_r = scale * xr, physical_i = scale * xi;
        double fr = zr[m] + physical_r, fi = zi[m] + physical_i;
        double f2 = fr * fr + fi * fi;
        if (f2 > r2) {
            out.kind = FDO_ESCAPED;
            out.n = n;
            out.zr = fr;
            out.zi = fi;
            out.dr = dr;
            out.di = di;
            out.dexp = derivative_exp;
            return out;
        }
        if (nucleus && n == nucleus + 1) {
            if (con_small(&con)) {
                out.kind = FDO_INTERIOR;
                out.n = n;
                return out;
            }
            nucleus = 0;
        }
        double delta2 = physical_r * physical_r + physical_i * physical_i;
        int rebase = EXT ? f2 <= delta2 : f2 < delta2;
        if (rebase || m == last) {
            if (EXT) {
                if (f2 > SQUARE_OK) {
                    xr = fr;
                    xi = fi;
                    delta_exp = 0;
                } else {
                    double t[2][2] = {{zr[m], zi[m]}, {xr, xi}};
                    int64_t te[2] = {0, delta_exp};
                    sum2(t, te, delta_exp, &xr, &xi, &delta_exp);
                }
                scale = exp2i(delta_exp);
                m = 0;
                if (scale < TINY) {
                    nucleus = 0;
                    if (D) {
                        zero_step_dz(xr, xi, delta_exp, dr, di, derivative_exp, &dr, &di,
                                     &derivative_exp);
                        inverse = exp2i(-derivative_exp);
                    }
                    zero_step(xr, xi, delta_exp, ar, ai, offset_exp, &xr, &xi, &delta_exp);
                    scale = exp2i(delta_exp);
                    m = 1;
                    n = n + 1;
                }
                double factor = exp2i(offset_exp - delta_exp);
                dc_r = ar * factor;
                dc_i = ai * factor;
            } else {
                xr = fr;
                xi = fi;
                m = 0;
            }
        } else if (EXT) {
            double magnitude = fmax(fabs(xr), fabs(xi));
            int64_t k = 0;
            if (magnitude > HI && delta_exp < 0)
                k = -delta_exp < 32 ? -delta_exp : 32;
            else if (magnitude < LO && magnitude != 0.0)
                k = -32;
            if (k != 0) {
                double factor = exp2i(-k);
                xr = xr * factor;
                xi = xi * factor;
                delta_exp = delta_exp + k;
                scale = exp2i(delta_exp);
                double f2c = exp2i(offset_exp - delta_exp);
                dc_r = ar * f2c;
                dc_i = ai * f2c;
            }
        }
        double check_r = EXT ? physical_r : xr, check_i = EXT ? physical_i : xi;
        int judge = EXT ? (returns < 4 && resolvable(check_r, check_i, f2))
                        : resolvable(check_r, check_i, f2);
        if (!judge) {
            if (EXT) {
                if (n >= chk) {
                    sr = fr, si = fi, sn = n, chk = chk * 2;
                }
            } else if (hold > 0) {
                if (n >= sn + hold) {
                    hold = 0, predicted = 0, pending = 1;
                }
            } else if (n >= chk) {
                pending = 1;
            }
            continue;
        }
        double dist = fabs(fr - sr) + fabs(fi - si);
        double size = fabs(sr) + fabs(si);
        if (EXT && returns < 4 && dist < 1e-13 * size + 1e-300) {
            if (contracting(zr, zi, len, m, check_r, check_i, ar * exp2i(offset_exp),
                            ai * exp2i(offset_exp), n - sn)) {
                out.kind = FDO_UNRESOLVED;
                return out;
            }
            returns += 1;
        }
        if (!EXT) {
            if (dist < 1e-3 * size) {
                uint64_t period = n - sn;
                if (hold == period && mul_below_one(&mu)) {
                    double mr, mi;
                    mul_value(&mu, &mr, &mi);
                    double gr = fr - sr, gi = fi - si;
                    double er = mr - 1.0, ei = mi;
                    double den = er * er + ei * ei;
                    double pr = sr - (gr * er + gi * ei) / den;
                    double pi = si - (gi * er - gr * ei) / den;
                    double tol = fmin(0.1 * sqrt(den), 1e-2) * qstep + 16.0 * U_RND * size;
                    if (predicted && fabs(pr - qr) + fabs(pi - qi) <= tol) {
                        out.kind = FDO_INTERIOR;
                        out.n = n;
                        return out;
                    }
                    qr = pr, qi = pi, qstep = (fabs(gr) + fabs(gi)) / sqrt(den);
                    predicted = 1;
                } else {
                    predicted = 0;
                }
                hold = period;
                sr = fr, si = fi, sn = n, pending = 0;
                mu.re = 1.0, mu.im = 0.0, mu.exp = 0;
                continue;
            }
            if (hold > 0) {
                if (n < sn + hold) continue;
                hold = 0, predicted = 0, pending = 1;
            }
            if (pending || n >= chk) {
                uint64_t np = next_pow2(n + 1);
                chk = chk > np ? chk : np;
                sr = fr, si = fi, sn = n, pending = 0;
            }
            continue;
        }
        if (n == chk || (EXT && n > chk)) {
            sr = fr, si = fi, sn = n, chk = chk * 2;
        }
    }
    return out;
}

#define FDO_INST(NAME, EXT, D, L)                                                              \
    static outcome_t NAME(const double *zr, const double *zi, uint64_t len, const bla_t *bla,   \
                          skip_t *skip, int have_c, double c_re, double c_im, double ar,       \
                          double ai, int64_t oe, uint64_t mi, double r2, uint64_t nuc,         \
                          const start_t *st) {                                                 \
        return iterate(EXT, D, L, zr, zi, len, bla, skip, have_c, c_re, c_im, ar, ai, oe, mi,  \
                       r2, nuc, st);                                                           \
    }
FDO_INST(it_o00, 0, 0, 0)
FDO_INST(it_o10, 0, 1, 0)
FDO_INST(it_o01, 0, 0, 1)
FDO_INST(it_o11, 0, 1, 1)
FDO_INST(it_e00, 1, 0, 0)
FDO_INST(it_e10, 1, 1, 0)
FDO_INST(it_e01, 1, 0, 1)
FDO_INST(it_e11, 1, 1, 1)

typedef outcome_t (*iter_fn)(const double *, const double *, uint64_t, const bla_t *, skip_t *,
                             int, double, double, double, double, int64_t, uint64_t, double,
                             uint64_t, const start_t *);

static iter_fn pick_iter(int ext, int d, int l) {
    static const iter_fn tab[8] = {it_o00, it_o01, it_o10, it_o11, it_e00, it_e01, it_e10, it_e11};
    return tab[(ext << 2) | (d << 1) | l];
}

int fdo_sample(const double *re, const double *im, uint64_t len, const double *bla_blocks,
               int bla_levels, int have_c, double c_re, double c_im, double ar, double ai,
               int extended, int64_t offset_exp, int derivative, uint64_t max_iter, double r2,
               uint64_t nucleus, uint64_t *n_out, double *z_out, double *d_out,
               int64_t *dexp_out, uint64_t *blocks_out, uint64_t *skipped_out) {
    bla_t t;
    memset(&t, 0, sizeof(t));
    int L = bla_blocks != NULL && bla_levels > 0;
    if (L) {
        uint64_t steps = len - 2, off = 0;
        t.nlevels = bla_levels;
        t.levels = (block_t **)malloc(bla_levels * sizeof(block_t *));
        t.r2 = (double **)malloc(bla_levels * sizeof(double *));
        t.count = (uint64_t *)malloc(bla_levels * sizeof(uint64_t));
        for (int i = 0; i < bla_levels; i++) {
            uint64_t c = steps >> (i + 1);
            t.count[i] = c;
            t.levels[i] = (block_t *)malloc((c ? c : 1) * sizeof(block_t));
            t.r2[i] = (double *)malloc((c ? c : 1) * sizeof(double));
            memcpy(t.levels[i], bla_blocks + 7 * off, c * sizeof(block_t));
            for (uint64_t k = 0; k < c; k++) t.r2[i][k] = t.levels[i][k].r * t.levels[i][k].r;
            off += c;
        }
    }
    skip_t sk = {0, 0};
    outcome_t o = pick_iter(extended, derivative, L)(re, im, len, L ? &t : NULL, &sk, have_c, c_re,
                                                    c_im, ar, ai, offset_exp, max_iter, r2, nucleus, NULL);
    if (L) bla_free(&t);
    if (n_out) *n_out = o.n;
    if (z_out) z_out[0] = o.zr, z_out[1] = o.zi;
    if (d_out) d_out[0] = o.dr, d_out[1] = o.di;
    if (dexp_out) *dexp_out = o.dexp;
    if (blocks_out) *blocks_out = sk.blocks;
    if (skipped_out) *skipped_out = sk.skipped;
    return o.kind;
}

/* fdo_sample over many offsets (ar[i], ai[i]) against one table, built once: outputs are
 * arrays of count entries (kind as the return value of fdo_sample). */
void fdo_sample_many(const double *re, const double *im, uint64_t len, const double *bla_blocks,
                     int bla_levels, double c_re, double c_im, int have_c, const double *ar,
                     const double *ai, uint64_t count, int extended, int64_t offset_exp,
                     int derivative, uint64_t max_iter, double r2, uint64_t nucleus,
                     uint8_t *kind_out, uint64_t *n_out, double *z_out, double *d_out,
                     int64_t *dexp_out) {
    bla_t t;
    memset(&t, 0, sizeof(t));
    int L = bla_blocks != NULL && bla_levels > 0;
    if (L) {
        uint64_t steps = len - 2, off = 0;
        t.nlevels = bla_levels;
        t.levels = (block_t **)malloc(bla_levels * sizeof(block_t *));
        t.r2 = (double **)malloc(bla_levels * sizeof(double *));
        t.count = (uint64_t *)malloc(bla_levels * sizeof(uint64_t));
        for (int i = 0; i < bla_levels; i++) {
            uint64_t c = steps >> (i + 1);
            t.count[i] = c;
            t.levels[i] = (block_t *)malloc((c ? c : 1) * sizeof(block_t));
            t.r2[i] = (double *)malloc((c ? c : 1) * sizeof(double));
            memcpy(t.levels[i], bla_blocks + 7 * off, c * sizeof(block_t));
            for (uint64_t k = 0; k < c; k++) t.r2[i][k] = t.levels[i][k].r * t.levels[i][k].r;
            off += c;
        }
    }
    for (uint64_t i = 0; i < count; i++) {
        skip_t sk = {0, 0};
        outcome_t o = pick_iter(extended, derivative, L)(re, im, len, L ? &t : NULL, &sk, have_c,
                                                        c_re, c_im, ar[i], ai[i], offset_exp,
                                                        max_iter, r2, nucleus, NULL);
        kind_out[i] = (uint8_t)o.kind;
        n_out[i] = o.n;
        z_out[2 * i] = o.zr, z_out[2 * i + 1] = o.zi;
        d_out[2 * i] = o.dr, d_out[2 * i + 1] = o.di;
        dexp_out[i] = o.dexp;
    }
    if (L) bla_free(&t);
}

/* ------------------------------------------------- certified error control (GPU route) */

/* reference.rs:122-148 Reference::error_radius: q[m] >= |A_m - Z_m| for the stored orbit Z
 * and the exact orbit A of the exact view centre; bits < 0: an f64 orbit from (c_re, c_im),
 * the correctly rounded centre, else the fixed-point fraction bits it was computed with. */
void fdo_error_radius(const double *re, const double *im, uint64_t len, int64_t bits,
                      double c_re, double c_im, double *q) {
    const double U = DBL_EPSILON / 2.0;
    double c1 = fabs(c_re) + fabs(c_im);
    double fixed = 0.0;
    if (bits >= 0) {
        fixed = exp2i(4 - bits);
        if (fixed < DBL_MIN) fixed = DBL_MIN; /* f64::max with MIN_POSITIVE */
    }
    double a = 0.0;
    for (uint64_t m = 0; m < len; m++) {
        double s = sqrt(re[m] * re[m] + im[m] * im[m]) * (1.0 + 4.0 * U);
        double round, defect;
        if (bits >= 0) {
            round = 2.0 * U * s + DBL_MIN;
            defect = fixed;
        } else {
            round = 0.0;
            defect = 8.0 * U * (s * s + c1) + 2.0 * U * c1 + DBL_MIN;
        }
        q[m] = a + round;
        a = (a * (2.0 * (s + round) + a) + defect) * (1.0 + 8.0 * U);
    }
}

/* bla/guarded.rs:9-18 norm, up, remainder */
static double g_norm(double re, double im) { return fabs(re) + fabs(im); }
static double g_up(double x) { return x * (1.0 + 64.0 * DBL_EPSILON) + DBL_MIN; }

typedef struct {
    block_t b;
    double dalpha, dbeta;
} gblock_t;

/* guarded.rs:34-50 GuardedBlock::single */
static gblock_t gblock_single(double zr, double zi, double q, double eps) {
    gblock_t g;
    memset(&g, 0, sizeof(g));
    block_t b = block_single(zr, zi, eps);
    if (b.r == 0.0 || !isfinite(q) || q < 0.0 || g_norm(zr, zi) + q > 2.0) return g;
    double da = g_up(2.0 * b.r + 2.0 * q);
    b.alpha = g_up(b.r + 2.0 * q);
    g.b = b;
    g.dalpha = da;
    g.dbeta = 0.0;
    return g;
}

/* guarded.rs:51-97 GuardedBlock::merge (x then y, |dc| <= c) */
static gblock_t gblock_merge(const gblock_t *dx, const gblock_t *dy, double c) {
    gblock_t zero;
    memset(&zero, 0, sizeof(zero));
    const block_t *x = &dx->b, *y = &dy->b;
    if (x->r == 0.0 || y->r == 0.0) return zero;
    double ax = g_norm(x->ar, x->ai), ay = g_norm(y->ar, y->ai), bx = g_norm(x->br, x->bi);
    double r = fmin(x->r, (y->r - g_up((bx + x->beta) * c)) / g_up(ax + x->alpha)) *
               (1.0 - 64.0 * DBL_EPSILON);
    gblock_t z;
    /* affine(x.a, y.a, 0) and affine(x.b, y.a, y.b), bla.rs:124-129 */
    z.b.ar = y->ar * x->ar - y->ai * x->ai + 0.0 + 0.0;
    z.b.ai = y->ar * x->ai + y->ai * x->ar + 0.0 + 0.0;
    z.b.br = y->ar * x->br - y->ai * x->bi + y->br + 0.0;
    z.b.bi = y->ar * x->bi + y->ai * x->br + y->bi + 0.0;
    double ra = g_up(32.0 * DBL_EPSILON * ay * ax);
    double rb = g_up(32.0 * DBL_EPSILON * (ay * bx + g_norm(y->br, y->bi)));
    /* remainder(ay, ax, bx, xa, xb, ya, yb) = (up(ay xa + ya (ax + xa)),
     *                                          up(ay xb + ya (bx + xb) + yb)) */
    double alpha = g_up(ay * x->alpha + y->alpha * (ax + x->alpha));
    double beta = g_up(ay * x->beta + y->alpha * (bx + x->beta) + y->beta);
    double da = g_up(ay * dx->dalpha + dy->dalpha * (ax + dx->dalpha));
    double db = g_up(ay * dx->dbeta + dy->dalpha * (bx + dx->dbeta) + dy->dbeta);
    z.b.r = r;
    z.b.alpha = g_up(alpha + ra);
    z.b.beta = g_up(beta + rb);
    z.dalpha = g_up(da + ra);
    z.dbeta = g_up(db + rb);
    if (r > 0.0 && isfinite(z.b.ar) && isfinite(z.b.ai) && isfinite(z.b.br) && isfinite(z.b.bi) &&
        isfinite(z.b.r) && isfinite(z.b.alpha) && isfinite(z.b.beta) && isfinite(z.dalpha) &&
        isfinite(z.dbeta))
        return z;
    return zero;
}

/* guarded.rs:103-134 GuardedBla::build: levels 1.. of guarded blocks, 9 doubles each
 * (a.re a.im b.re b.im r alpha beta dalpha dbeta), level-major into out (cap blocks);
 * returns the level count, or -1 for a refused request (bla.rs check, non-finite or
 * negative radii). */
int fdo_guarded_build(const double *re, const double *im, const double *q, uint64_t len,
                      double eps, double dc_max, double *out, uint64_t cap) {
    if (!(eps > 0.0 && eps <= 1.0 / 16777216.0 && dc_max > 0.0 && isfinite(dc_max))) return -1;
    for (uint64_t k = 0; k < len; k++)
        if (!isfinite(q[k]) || q[k] < 0.0) return -1;
    uint64_t n = len >= 2 ? len - 2 : 0;
    gblock_t *next = (gblock_t *)malloc((n ? n : 1) * sizeof(gblock_t));
    for (uint64_t k = 1; k <= n; k++) next[k - 1] = gblock_single(re[k], im[k], q[k], eps);
    int levels = 0;
    uint64_t w = 0;
    for (;;) {
        uint64_t m = n / 2;
        int any = 0;
        for (uint64_t i = 0; i < m; i++) {
            next[i] = gblock_merge(&next[2 * i], &next[2 * i + 1], dc_max);
            any |= next[i].b.r > 0.0;
        }
        n = m;
        if (!any) break;
        for (uint64_t i = 0; i < n; i++, w++) {
            if (w >= cap) continue;
            double *o = out + 9 * w;
            o[0] = next[i].b.ar, o[1] = next[i].b.ai, o[2] = next[i].b.br, o[3] = next[i].b.bi;
            o[4] = next[i].b.r, o[5] = next[i].b.alpha, o[6] = next[i].b.beta;
            o[7] = next[i].dalpha, o[8] = next[i].dbeta;
        }
        levels++;
    }
    free(next);
    return levels;
}

static float de_px(double az, double log2z, double dr, double di, int64_t dexp, double px_m,
                   int64_t px_e);

/* store.rs:119-139 bounds: (lo, bound) or 0 when the radii do not prove escape. */
static int escape_bounds(uint64_t n, double az, double ez, double ad, double ed, double *lo_out,
                         double *bound_out) {
    const double U = DBL_EPSILON / 2.0;
    double lo = az - ez;
    if (!(lo > 2.0 && ed < 0.5 * ad)) return 0;
    double lnz = log(az), llo = log(lo), lhi = log(az + ez);
    double nu = log1p(-log1p(-ez / az) / llo) / 0.693147180559945309417232121458176568 +
                8.0 * U * ((double)n + 8.0);
    double g = lhi * exp2(-(double)n);
    double de_lo = (lo * llo) / (az * lnz) * ad / (ad + ed) * exp(-g) * (1.0 - 4.0 / (lo * lo)) *
                   (1.0 - 1e-6);
    double de_hi = ((az + ez) * lhi) / (az * lnz) * ad / (ad - ed) * (1.0 + 1e-6);
    *lo_out = de_lo;
    *bound_out = nu * de_hi / de_lo * (1.0 + 1e-6);
    return 1;
}

/* store.rs:155-198 external_escape_ok: admit an externally computed escaped state with
 * radii ez (z units) and ed (derivative mantissa units) against output pixel px_m 2^px_e. */
int fdo_external_escape_ok(uint64_t n, double zr, double zi, double dr, double di, int64_t dexp,
                           double ez, double ed, double px_m, int64_t px_e) {
    const double LN2 = 0.693147180559945309417232121458176568;
    const double TAU = 6.283185307179586476925286766559005768;
    double az = hypot(zr, zi), ad = hypot(dr, di);
    if (!(ez >= 0.0 && ed >= 0.0 && isfinite(ez) && isfinite(ed) && isfinite(az) &&
          isfinite(ad)))
        return 0;
    double lo, bound;
    if (!escape_bounds(n, az, ez, ad, ed, &lo, &bound)) return 0;
    double de = (double)de_px(az, log2(az), dr, di, dexp, px_m, px_e);
    double displacement = bound * de * lo * LN2 / 2.0;
    double base = az * log(az);
    double de_hi = (az + ez) * log(az + ez) / base * ad / (ad - ed);
    double de_lo = (az - ez) * log(az - ez) / base * ad / (ad + ed);
    double angle = asin(ez / az) + asin(ed / ad);
    return isfinite(displacement) && displacement <= 5e-4 && isfinite(de) &&
           de_hi - 1.0 + (double)FLT_EPSILON <= 1e-3 && 1.0 - de_lo + (double)FLT_EPSILON <= 1e-3 &&
           angle + TAU / 65536.0 <= 0.1 * (3.141592653589793238462643383279502884 / 180.0);
}

/* ---------------------------------------------------------------- series.rs */

/* series.rs:99-234 Series::build_stalled: the growing-order parameter series
 * S_n(x) = sum b[k] x^k with dc = rho x, b[1]' = 2 Z b[1] + rho, b[k]' = 2 Z b[k] + sum(i+j=k)
 * b[i] b[j], its state / derivative majorants E, F and the admission guard. */
int fdo_series_build(const double *re, const double *im, uint64_t len, double rho, double pixel,
                     uint64_t max_iter, const fdo_series_config *cfg, uint64_t stalled_steps,
                     fdo_series *out) {
    uint32_t k_max = cfg->max_order < 1 ? 1 : cfg->max_order;
    uint32_t active = cfg->start_order < 1 ? 1 : cfg->start_order;
    if (active > k_max) active = k_max;
    if (out->b_cap < k_max) return -1;
    double *br = (double *)calloc(k_max + 1, sizeof(double));
    double *bi = (double *)calloc(k_max + 1, sizeof(double));
    double *sr = (double *)calloc(2 * k_max + 1, sizeof(double));
    double *si = (double *)calloc(2 * k_max + 1, sizeof(double));
    double *best_b = (double *)calloc(2 * k_max, sizeof(double));
    double e = 0.0, f = 0.0;
    out->n = 0;
    out->rho = rho;
    out->points = len;
    out->b_len = 0;
    out->state_bound = out->derivative_bound = out->shift_px = 0.0;
    out->built_steps = 0;
    out->stop = FDO_STOP_REFERENCE;
    out->coefficient_pairs = 0;
    out->growth_steps = 0;
    uint64_t a = len >= 2 ? len - 2 : 0, b = max_iter >= 1 ? max_iter - 1 : 0;
    uint64_t limit = a < b ? a : b;
    int have_best = 0;
    uint64_t best_n = 0;
    uint32_t best_len = 0;
    double best_e = 0, best_f = 0, budget_e = 0, budget_f = 0, best_dmin = 0;
    if (!(rho > 0.0 && isfinite(rho) && pixel > 0.0)) {
        out->stop = FDO_STOP_RADIUS;
        goto done;
    }
    double cap_e = cfg->shift_px * pixel * 2.0 / rho, cap_f = cfg->derivative_rel * 2.0;
    for (uint64_t n = 0; n < limit; n++) {
        double zr = re[n], zi = im[n];
        double az = hypot(zr, zi);
        for (uint32_t k = 0; k <= 2 * active; k++) sr[k] = si[k] = 0.0;
        for (uint32_t i = 1; i <= active; i++) {
            double ar = br[i], ai = bi[i];
            if (ar == 0.0 && ai == 0.0) continue;
            sr[2 * i] += ar * ar - ai * ai;
            si[2 * i] += 2.0 * ar * ai;
            ar = 2.0 * ar;
            ai = 2.0 * ai;
            for (uint32_t j = i + 1; j <= active; j++) {
                sr[i + j] += ar * br[j] - ai * bi[j];
                si[i + j] += ar * bi[j] + ai * br[j];
            }
        }
        out->coefficient_pairs += (uint64_t)active * (active + 1) / 2;
        double m = 0.0, d = 0.0, t = 0.0, td = 0.0;
        for (uint32_t k = 1; k <= active; k++) {
            double v = hypot(br[k], bi[k]);
            m += v;
            d += (double)k * v;
        }
        for (uint32_t k = active + 1; k <= 2 * active; k++) {
            double v = hypot(sr[k], si[k]);
            t += v;
            td += (double)k * v;
        }
        double e_next = 2.0 * (az + m) * e + e * e + t;
        f = 2.0 * (az + m + e) * f + 2.0 * e * d + td;
        e = e_next;
        for (uint32_t k = 1; k <= active; k++) {
            double xr = br[k], xi = bi[k];
            br[k] = 2.0 * (zr * xr - zi * xi) + sr[k];
            bi[k] = 2.0 * (zr * xi + zi * xr) + si[k];
        }
        br[1] += rho;
        uint64_t step = n + 1;
        out->built_steps = step;
        zr = re[step];
        zi = im[step];
        double tail_d = 0.0;
        m = 0.0;
        for (uint32_t k = 1; k <= active; k++) {
            double v = hypot(br[k], bi[k]);
            m += v;
            if (k >= 2) tail_d += (double)k * v;
        }
        if (!(isfinite(e) && isfinite(f))) {
            out->stop = FDO_STOP_NONFINITE;
            break;
        }
        if (hypot(zr, zi) + m + e >= 2.0) {
            out->stop = FDO_STOP_ESCAPE;
            break;
        }
        double dmin = hypot(br[1], bi[1]) - tail_d - f;
        double budget = dmin > 0.0 ? cfg->shift_px * pixel * dmin / rho : 0.0;
        if (dmin > 0.0 && e <= budget && f <= cfg->derivative_rel * dmin) {
            have_best = 1;
            best_n = step;
            best_len = active;
            for (uint32_t k = 1; k <= active; k++) {
                best_b[2 * (k - 1)] = br[k];
                best_b[2 * (k - 1) + 1] = bi[k];
            }
            best_e = e;
            best_f = f;
            budget_e = budget;
            budget_f = cfg->derivative_rel * dmin;
            best_dmin = dmin;
        } else if (e > cap_e || f > cap_f) {
            out->stop = FDO_STOP_BUDGET;
            break;
        }
        if (active < k_max && e > cfg->grow_at * budget) {
            active = 2 * active < k_max ? 2 * active : k_max;
            out->growth_steps++;
        }
        if (stalled_steps > 0 && have_best && step - best_n >= stalled_steps) {
            out->stop = FDO_STOP_STALL;
            break;
        }
    }
    if (have_best) {
        /* series.rs:215-229: trim high orders that fit in half the remaining slack */
        double slack_e = budget_e - best_e, slack_f = budget_f - best_f;
        double te = 0.0, tf = 0.0;
        while (best_len > 1) {
            uint32_t k = best_len;
            double v = hypot(best_b[2 * (k - 1)], best_b[2 * (k - 1) + 1]);
            if (te + v > 0.5 * slack_e || tf + (double)k * v > 0.5 * slack_f) break;
            te += v;
            tf += (double)k * v;
            best_len--;
        }
        out->n = best_n;
        out->b_len = best_len;
        memcpy(out->b, best_b, 2 * best_len * sizeof(double));
        out->state_bound = best_e + te;
        out->derivative_bound = best_f + tf;
        out->shift_px = out->state_bound * rho / (best_dmin * pixel);
    }
done:
    free(br);
    free(bi);
    free(sr);
    free(si);
    free(best_b);
    return 0;
}

/* series.rs:236-257 Series::eval: Horner on x = dc / rho for S = x p(x) and S' = p + x p'. */
void fdo_series_eval(const fdo_series *s, int derivative, double ar, double ai, double out[4]) {
    double sc = 1.0 / s->rho;
    double xr = ar * sc, xi = ai * sc;
    double pr = 0.0, pi = 0.0, qr = 0.0, qi = 0.0;
    for (uint32_t k = s->b_len; k-- > 0;) {
        double cr = s->b[2 * k], ci = s->b[2 * k + 1];
        if (derivative) {
            double tr = qr * xr - qi * xi + pr;
            qi = qr * xi + qi * xr + pi;
            qr = tr;
        }
        double tr = pr * xr - pi * xi + cr;
        pi = pr * xi + pi * xr + ci;
        pr = tr;
    }
    out[0] = pr * xr - pi * xi;
    out[1] = pr * xi + pi * xr;
    if (!derivative) {
        out[2] = out[3] = 0.0;
        return;
    }
    double dr = pr + qr * xr - qi * xi, di = pi + qr * xi + qi * xr;
    out[2] = dr * sc;
    out[3] = di * sc;
}

/* ---------------------------------------------------------------- store.rs */

/* samples.rs:60-65 Samples::angle */
static uint16_t sample_angle(double x, double y) {
    double t = atan2(y, x) * (65536.0 / 6.283185307179586);
    double r = round(t);
    if (isnan(r)) return 0;
    int64_t v = r >= 9.2233720368547758e18 ? INT64_MAX : (r <= -9.2233720368547758e18 ? INT64_MIN : (int64_t)r);
    return (uint16_t)(uint64_t)v;
}

/* store.rs:141-150 de_px */
static float de_px(double az, double log2z, double dr, double di, int64_t dexp, double px_m,
                   int64_t px_e) {
    double lnz = log2z * 0.693147180559945309417232121458176568;
    double base = 2.0 * az * lnz / (hypot(dr, di) * px_m);
    int64_t k = -(dexp + px_e);
    if (k < -4000) k = -4000;
    if (k > 4000) k = 4000;
    return (float)(base * exp2i(k / 2) * exp2i(k - k / 2));
}

/* store.rs:66-112 Store::put (heuristic evidence: no Bound column) */
static void store_put(const fdo_plane *pl, uint32_t ss, const fdo_columns *c, size_t idx,
                      const outcome_t *o) {
    if (o->kind != FDO_ESCAPED) {
        if (c->class_) c->class_[idx] = (uint8_t)(o->kind == FDO_UNRESOLVED ? 2 : 1);
        return;
    }
    double z2 = o->zr * o->zr + o->zi * o->zi;
    double log2z = 0.5 * log2(z2);
    if (c->class_) c->class_[idx] = 0;
    if (c->nu) c->nu[idx] = (double)o->n + 1.0 - log2(log2z);
    if (c->de) c->de[idx] = de_px(sqrt(z2), log2z, o->dr, o->di, o->dexp, pl->h_m * (double)ss, pl->h_e);
    if (c->normal) {
        double a = o->zr * o->dr + o->zi * o->di, b = o->zi * o->dr - o->zr * o->di;
        double x = pl->cos_r * a + pl->sin_r * b, y = pl->sin_r * a - pl->cos_r * b;
        c->normal[idx] = sample_angle(x, y);
    }
}

/* ---------------------------------------------------------------- grid.rs */

typedef struct {
    const fdo_plane *pl;
    int tier, ext, deriv, L;
    const double *zr, *zi;
    uint64_t len;
    const bla_t *bla;
    uint64_t max_iter, nucleus;
    double r2, h;
    uint32_t nx, ny, ss;
    const uint8_t *mask;
    const fdo_series *series; /* admitted series start (n > 0), or NULL */
    fdo_outcomes *oc;
    const fdo_columns *cols;
    pthread_mutex_t lock;
    uint32_t next_row;
    uint64_t iterations, equivalent, blocks, skipped, fallback, closed;
} job_t;

/* grid.rs:652-725 one_from: the sample's kernel call from start `st` */
static outcome_t one_from(const job_t *job, iter_fn f, double ux, double uy, skip_t *sk,
                          const start_t *st) {
    const fdo_plane *pl = job->pl;
    if (job->tier == FDO_TIER_F64) {
        double ar = ux * job->h, ai = uy * job->h;
        return f(job->zr, job->zi, job->len, job->bla, sk, 1, pl->c_re + ar, pl->c_im + ai, ar,
                 ai, 0, job->max_iter, job->r2, job->nucleus, st);
    }
    if (job->tier == FDO_TIER_FIXED)
        return f(job->zr, job->zi, job->len, job->bla, sk, 0, 0, 0, ux * job->h, uy * job->h, 0,
                 job->max_iter, job->r2, job->nucleus, st);
    return f(job->zr, job->zi, job->len, job->bla, sk, 0, 0, 0, ux * pl->h_m, uy * pl->h_m,
             pl->h_e, job->max_iter, job->r2, job->nucleus, st);
}

/* grid.rs:623-649 one + grid.rs:748-775 fill */
static void fill_row(job_t *job, uint32_t j, uint64_t acc[6]) {
    iter_fn f = pick_iter(job->ext, job->deriv, job->L);
    const fdo_plane *pl = job->pl;
    for (uint32_t i = 0; i < job->nx; i++) {
        size_t idx = (size_t)j * job->nx + i;
        if (job->mask && !job->mask[idx]) continue;
        double ux, uy;
        unit_offset(pl, i, j, &ux, &uy);
        skip_t sk = {0, 0};
        start_t st;
        if (job->series) {
            /* grid.rs:663-673: the series state at the sample's offset (ordinary tiers) */
            double v[4];
            fdo_series_eval(job->series, job->deriv, ux * job->h, uy * job->h, v);
            st.n = job->series->n;
            st.xr = v[0];
            st.xi = v[1];
            st.dr = v[2];
            st.di = v[3];
        }
        outcome_t o = one_from(job, f, ux, uy, &sk, job->series ? &st : NULL);
        /* grid.rs:630-642: the prefix is not counted as work; an Unresolved series start
         * retries from zero (keeping the nucleus heuristic the start declined) */
        uint64_t prefix = job->series && !(o.kind == FDO_INTERIOR && o.n == 0) ? job->series->n : 0;
        uint64_t its = (o.kind == FDO_UNRESOLVED ? job->max_iter : o.n) - prefix;
        if (prefix > 0 && o.kind == FDO_UNRESOLVED) {
            o = one_from(job, f, ux, uy, &sk, NULL);
            its += o.kind == FDO_UNRESOLVED ? job->max_iter : o.n;
        }
        acc[1] += its;
        if (job->L) {
            its = its - sk.skipped + sk.blocks;
            acc[2] += sk.blocks;
            acc[3] += sk.skipped;
            int closed = o.kind == FDO_INTERIOR && o.n == 0;
            acc[5] += closed;
            acc[4] += sk.blocks == 0 && !closed;
        }
        acc[0] += its;
        if (job->oc) {
            if (job->oc->kind) job->oc->kind[idx] = (uint8_t)o.kind;
            if (job->oc->n) job->oc->n[idx] = o.n;
            if (job->oc->zr) job->oc->zr[idx] = o.zr;
            if (job->oc->zi) job->oc->zi[idx] = o.zi;
            if (job->oc->dr) job->oc->dr[idx] = o.dr;
            if (job->oc->di) job->oc->di[idx] = o.di;
            if (job->oc->dexp) job->oc->dexp[idx] = o.dexp;
        }
        if (job->cols) store_put(pl, job->ss, job->cols, idx, &o);
    }
}

/* grid.rs:494-519 worker pulling rows from a shared queue */
static void *worker(void *arg) {
    job_t *job = (job_t *)arg;
    uint64_t acc[6] = {0, 0, 0, 0, 0, 0};
    for (;;) {
        pthread_mutex_lock(&job->lock);
        uint32_t j = job->next_row < job->ny ? job->next_row++ : job->ny;
        pthread_mutex_unlock(&job->lock);
        if (j >= job->ny) break;
        fill_row(job, j, acc);
    }
    pthread_mutex_lock(&job->lock);
    job->iterations += acc[0];
    job->equivalent += acc[1];
    job->blocks += acc[2];
    job->skipped += acc[3];
    job->fallback += acc[4];
    job->closed += acc[5];
    pthread_mutex_unlock(&job->lock);
    return NULL;
}

/* grid.rs:442-444 dc_max and grid.rs:423-430 dc_bound */
static void dc_bound(const fdo_plane *pl, int tier, uint32_t nx, uint32_t ny, double *dc,
                     int64_t *e) {
    double half = hypot((double)nx / 2.0, (double)ny / 2.0);
    if (tier == FDO_TIER_SCALED) {
        *dc = pl->h_m * half;
        *e = pl->h_e;
    } else {
        *dc = plane_h(pl) * half;
        *e = 0;
    }
}

/* grid.rs:779-799 setup */
static int setup(const char *re, const char *im, const char *width, double rot, uint32_t nx,
                 uint32_t ny, uint32_t ss, int tier_req, fdo_plane *pl, int *tier, char *err,
                 size_t errcap) {
    if (ss == 0 || nx % ss != 0 || ny % ss != 0) {
        set_err(err, errcap, "grid must be a whole number of pixels");
        return -1;
    }
    if (fdo_plane_new(re, im, width, rot, nx, ny, pl, err, errcap) != 0) return -1;
    int t = tier_req < 0 ? pl->tier : tier_req;
    int allows = t == FDO_TIER_F64 ? pl->tier == FDO_TIER_F64
                                   : (t == FDO_TIER_FIXED ? pl->tier != FDO_TIER_SCALED : 1);
    if (!allows) {
        set_err(err, errcap, "tier is not valid at this depth");
        return -1;
    }
    *tier = t;
    return 0;
}

/* grid.rs:802-815 compute */
static int compute_reference(const char *re, const char *im, const fdo_plane *pl, int tier,
                             uint64_t max_iter, double radius, orbit_t *o) {
    memset(o, 0, sizeof(*o));
    if (tier == FDO_TIER_F64) {
        reference_f64(o, pl->c_re, pl->c_im, max_iter, radius);
        return 0;
    }
    return reference_fixed(o, re, im, pl->bits, max_iter);
}

int64_t fdo_reference(const char *re, const char *im, const char *width, double rotation,
                      uint32_t nx, uint32_t ny, uint64_t max_iter, double escape_radius,
                      int tier, double *out_re, double *out_im, uint64_t cap,
                      uint32_t *bits, char *err, size_t errcap) {
    fdo_plane pl;
    int t;
    if (setup(re, im, width, rotation, nx, ny, 1, tier, &pl, &t, err, errcap) != 0) return -1;
    orbit_t o;
    if (compute_reference(re, im, &pl, t, max_iter, escape_radius, &o) != 0) {
        set_err(err, errcap, "bad centre");
        return -1;
    }
    if (bits) *bits = t == FDO_TIER_F64 ? 53 : (uint32_t)pl.bits;
    uint64_t n = o.len < cap ? o.len : cap;
    if (out_re) memcpy(out_re, o.re, n * 8);
    if (out_im) memcpy(out_im, o.im, n * 8);
    free(o.re);
    free(o.im);
    return (int64_t)o.len;
}

/* grid.rs:446-535 run (+ control.rs:356-366 per-frame table for bla = 1) */
/* grid.rs:446-477 run: the supplied-orbit checks and plane.bits = supplied bits. */
static int supplied_orbit(fdo_plane *pl, int tier, const double *re, const double *im,
                          uint64_t len, uint32_t bits, orbit_t *o, char *err, size_t errcap) {
    uint64_t need = tier == FDO_TIER_F64 ? 53 : pl->bits;
    if (bits != need && (tier == FDO_TIER_F64 || bits < need)) {
        char m[160];
        snprintf(m, sizeof(m), "supplied reference has %u bits; this view needs %s%llu", bits,
                 tier == FDO_TIER_F64 ? "" : "at least ", (unsigned long long)need);
        set_err(err, errcap, m);
        return -1;
    }
    if (len == 0 || re[0] != 0.0 || im[0] != 0.0) {
        set_err(err, errcap,
                "supplied reference orbit must start at Z_0 = 0 with equal re/im lengths");
        return -1;
    }
    o->re = (double *)malloc(len * sizeof(double));
    o->im = (double *)malloc(len * sizeof(double));
    memcpy(o->re, re, len * sizeof(double));
    memcpy(o->im, im, len * sizeof(double));
    o->len = o->cap = len;
    pl->bits = bits;
    return 0;
}

int fdo_render(const char *re, const char *im, const char *width, double rotation,
               uint32_t nx, uint32_t ny, uint32_t ss, uint64_t max_iter, double escape_radius,
               uint32_t columns, int tier_req, int bla, double bla_eps, int threads,
               const uint8_t *mask, fdo_outcomes *oc, fdo_columns *cols, fdo_stats *stats,
               char *kernel_id, size_t kcap, char *err, size_t errcap) {
    return fdo_render_with(re, im, width, rotation, nx, ny, ss, max_iter, escape_radius, columns,
                           tier_req, bla, bla_eps, NULL, NULL, 0, 0, 0.0, 0, threads, mask, oc,
                           cols, stats, kernel_id, kcap, err, errcap);
}

static int render_impl(const char *re, const char *im, const char *width, double rotation,
                       uint32_t nx, uint32_t ny, uint32_t ss, uint64_t max_iter,
                       double escape_radius, uint32_t columns, int tier_req, int bla,
                       double bla_eps, const double *orbit_re, const double *orbit_im,
                       uint64_t orbit_len, uint32_t orbit_bits, double table_dc,
                       int64_t table_dc_exp, const fdo_series *series, int threads,
                       const uint8_t *mask, fdo_outcomes *oc, fdo_columns *cols,
                       fdo_stats *stats, char *kernel_id, size_t kcap, char *err,
                       size_t errcap) {
    fdo_plane pl;
    int tier;
    if (setup(re, im, width, rotation, nx, ny, ss, tier_req, &pl, &tier, err, errcap) != 0)
        return -1;
    if (columns & FDO_COL_BOUND) {
        set_err(err, errcap, "the Bound column is not restated by this oracle");
        return -1;
    }
    if (series && tier == FDO_TIER_SCALED) {
        /* series.rs:353-356 */
        set_err(err, errcap, "a series start needs f64 deltas and no Bound column");
        return -1;
    }
    double t0 = now_s();
    orbit_t o;
    if (orbit_re) {
        if (supplied_orbit(&pl, tier, orbit_re, orbit_im, orbit_len, orbit_bits, &o, err,
                           errcap) != 0)
            return -1;
    } else if (compute_reference(re, im, &pl, tier, max_iter, escape_radius, &o) != 0) {
        set_err(err, errcap, "bad centre");
        return -1;
    }
    if (series) {
        /* series.rs:357-366: built over this orbit, and a radius covering the view */
        double dc = plane_h(&pl) * hypot((double)nx / 2.0, (double)ny / 2.0);
        const char *why = NULL;
        char m[200];
        if (series->points != o.len) {
            why = "series built over another orbit";
        } else if (isnan(dc) || dc > series->rho) {
            snprintf(m, sizeof(m), "view reaches |dc| = %e, beyond the series radius %e", dc,
                     series->rho);
            why = m;
        }
        if (why) {
            free(o.re);
            free(o.im);
            set_err(err, errcap, why);
            return -1;
        }
    }
    double t1 = now_s();
    bla_t table;
    memset(&table, 0, sizeof(table));
    int use_bla = 0;
    if (bla) {
        double dc;
        int64_t de;
        dc_bound(&pl, tier, nx, ny, &dc, &de);
        if (table_dc > 0.0) {
            /* grid.rs:399-411 bla_checks: the view's |dc| bound against the table's */
            int over = de == table_dc_exp
                           ? dc > table_dc
                           : log2(dc) + (double)de > log2(table_dc) + (double)table_dc_exp;
            if (isnan(dc) || over) {
                free(o.re);
                free(o.im);
                set_err(err, errcap, "view reaches |dc| beyond the BLA table's dc_max");
                return -1;
            }
            dc = table_dc;
            de = table_dc_exp;
        }
        if (bla_build(&table, o.re, o.im, o.len, bla_eps, dc, de) != 0) {
            free(o.re);
            free(o.im);
            set_err(err, errcap, "BLA needs 0 < eps <= 2^-24 and a finite dc_max > 0");
            return -1;
        }
        if (escape_radius < 4.0 || isnan(escape_radius)) {
            bla_free(&table);
            free(o.re);
            free(o.im);
            set_err(err, errcap, "escape radius is below 4");
            return -1;
        }
        use_bla = table.nlevels > 0;
    }
    double t2 = now_s();
    job_t job;
    memset(&job, 0, sizeof(job));
    job.pl = &pl;
    job.tier = tier;
    job.ext = tier == FDO_TIER_SCALED;
    job.deriv = (columns & (FDO_COL_DE | FDO_COL_NORMAL)) != 0;
    job.L = use_bla;
    job.zr = o.re;
    job.zi = o.im;
    job.len = o.len;
    job.bla = use_bla ? &table : NULL;
    job.max_iter = max_iter;
    job.nucleus = fdo_nucleus_period(o.re, o.im, o.len);
    job.r2 = escape_radius * escape_radius;
    job.h = plane_h(&pl);
    job.nx = nx;
    job.ny = ny;
    job.ss = ss;
    job.mask = mask;
    job.series = series && series->n > 0 ? series : NULL;
    job.oc = oc;
    job.cols = cols;
    pthread_mutex_init(&job.lock, NULL);
    int nt = threads < 1 ? 1 : threads;
    if ((uint32_t)nt > ny) nt = (int)ny;
    pthread_t *th = (pthread_t *)malloc(nt * sizeof(pthread_t));
    for (int k = 1; k < nt; k++) pthread_create(&th[k], NULL, worker, &job);
    worker(&job);
    for (int k = 1; k < nt; k++) pthread_join(th[k], NULL);
    free(th);
    pthread_mutex_destroy(&job.lock);
    double t3 = now_s();
    if (stats) {
        memset(stats, 0, sizeof(*stats));
        stats->reference_len = o.len;
        stats->iterations = job.iterations;
        stats->equivalent = job.equivalent;
        stats->blocks = job.blocks;
        stats->skipped = job.skipped;
        stats->fallback_samples = job.fallback;
        stats->closed_form_samples = job.closed;
        stats->bla_levels = (uint32_t)table.nlevels;
        stats->reference_seconds = t1 - t0;
        stats->bla_seconds = t2 - t1;
        stats->render_seconds = t3 - t2;
    }
    if (kernel_id && kcap) {
        /* view.rs:90-96 kernel, grid.rs:527-529 " bla/1" */
        /* series.rs:371-372 " series/0" after it */
        const char *sfx = use_bla ? (series ? " bla/1 series/0" : " bla/1")
                                  : (series ? " series/0" : "");
        if (tier == FDO_TIER_F64)
            snprintf(kernel_id, kcap, "pert-f64/1%s", sfx);
        else if (tier == FDO_TIER_FIXED)
            snprintf(kernel_id, kcap, "pert-fx/1 bits=%llu%s", (unsigned long long)pl.bits, sfx);
        else
            snprintf(kernel_id, kcap, "pert-fx-scaled/1 bits=%llu%s", (unsigned long long)pl.bits, sfx);
    }
    bla_free(&table);
    free(o.re);
    free(o.im);
    return 0;
}

int fdo_render_with(const char *re, const char *im, const char *width, double rotation,
                    uint32_t nx, uint32_t ny, uint32_t ss, uint64_t max_iter,
                    double escape_radius, uint32_t columns, int tier_req, int bla, double bla_eps,
                    const double *orbit_re, const double *orbit_im, uint64_t orbit_len,
                    uint32_t orbit_bits, double table_dc, int64_t table_dc_exp, int threads,
                    const uint8_t *mask, fdo_outcomes *oc, fdo_columns *cols, fdo_stats *stats,
                    char *kernel_id, size_t kcap, char *err, size_t errcap) {
    return render_impl(re, im, width, rotation, nx, ny, ss, max_iter, escape_radius, columns,
                       tier_req, bla, bla_eps, orbit_re, orbit_im, orbit_len, orbit_bits,
                       table_dc, table_dc_exp, NULL, threads, mask, oc, cols, stats, kernel_id,
                       kcap, err, errcap);
}

/* grid.rs:218-226 render_series / series.rs:342-374 Series::render_masked */
int fdo_render_series(const char *re, const char *im, const char *width, double rotation,
                      uint32_t nx, uint32_t ny, uint32_t ss, uint64_t max_iter,
                      double escape_radius, uint32_t columns, int tier_req, int bla,
                      double bla_eps, const fdo_series *series, int threads,
                      const uint8_t *mask, fdo_outcomes *oc, fdo_columns *cols, fdo_stats *stats,
                      char *kernel_id, size_t kcap, char *err, size_t errcap) {
    return render_impl(re, im, width, rotation, nx, ny, ss, max_iter, escape_radius, columns,
                       tier_req, bla, bla_eps, NULL, NULL, 0, 0, 0.0, 0, series, threads, mask,
                       oc, cols, stats, kernel_id, kcap, err, errcap);
}

/* ---------------------------------------------------------------- jetaa.rs */

/* crates/fd-cli/src/jetaa.rs:38-44 first */
void fdo_jetaa_first(uint64_t w, uint64_t h, uint8_t *mask) {
    for (uint64_t i = 0; i < 4 * w * h; i++) mask[i] = (i % (2 * w)) % 2 == 0 && (i / (2 * w)) % 2 == 0;
}

typedef struct { double r, i; } jc;
static jc jc_mul(jc a, jc b) { jc o = {a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r}; return o; }
static jc jc_add(jc a, jc b) { jc o = {a.r + b.r, a.i + b.i}; return o; }
static jc jc_sub(jc a, jc b) { jc o = {a.r - b.r, a.i - b.i}; return o; }
static jc jc_scale(jc a, double k) { jc o = {a.r * k, a.i * k}; return o; }

/* jetaa.rs:63-76 Off */
typedef struct { jc d, d2h, d3s; } joff;
static joff joff_new(double dx, double dy) {
    joff o;
    o.d.r = dx;
    o.d.i = -dy;
    jc d2 = jc_mul(o.d, o.d);
    o.d2h = jc_scale(d2, 0.5);
    o.d3s = jc_scale(jc_mul(d2, o.d), 1.0 / 6.0);
    return o;
}

/* jetaa.rs:78-103 Jet / Jet::at */
typedef struct { int ok; double nu; jc f1, f2, f3; } jjet;
static int jjet_at(const jjet *j, const joff *o, double *nu, float *de, double *x, double *y) {
    double re = 1.0 + jc_mul(j->f1, o->d).r + jc_mul(j->f2, o->d2h).r + jc_mul(j->f3, o->d3s).r;
    jc fp = jc_add(jc_add(j->f1, jc_mul(j->f2, o->d)), jc_mul(j->f3, o->d2h));
    double dd = 2.0 * re / sqrt(fp.r * fp.r + fp.i * fp.i);
    double n2 = j->nu - log2(re);
    if (!(re > 0.0 && isfinite(n2) && isfinite(dd) && dd > 0.0)) return 0;
    *nu = n2;
    *de = (float)dd;
    *x = fp.r;
    *y = fp.i;
    return 1;
}

/* samples.rs:53-58 Samples::unit */
static void samples_unit(uint16_t a, float *c, float *s) {
    float t = (float)a * (6.2831855f / 65536.0f);
    *s = sinf(t);
    *c = cosf(t);
}

/* jetaa.rs:123-245 rebuild */
void fdo_jetaa_rebuild(uint8_t *cls, double *nu, float *de, uint16_t *normal, uint64_t w,
                       uint64_t h, uint8_t *todo) {
    const uint64_t nx = 2 * w, n = w * h;
#define JIDX(px, py, k) ((2 * (py) + (k) / 2) * nx + 2 * (px) + (k) % 2)
#define JREAL(q) JIDX((q) % w, (q) / w, 0)
    int *gok = (int *)calloc(n ? n : 1, sizeof(int));
    double *gnu = (double *)calloc(n ? n : 1, sizeof(double));
    jc *gc = (jc *)calloc(n ? n : 1, sizeof(jc));
    jjet *jets = (jjet *)calloc(n ? n : 1, sizeof(jjet));
    for (uint64_t q = 0; q < n; q++) {
        uint64_t i = JREAL(q);
        if ((cls[i] & 3) != FDO_ESCAPED) continue;
        float x, y;
        samples_unit(normal[i], &x, &y);
        gok[q] = 1;
        gnu[q] = nu[i];
        jc v = {(double)x, (double)y};
        gc[q] = jc_scale(v, 2.0 / (double)de[i]);
    }
    for (uint64_t q = 0; q < n; q++) {
        uint64_t px = q % w, py = q / w;
        if (px == 0 || py == 0 || px + 1 >= w || py + 1 >= h || !gok[q]) continue;
        uint64_t nbr[4] = {q + 1, q - 1, q - w, q + w};
        jc f[4];
        int all = 1;
        for (int k = 0; k < 4; k++) {
            if (!gok[nbr[k]]) { all = 0; break; }
            f[k] = jc_scale(gc[nbr[k]], exp2(gnu[q] - gnu[nbr[k]]));
        }
        if (!all) continue;
        jc a = jc_sub(f[0], f[1]), b = jc_sub(f[2], f[3]);
        jjet *j = &jets[q];
        j->nu = gnu[q];
        j->f1 = gc[q];
        jc t = {a.r + b.i, a.i - b.r};
        j->f2 = jc_scale(t, 0.25);
        j->f3 = jc_scale(jc_sub(jc_sub(jc_add(f[0], f[1]), f[2]), f[3]), 0.5);
        j->ok = isfinite(j->f2.r) && isfinite(j->f2.i) && isfinite(j->f3.r) && isfinite(j->f3.i);
    }
    joff own[3] = {joff_new(0.5, 0.0), joff_new(0.0, 0.5), joff_new(0.5, 0.5)};
    joff oth[3] = {joff_new(-0.5, 0.0), joff_new(0.0, -0.5), joff_new(-0.5, -0.5)};
    const double M = 3.0;
    /* pred: 0 compute, 1 own kind, 2 escaped (jetaa.rs:119-121) */
    uint8_t *what = (uint8_t *)calloc(3 * (n ? n : 1), 1);
    double *pnu = (double *)calloc(3 * (n ? n : 1), sizeof(double));
    float *pde = (float *)calloc(3 * (n ? n : 1), sizeof(float));
    uint16_t *pnm = (uint16_t *)calloc(3 * (n ? n : 1), sizeof(uint16_t));
    for (uint64_t q = 0; q < n; q++) {
        uint64_t px = q % w, py = q / w;
        if (px == 0 || py == 0 || px + 2 >= w || py + 2 >= h) continue;
        uint64_t r0 = JREAL(q);
        if (!gok[q]) {
            int k0 = cls[r0] & 3, same = 1;
            uint64_t rs[5] = {q + 1, q + w, q + w + 1, q - 1, q - w};
            for (int k = 0; k < 5; k++) same &= (cls[JREAL(rs[k])] & 3) == k0;
            if (same) what[3 * q] = what[3 * q + 1] = what[3 * q + 2] = 1;
            continue;
        }
        if (de[r0] < 0.3f || !jets[q].ok) continue;
        uint64_t nb[3] = {q + 1, q + w, q + w + 1};
        for (int k = 0; k < 3; k++) {
            double an, ax, ay, bn, bx, by;
            float ad, bd;
            if (!jjet_at(&jets[q], &own[k], &an, &ad, &ax, &ay)) continue;
            if (!jets[nb[k]].ok || !jjet_at(&jets[nb[k]], &oth[k], &bn, &bd, &bx, &by)) continue;
            if (fabs(an - bn) >= M * 1e-3 || fabs(log((double)(ad / bd))) >= M * 1e-2) continue;
            double na = sqrt(ax * ax + ay * ay), nb2 = sqrt(bx * bx + by * by);
            jc ua = {ax / na, ay / na}, ub = {bx / nb2, by / nb2};
            double cross = ua.r * ub.i - ua.i * ub.r, dot = ua.r * ub.r + ua.i * ub.i;
            if (dot > 0.0 && fabs(cross) < sin(M * 3e-3)) {
                what[3 * q + k] = 2;
                pnu[3 * q + k] = 0.5 * (an + bn);
                pde[3 * q + k] = 0.5f * (ad + bd);
                pnm[3 * q + k] = sample_angle(ua.r + ub.r, ua.i + ub.i);
            }
        }
    }
    memset(todo, 0, 4 * n);
    for (uint64_t q = 0; q < n; q++) {
        uint64_t px = q % w, py = q / w;
        uint8_t inside = cls[JREAL(q)];
        for (int k = 0; k < 3; k++) {
            uint64_t i = JIDX(px, py, k + 1);
            if (what[3 * q + k] == 0) {
                todo[i] = 1;
            } else if (what[3 * q + k] == 1) {
                cls[i] = (inside & 3) == 3 ? inside : (uint8_t)(inside & 3);
            } else {
                cls[i] = FDO_ESCAPED;
                nu[i] = pnu[3 * q + k];
                de[i] = pde[3 * q + k];
                normal[i] = pnm[3 * q + k];
            }
        }
    }
#undef JIDX
#undef JREAL
    free(gok);
    free(gnu);
    free(gc);
    free(jets);
    free(what);
    free(pnu);
    free(pde);
    free(pnm);
}

/* jetaa.rs:248-261 merge */
void fdo_jetaa_merge(uint8_t *cls, double *nu, float *de, uint16_t *normal, const uint8_t *fcls,
                     const double *fnu, const float *fde, const uint16_t *fnormal,
                     const uint8_t *mask, uint64_t n) {
    for (uint64_t i = 0; i < n; i++) {
        if (!mask[i]) continue;
        cls[i] = fcls[i];
        nu[i] = fnu[i];
        de[i] = fde[i];
        normal[i] = fnormal[i];
    }
}

/* ---------------------------------------------------------------- fd-samples */

typedef struct {
    uint8_t *buf;
    uint64_t cap, n;
} wbuf;

static void wb_bytes(wbuf *w, const void *p, uint64_t n) {
    if (w->n + n <= w->cap) memcpy(w->buf + w->n, p, n);
    w->n += n;
}
static void wb_pad8(wbuf *w) {
    static const uint8_t z[8] = {0};
    uint64_t a = (w->n + 7) & ~7ull;
    wb_bytes(w, z, a - w->n);
}
static void wb_str(wbuf *w, const char *s) {
    uint16_t n = (uint16_t)strlen(s);
    wb_bytes(w, &n, 2);
    wb_bytes(w, s, n);
}

/* write.rs:8-57 write with header.rs:311-330 Header::encode (little-endian host) */
uint64_t fdo_write_fds(uint8_t *buf, uint64_t cap, const char *re, const char *im,
                       const char *width, double rotation, uint32_t nx, uint32_t ny, uint32_t ss,
                       uint64_t max_iter, double escape_radius, uint32_t columns,
                       const char *kernel_id, const fdo_columns *c) {
    wbuf w = {buf, cap, 0};
    uint16_t major = 1, minor = 1;
    uint32_t zero = 0, cols = columns | FDO_COL_CLASS;
    wb_bytes(&w, "FDSAMPLE", 8);
    wb_bytes(&w, &major, 2);
    wb_bytes(&w, &minor, 2);
    wb_bytes(&w, &cols, 4);
    wb_bytes(&w, &nx, 4);
    wb_bytes(&w, &ny, 4);
    wb_bytes(&w, &ss, 4);
    wb_bytes(&w, &zero, 4);
    wb_bytes(&w, &max_iter, 8);
    wb_bytes(&w, &escape_radius, 8);
    wb_bytes(&w, &rotation, 8);
    wb_str(&w, re);
    wb_str(&w, im);
    wb_str(&w, width);
    wb_str(&w, kernel_id);
    wb_pad8(&w);
    uint64_t n = (uint64_t)nx * ny;
    wb_bytes(&w, c->class_, n);
    wb_pad8(&w);
    if (cols & FDO_COL_NU) {
        wb_bytes(&w, c->nu, 8 * n);
        wb_pad8(&w);
    }
    if (cols & FDO_COL_DE) {
        wb_bytes(&w, c->de, 4 * n);
        wb_pad8(&w);
    }
    if (cols & FDO_COL_NORMAL) {
        wb_bytes(&w, c->normal, 2 * n);
        wb_pad8(&w);
    }
    return w.n;
}

/* Bla::find over a freshly built table, for the reference's find tests (bla.rs:586-642).
 * out_level[k] = level index i (block length 2 << i) or -1. Returns levels or -1. */
int fdo_bla_find_many(const double *re, const double *im, uint64_t len, double eps, double dc_max,
                      int64_t dc_exp, const uint64_t *ms, uint64_t nm, double d2, uint64_t budget,
                      int32_t *out_level) {
    bla_t t;
    if (bla_build(&t, re, im, len, eps, dc_max, dc_exp) != 0) return -1;
    for (uint64_t k = 0; k < nm; k++) out_level[k] = bla_find(&t, ms[k], d2, budget);
    int n = t.nlevels;
    bla_free(&t);
    return n;
}

/* Fixed::parse (decimal.rs:63-96): limbs little-endian into out; returns the sign or -1. */
int fdo_fixed_parse(const char *s, uint64_t limbs, uint64_t *out) {
    int neg;
    if (fixed_parse(s, (size_t)limbs, &neg, out) != 0) return -1;
    return neg;
}
