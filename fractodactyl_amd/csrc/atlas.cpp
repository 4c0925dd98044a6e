// Canonical chunk formats for reference orbits and BLA tables (docs/spec/ATLAS.md:127-333,
// crates/fd-atlas/src/{chunk,slab,bla,manifest,sha256}.rs), so orbits and tables can be
// stored and reused across runs and processes, and read by the reference's own atlas.
//
// A chunk is a 32-byte header stating its numeric contract, then the payload zero-padded
// to 8 bytes; every field is little-endian, every f64 canonical (+0.0 for either zero,
// one NaN pattern), and a chunk is named by the SHA-256 of its bytes.  Readers re-encode
// and refuse non-canonical bytes.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "abi_guard.hpp"
#include "fdg.h"

namespace {

// ---------------------------------------------------------------- SHA-256 (FIPS 180-4)

struct Sha256 {
    uint32_t h[8] = {0x6a09e667, 0xbb67ae85, 0x3c6ef372, 0xa54ff53a,
                     0x510e527f, 0x9b05688c, 0x1f83d9ab, 0x5be0cd19};
    uint8_t buf[64];
    size_t fill = 0;
    uint64_t total = 0;

    static uint32_t rotr(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }

    void block(const uint8_t* p) {
        static const uint32_t k[64] = {
            0x428a2f98, 0x71374491, 0xb5c0fbcf, 0xe9b5dba5, 0x3956c25b, 0x59f111f1, 0x923f82a4,
            0xab1c5ed5, 0xd807aa98, 0x12835b01, 0x243185be, 0x550c7dc3, 0x72be5d74, 0x80deb1fe,
            0x9bdc06a7, 0xc19bf174, 0xe49b69c1, 0xefbe4786, 0x0fc19dc6, 0x240ca1cc, 0x2de92c6f,
            0x4a7484aa, 0x5cb0a9dc, 0x76f988da, 0x983e5152, 0xa831c66d, 0xb00327c8, 0xbf597fc7,
            0xc6e00bf3, 0xd5a79147, 0x06ca6351, 0x14292967, 0x27b70a85, 0x2e1b2138, 0x4d2c6dfc,
            0x53380d13, 0x650a7354, 0x766a0abb, 0x81c2c92e, 0x92722c85, 0xa2bfe8a1, 0xa81a664b,
            0xc24b8b70, 0xc76c51a3, 0xd192e819, 0xd6990624, 0xf40e3585, 0x106aa070, 0x19a4c116,
            0x1e376c08, 0x2748774c, 0x34b0bcb5, 0x391c0cb3, 0x4ed8aa4a, 0x5b9cca4f, 0x682e6ff3,
            0x748f82ee, 0x78a5636f, 0x84c87814, 0x8cc70208, 0x90befffa, 0xa4506ceb, 0xbef9a3f7,
            0xc67178f2};
        uint32_t w[64];
        for (int i = 0; i < 16; i++)
            w[i] = (uint32_t)p[4 * i] << 24 | (uint32_t)p[4 * i + 1] << 16 |
                   (uint32_t)p[4 * i + 2] << 8 | p[4 * i + 3];
        for (int i = 16; i < 64; i++) {
            uint32_t s0 = rotr(w[i - 15], 7) ^ rotr(w[i - 15], 18) ^ (w[i - 15] >> 3);
            uint32_t s1 = rotr(w[i - 2], 17) ^ rotr(w[i - 2], 19) ^ (w[i - 2] >> 10);
            w[i] = w[i - 16] + s0 + w[i - 7] + s1;
        }
        uint32_t a = h[0], b = h[1], c = h[2], d = h[3], e = h[4], f = h[5], g = h[6], hh = h[7];
        for (int i = 0; i < 64; i++) {
            uint32_t t1 = hh + (rotr(e, 6) ^ rotr(e, 11) ^ rotr(e, 25)) + ((e & f) ^ (~e & g)) +
                          k[i] + w[i];
            uint32_t t2 = (rotr(a, 2) ^ rotr(a, 13) ^ rotr(a, 22)) + ((a & b) ^ (a & c) ^ (b & c));
            hh = g;
            g = f;
            f = e;
            e = d + t1;
            d = c;
            c = b;
            b = a;
            a = t1 + t2;
        }
        h[0] += a;
        h[1] += b;
        h[2] += c;
        h[3] += d;
        h[4] += e;
        h[5] += f;
        h[6] += g;
        h[7] += hh;
    }
    void update(const uint8_t* p, size_t n) {
        total += n;
        while (n > 0) {
            if (fill == 0 && n >= 64) {
                block(p);
                p += 64;
                n -= 64;
                continue;
            }
            size_t take = std::min(n, 64 - fill);
            std::memcpy(buf + fill, p, take);
            fill += take;
            p += take;
            n -= take;
            if (fill == 64) {
                block(buf);
                fill = 0;
            }
        }
    }
    void finish(uint8_t out[32]) {
        uint64_t bits = total * 8;
        uint8_t pad = 0x80;
        update(&pad, 1);
        uint8_t zero = 0;
        while (fill != 56) update(&zero, 1);
        uint8_t len[8];
        for (int i = 0; i < 8; i++) len[i] = (uint8_t)(bits >> (56 - 8 * i));
        update(len, 8);
        for (int i = 0; i < 8; i++)
            for (int j = 0; j < 4; j++) out[4 * i + j] = (uint8_t)(h[i] >> (24 - 8 * j));
    }
};

// ---------------------------------------------------------------- chunks (chunk.rs)

constexpr uint8_t kMagic[8] = {'F', 'D', 'C', 'H', 'U', 'N', 'K', 0};
constexpr size_t kHeader = 32;
enum : uint16_t { KIND_ORBIT_SLAB = 1, KIND_BLA = 2, KIND_ORBIT_MANIFEST = 9 };
constexpr uint16_t kFormulaMandelbrot = 1;
constexpr uint8_t kRoundingNearest = 1;

struct Contract {
    uint16_t kind, encoding, formula;
    uint32_t precision_bits;
    uint8_t rounding;
    bool operator==(const Contract& o) const {
        return kind == o.kind && encoding == o.encoding && formula == o.formula &&
               precision_bits == o.precision_bits && rounding == o.rounding;
    }
};

uint64_t canonical_f64(double v) {
    if (v == 0.0) return 0;
    if (std::isnan(v)) return 0x7ff8000000000000ull;
    uint64_t b;
    std::memcpy(&b, &v, 8);
    return b;
}

struct Builder {
    Contract c;
    std::vector<uint8_t> payload;
    void raw(const void* p, size_t n) {
        const uint8_t* q = static_cast<const uint8_t*>(p);
        payload.insert(payload.end(), q, q + n);
    }
    void u32(uint32_t v) {
        for (int i = 0; i < 4; i++) payload.push_back((uint8_t)(v >> (8 * i)));
    }
    void u64(uint64_t v) {
        for (int i = 0; i < 8; i++) payload.push_back((uint8_t)(v >> (8 * i)));
    }
    void f64(double v) { u64(canonical_f64(v)); }
    void str(const char* s) {
        size_t n = std::strlen(s);
        u64(n);
        raw(s, n);
    }
    std::vector<uint8_t> finish() const {
        std::vector<uint8_t> b(kMagic, kMagic + 8);
        auto put16 = [&](uint16_t v) {
            b.push_back((uint8_t)v);
            b.push_back((uint8_t)(v >> 8));
        };
        put16(1);  // major
        put16(0);  // minor
        put16(c.kind);
        put16(c.encoding);
        put16(c.formula);
        b.push_back(c.rounding);
        b.push_back(0);
        for (int i = 0; i < 4; i++) b.push_back((uint8_t)(c.precision_bits >> (8 * i)));
        uint64_t len = payload.size();
        for (int i = 0; i < 8; i++) b.push_back((uint8_t)(len >> (8 * i)));
        b.insert(b.end(), payload.begin(), payload.end());
        b.resize((b.size() + 7) & ~(size_t)7, 0);
        return b;
    }
};

uint16_t rd16(const uint8_t* p) { return (uint16_t)(p[0] | p[1] << 8); }
uint32_t rd32(const uint8_t* p) {
    return (uint32_t)p[0] | (uint32_t)p[1] << 8 | (uint32_t)p[2] << 16 | (uint32_t)p[3] << 24;
}
uint64_t rd64(const uint8_t* p) {
    uint64_t v = 0;
    for (int i = 7; i >= 0; i--) v = v << 8 | p[i];
    return v;
}

// Chunk::from_bytes (chunk.rs:181-222): the contract and the payload span, or false.
bool parse(const uint8_t* b, uint64_t n, Contract* c, const uint8_t** payload, uint64_t* len) {
    if (n < kHeader || std::memcmp(b, kMagic, 8) != 0 || rd16(b + 8) != 1) return false;
    c->kind = rd16(b + 12);
    c->encoding = rd16(b + 14);
    c->formula = rd16(b + 16);
    c->rounding = b[18];
    c->precision_bits = rd32(b + 20);
    if (c->kind == 0 || c->rounding > 2 || b[19] != 0) return false;
    *len = rd64(b + 24);
    if (*len > n || ((*len + 7) & ~(uint64_t)7) != n - kHeader) return false;
    for (uint64_t i = kHeader + *len; i < n; i++)
        if (b[i]) return false;
    *payload = b + kHeader;
    return true;
}

int64_t emit(const std::vector<uint8_t>& bytes, uint8_t* buf, uint64_t cap) {
    if (buf && cap >= bytes.size()) std::memcpy(buf, bytes.data(), bytes.size());
    return (int64_t)bytes.size();
}

Contract slab_contract(uint32_t bits) {
    return {KIND_ORBIT_SLAB, 1, kFormulaMandelbrot, bits, kRoundingNearest};
}

// OrbitSlab::to_chunk (slab.rs:66-86)
bool slab_bytes(const double* re, const double* im, uint64_t start, uint32_t len, uint32_t size,
                uint32_t bits, std::vector<uint8_t>* out) {
    if (size == 0 || len < 1 || len > size || start % size != 0) return false;
    for (uint32_t i = 0; i < len; i++)
        if (!std::isfinite(re[i]) || !std::isfinite(im[i])) return false;
    Builder b{slab_contract(bits), {}};
    b.u32(size);
    b.u32(len);
    b.u64(start);
    for (uint32_t i = 0; i < len; i++) b.f64(re[i]);
    for (uint32_t i = 0; i < len; i++) b.f64(im[i]);
    *out = b.finish();
    return true;
}

Contract manifest_contract(uint32_t bits) {
    return {KIND_ORBIT_MANIFEST, 1, kFormulaMandelbrot, bits, kRoundingNearest};
}

const Contract kBlaContract = {KIND_BLA, 1, kFormulaMandelbrot, 53, kRoundingNearest};

// BlaTable::to_chunk (crates/fd-atlas/src/bla.rs:34-70); blocks level-major, 7 doubles per
// block (a.re a.im b.re b.im r alpha beta), level j (1-based) holding (points - 2) >> j.
bool bla_bytes(const uint8_t orbit[32], double eps, double dc_max, uint64_t points,
               uint32_t levels, const double* blocks, std::vector<uint8_t>* out) {
    const uint64_t steps = points >= 2 ? points - 2 : 0;
    if (levels >= 63 || !(eps > 0.0 && eps < 1.0 && dc_max > 0.0 && std::isfinite(dc_max)))
        return false;
    uint64_t off = 0;
    for (uint32_t j = 1; j <= levels; j++) {
        uint64_t cnt = steps >> j;
        bool used = false;
        for (uint64_t t = 0; t < cnt; t++) {
            const double* blk = blocks + 7 * (off + t);
            bool zero = true;
            for (int f = 0; f < 7; f++) {
                if (!std::isfinite(blk[f])) return false;
                zero = zero && blk[f] == 0.0;
            }
            if (!(blk[4] > 0.0 || zero)) return false;
            used = used || blk[4] > 0.0;
        }
        if (!used) return false;
        off += cnt;
    }
    Builder b{kBlaContract, {}};
    b.raw(orbit, 32);
    b.f64(eps);
    b.f64(dc_max);
    b.u64(points);
    b.u64(levels);
    off = 0;
    for (uint32_t j = 1; j <= levels; j++) {
        uint64_t cnt = steps >> j;
        for (int f = 0; f < 7; f++)
            for (uint64_t t = 0; t < cnt; t++) b.f64(blocks[7 * (off + t) + f]);
        off += cnt;
    }
    *out = b.finish();
    return true;
}

bool same_bytes(const std::vector<uint8_t>& a, const uint8_t* b, uint64_t n) {
    return a.size() == n && std::memcmp(a.data(), b, n) == 0;
}

}  // namespace

using fdg::guarded;

extern "C" {

int fdg_chunk_id(const uint8_t* bytes, uint64_t n, uint8_t id[32]) {
    return guarded<int>(nullptr, [&]() -> int {
        if ((!bytes && n) || !id) return FDG_EINVAL;
        Sha256 s;
        s.update(bytes, n);
        s.finish(id);
        return FDG_OK;
    });
}

int64_t fdg_chunk_orbit_slab(const double* re, const double* im, uint64_t start, uint32_t len,
                             uint32_t size, uint32_t bits, uint8_t* buf, uint64_t cap) {
    return guarded<int64_t>(nullptr, [&]() -> int64_t {
        if (!re || !im) return FDG_EINVAL;
        std::vector<uint8_t> b;
        if (!slab_bytes(re, im, start, len, size, bits, &b)) return FDG_EINVAL;
        return emit(b, buf, cap);
    });
}

int fdg_chunk_orbit_slab_read(const uint8_t* bytes, uint64_t n, uint64_t* start, uint32_t* len,
                              uint32_t* size, uint32_t* bits, double* re, double* im,
                              uint64_t cap) {
    return guarded<int>(nullptr, [&]() -> int {
        Contract c;
        const uint8_t* p;
        uint64_t plen;
        if (!bytes || !parse(bytes, n, &c, &p, &plen) || !(c == slab_contract(c.precision_bits)) ||
            plen < 16 || rd16(bytes + 10) != 0)
            return FDG_EINVAL;
        uint32_t sz = rd32(p), ln = rd32(p + 4);
        uint64_t st = rd64(p + 8);
        if (plen != 16 + 16 * (uint64_t)ln) return FDG_EINVAL;
        std::vector<double> r(ln), m(ln);
        for (uint32_t i = 0; i < ln; i++) {
            uint64_t a = rd64(p + 16 + 8 * (uint64_t)i), b = rd64(p + 16 + 8 * ((uint64_t)ln + i));
            std::memcpy(&r[i], &a, 8);
            std::memcpy(&m[i], &b, 8);
        }
        std::vector<uint8_t> again;
        if (!slab_bytes(r.data(), m.data(), st, ln, sz, c.precision_bits, &again) ||
            !same_bytes(again, bytes, n))
            return FDG_EINVAL;  // not canonical
        if (start) *start = st;
        if (len) *len = ln;
        if (size) *size = sz;
        if (bits) *bits = c.precision_bits;
        if (re && im && cap >= ln) {
            std::memcpy(re, r.data(), 8 * (size_t)ln);
            std::memcpy(im, m.data(), 8 * (size_t)ln);
        }
        return FDG_OK;
    });
}

int64_t fdg_chunk_orbit_manifest(const char* center_re, const char* center_im, uint32_t bits,
                                 const uint8_t* slab_ids, uint64_t n, uint8_t* buf,
                                 uint64_t cap) {
    return guarded<int64_t>(nullptr, [&]() -> int64_t {
        if (!center_re || !center_im || !slab_ids || n == 0) return FDG_EINVAL;
        Builder b{manifest_contract(bits), {}};
        b.str(center_re);
        b.str(center_im);
        b.u64(n);
        b.raw(slab_ids, 32 * n);
        return emit(b.finish(), buf, cap);
    });
}

int fdg_chunk_orbit_manifest_read(const uint8_t* bytes, uint64_t n, char* center_re,
                                  char* center_im, uint64_t str_cap, uint32_t* bits,
                                  uint8_t* slab_ids, uint64_t ids_cap, uint64_t* count) {
    return guarded<int>(nullptr, [&]() -> int {
        Contract c;
        const uint8_t* p;
        uint64_t plen;
        if (!bytes || !parse(bytes, n, &c, &p, &plen) ||
            !(c == manifest_contract(c.precision_bits)) || rd16(bytes + 10) != 0)
            return FDG_EINVAL;
        uint64_t at = 0;
        std::string s[2];
        for (auto& x : s) {
            if (at + 8 > plen) return FDG_EINVAL;
            uint64_t k = rd64(p + at);
            at += 8;
            if (k > plen - at) return FDG_EINVAL;
            x.assign(reinterpret_cast<const char*>(p + at), k);
            at += k;
        }
        if (at + 8 > plen) return FDG_EINVAL;
        uint64_t ns = rd64(p + at);
        at += 8;
        if (ns == 0 || ns > (plen - at) / 32 || at + 32 * ns != plen) return FDG_EINVAL;
        if (center_re && center_im && str_cap > std::max(s[0].size(), s[1].size())) {
            std::memcpy(center_re, s[0].c_str(), s[0].size() + 1);
            std::memcpy(center_im, s[1].c_str(), s[1].size() + 1);
        }
        if (bits) *bits = c.precision_bits;
        if (count) *count = ns;
        if (slab_ids && ids_cap >= ns) std::memcpy(slab_ids, p + at, 32 * ns);
        return FDG_OK;
    });
}

int64_t fdg_chunk_bla(const uint8_t orbit_id[32], double eps, double dc_max, uint64_t points,
                      uint32_t levels, const double* blocks, uint8_t* buf, uint64_t cap) {
    return guarded<int64_t>(nullptr, [&]() -> int64_t {
        if (!orbit_id || (!blocks && levels)) return FDG_EINVAL;
        std::vector<uint8_t> b;
        if (!bla_bytes(orbit_id, eps, dc_max, points, levels, blocks, &b)) return FDG_EINVAL;
        return emit(b, buf, cap);
    });
}

int fdg_chunk_bla_read(const uint8_t* bytes, uint64_t n, uint8_t orbit_id[32], double* eps,
                       double* dc_max, uint64_t* points, uint32_t* levels, double* blocks,
                       uint64_t cap_blocks) {
    return guarded<int>(nullptr, [&]() -> int {
        Contract c;
        const uint8_t* p;
        uint64_t plen;
        if (!bytes || !parse(bytes, n, &c, &p, &plen) || !(c == kBlaContract) ||
            rd16(bytes + 10) != 0 || plen < 64)
            return FDG_EINVAL;
        uint64_t a = rd64(p + 32), d = rd64(p + 40), pts = rd64(p + 48), nl = rd64(p + 56);
        if (nl >= 63) return FDG_EINVAL;
        double e, dc;
        std::memcpy(&e, &a, 8);
        std::memcpy(&dc, &d, 8);
        const uint64_t steps = pts >= 2 ? pts - 2 : 0;
        uint64_t total = 0;
        // The block count comes from the chunk's own header: bound it by the payload before
        // any product with it (56 * total can wrap for a forged points field), as the
        // reference checks lengths before reading (crates/fd-atlas/src/bla.rs:100).
        const uint64_t fits = (plen - 64) / 56;
        for (uint64_t j = 1; j <= nl; j++) {
            total += steps >> j;
            if (total > fits) return FDG_EINVAL;
        }
        if (plen != 64 + 56 * total) return FDG_EINVAL;
        std::vector<double> blk(7 * total);
        uint64_t at = 64, off = 0;
        for (uint64_t j = 1; j <= nl; j++) {
            uint64_t cnt = steps >> j;
            for (int f = 0; f < 7; f++)
                for (uint64_t t = 0; t < cnt; t++, at += 8) {
                    uint64_t v = rd64(p + at);
                    std::memcpy(&blk[7 * (off + t) + f], &v, 8);
                }
            off += cnt;
        }
        std::vector<uint8_t> again;
        if (!bla_bytes(p, e, dc, pts, (uint32_t)nl, blk.data(), &again) || !same_bytes(again, bytes, n))
            return FDG_EINVAL;  // not canonical (or not a valid table)
        if (orbit_id) std::memcpy(orbit_id, p, 32);
        if (eps) *eps = e;
        if (dc_max) *dc_max = dc;
        if (points) *points = pts;
        if (levels) *levels = (uint32_t)nl;
        if (blocks && cap_blocks >= total) std::memcpy(blocks, blk.data(), 8 * blk.size());
        return FDG_OK;
    });
}

}  // extern "C"
