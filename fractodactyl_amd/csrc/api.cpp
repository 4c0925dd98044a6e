// C ABI (include/fdg.h): contexts, single frames (setup on the host, resident device
// buffers, kernel launches on one HIP stream, the column encoding), host-only helpers.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "internal.hpp"

namespace fdg {

thread_local std::string g_create_error;

int Buffers::outputs(fdg_ctx* ctx, size_t n) {
    FDG_TRY(ctx, o_kind.ensure(n), "hipMalloc");
    FDG_TRY(ctx, o_n.ensure(n * 8), "hipMalloc");
    FDG_TRY(ctx, o_zr.ensure(n * 8), "hipMalloc");
    FDG_TRY(ctx, o_zi.ensure(n * 8), "hipMalloc");
    FDG_TRY(ctx, o_dr.ensure(n * 8), "hipMalloc");
    FDG_TRY(ctx, o_di.ensure(n * 8), "hipMalloc");
    FDG_TRY(ctx, o_dexp.ensure(n * 8), "hipMalloc");
    FDG_TRY(ctx, o_ops.ensure(n * 4), "hipMalloc");
    FDG_TRY(ctx, counters.ensure(256), "hipMalloc");
    FDG_TRY(ctx, fdev.ensure(sizeof(FrameDev)), "hipMalloc");
    return FDG_OK;
}

int setup(fdg_ctx* ctx, const fdg_view* v, const fdg_params* p, Frame* fr) {
    if (!v || !p || !v->center_re || !v->center_im || !v->width)
        return ctx->fail(FDG_EINVAL, "null view or params");
    // grid.rs:779-799 setup
    if (p->ss == 0 || p->nx % p->ss != 0 || p->ny % p->ss != 0)
        return ctx->fail(FDG_EINVAL, "grid must be a whole number of pixels");
    std::string e;
    if (!Plane::make(v->center_re, v->center_im, v->width, v->rotation, p->nx, p->ny, &fr->pl, &e))
        return ctx->fail(FDG_EINVAL, e);
    Tier t = p->tier < 0 ? fr->pl.tier : (Tier)p->tier;
    if (p->tier > 2) return ctx->fail(FDG_EINVAL, "unknown tier");
    bool allows = t == TIER_F64 ? fr->pl.tier == TIER_F64
                                : (t == TIER_FIXED ? fr->pl.tier != TIER_SCALED : true);
    if (!allows) return ctx->fail(FDG_EINVAL, "tier is not valid at this depth");
    fr->tier = t;
    return FDG_OK;
}

int compute_orbit(fdg_ctx* ctx, const fdg_view* v, const fdg_params* p, const Frame& fr,
                  Orbit* o) {
    // grid.rs:802-815 compute
    if (fr.tier == TIER_F64) {
        *o = Orbit::f64(fr.pl.c_re, fr.pl.c_im, p->max_iter, p->escape_radius);
        return FDG_OK;
    }
    if (!Orbit::fixed(v->center_re, v->center_im, fr.pl.bits, p->max_iter, o))
        return ctx->fail(FDG_EINVAL, "bad centre for the fixed-point reference");
    return FDG_OK;
}

uint64_t next_pow2_host(uint64_t x) {
    uint64_t s = 1;
    while (s < x) s <<= 1;
    return s;
}

int upload_orbit(fdg_ctx* ctx, DevBuf* dst, const Orbit& o, double* seconds) {
    double t0 = now_s();
    // The kernels read up to two points past the orbit's end (their next-step prefetch);
    // the 64-byte pad after the last point is uploaded as zeros with it.
    std::vector<double> inter(2 * o.len() + 8, 0.0);
    for (uint64_t k = 0; k < o.len(); k++) {
        inter[2 * k] = o.re[k];
        inter[2 * k + 1] = o.im[k];
    }
    FDG_TRY(ctx, dst->ensure(inter.size() * 8), "hipMalloc orbit");
    FDG_TRY(ctx,
            hipMemcpyAsync(dst->p, inter.data(), inter.size() * 8, hipMemcpyHostToDevice,
                           ctx->stream),
            "orbit upload");
    FDG_TRY(ctx, hipStreamSynchronize(ctx->stream), "orbit upload");
    *seconds += now_s() - t0;
    return FDG_OK;
}

// Leading and trailing zero digits are dropped before comparing (orbit.rs:235-247).
bool same_decimal(const std::string& a, const std::string& b) {
    Decimal x, y;
    if (!Decimal::parse(a, &x) || !Decimal::parse(b, &y)) return false;
    auto canon = [](Decimal& d) {
        size_t lead = 0;
        while (lead < d.digits.size() && d.digits[lead] == 0) lead++;
        d.digits.erase(d.digits.begin(), d.digits.begin() + lead);
        while (!d.digits.empty() && d.digits.back() == 0) {
            d.digits.pop_back();
            d.exp10 += 1;
        }
        if (d.digits.empty()) d.exp10 = 0;
    };
    canon(x);
    canon(y);
    return x.neg == y.neg && x.digits == y.digits && x.exp10 == y.exp10;
}

int read_encode(fdg_ctx* ctx, const Buffers& b, size_t off, size_t n, const Plane& pl,
                const fdg_params* p, fdg_columns* out, fdg_outcomes* raw, double* transfer_s,
                double* encode_s, std::vector<uint64_t>* n_out, std::vector<uint8_t>* kind_out,
                std::vector<uint32_t>* ops_out) {
    double t1 = now_s();
    // Outcomes land in the context's pinned staging area (page-locked host memory: the
    // copies run at full PCIe rate and the encoder reads them in place), one
    // asynchronous copy per array on the context's stream and one synchronisation.
    bool want_ops = (raw && raw->ops) || ops_out;
    size_t kbytes = (n + 7) & ~(size_t)7;
    size_t need = kbytes + 6 * 8 * n + (want_ops ? 4 * n : 0);
    int rc = ctx->staging(need);
    if (rc) return rc;
    uint8_t* base = static_cast<uint8_t*>(ctx->pinned);
    uint8_t* kind = base;
    uint64_t* nn = reinterpret_cast<uint64_t*>(base + kbytes);
    double* zr = reinterpret_cast<double*>(nn + n);
    double* zi = zr + n;
    double* dr = zi + n;
    double* di = dr + n;
    int64_t* dexp = reinterpret_cast<int64_t*>(di + n);
    uint32_t* ops = want_ops ? reinterpret_cast<uint32_t*>(dexp + n) : nullptr;
    auto get = [&](void* dst, const DevBuf& src, size_t elem) {
        return hipMemcpyAsync(dst, static_cast<const uint8_t*>(src.p) + off * elem, n * elem,
                              hipMemcpyDeviceToHost, ctx->stream);
    };
    FDG_TRY(ctx, get(kind, b.o_kind, 1), "readback");
    FDG_TRY(ctx, get(nn, b.o_n, 8), "readback");
    FDG_TRY(ctx, get(zr, b.o_zr, 8), "readback");
    FDG_TRY(ctx, get(zi, b.o_zi, 8), "readback");
    FDG_TRY(ctx, get(dr, b.o_dr, 8), "readback");
    FDG_TRY(ctx, get(di, b.o_di, 8), "readback");
    FDG_TRY(ctx, get(dexp, b.o_dexp, 8), "readback");
    if (want_ops) FDG_TRY(ctx, get(ops, b.o_ops, 4), "readback");
    FDG_TRY(ctx, hipStreamSynchronize(ctx->stream), "readback");
    *transfer_s += now_s() - t1;
    if (n_out) n_out->assign(nn, nn + n);
    if (kind_out) kind_out->assign(kind, kind + n);
    if (ops_out) ops_out->assign(ops, ops + n);
    // The device holds the stop iterate of Unresolved samples too; Outcome::Unresolved
    // carries none (sample.rs:21-40).
    for (size_t i = 0; i < n; i++) {
        if (kind[i] == KIND_UNRESOLVED) nn[i] = 0;
        if (kind[i] == kKindUnset) {
            // outside a sample mask: left unset, as render_with_masked leaves it
            nn[i] = 0;
            zr[i] = zi[i] = dr[i] = di[i] = 0.0;
            dexp[i] = 0;
            if (ops) ops[i] = 0;
        }
    }
    double t2 = now_s();
    Encoder enc(pl, p->ss);
    double* nu = (p->columns & FDG_COL_NU) ? out->nu : nullptr;
    float* de = (p->columns & FDG_COL_DE) ? out->de : nullptr;
    uint16_t* nm = (p->columns & FDG_COL_NORMAL) ? out->normal : nullptr;
    unsigned nt = std::max(1u, std::min(32u, std::thread::hardware_concurrency()));
    run_parts(nt, [&](unsigned w) {
        size_t lo = n * w / nt, hi = n * (w + 1) / nt;
        for (size_t i = lo; i < hi; i++) {
            if (nu) nu[i] = 0.0;
            if (de) de[i] = 0.0f;
            if (nm) nm[i] = 0;
            if (kind[i] == kKindUnset) {
                out->class_[i] = 0;
                continue;
            }
            enc.put(i, kind[i], nn[i], zr[i], zi[i], dr[i], di[i], dexp[i], out->class_, nu,
                    de, nm);
        }
    });
    *encode_s = now_s() - t2;
    if (raw) {
        if (raw->kind) std::memcpy(raw->kind, kind, n);
        if (raw->n) std::memcpy(raw->n, nn, n * 8);
        if (raw->zr) std::memcpy(raw->zr, zr, n * 8);
        if (raw->zi) std::memcpy(raw->zi, zi, n * 8);
        if (raw->dr) std::memcpy(raw->dr, dr, n * 8);
        if (raw->di) std::memcpy(raw->di, di, n * 8);
        if (raw->dexp) std::memcpy(raw->dexp, dexp, n * 8);
        if (raw->ops) std::memcpy(raw->ops, ops, n * 4);
    }
    return FDG_OK;
}

std::vector<uint32_t> mask_to_list(const uint8_t* mask, size_t n) {
    // per-chunk counts, then each chunk writes its own slice
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<size_t> cnt(nt + 1, 0);
    run_parts(nt, [&](unsigned w) {
        size_t lo = n * w / nt, hi = n * (w + 1) / nt, c = 0;
        for (size_t i = lo; i < hi; i++) c += mask[i] != 0;
        cnt[w + 1] = c;
    });
    for (unsigned w = 0; w < nt; w++) cnt[w + 1] += cnt[w];
    std::vector<uint32_t> list(cnt[nt]);
    run_parts(nt, [&](unsigned w) {
        size_t lo = n * w / nt, hi = n * (w + 1) / nt, o = cnt[w];
        for (size_t i = lo; i < hi; i++)
            if (mask[i]) list[o++] = (uint32_t)i;
    });
    return list;
}

int read_encode_list(fdg_ctx* ctx, Buffers& b, const std::vector<uint32_t>& list, size_t n,
                     const Plane& pl, const fdg_params* p, fdg_columns* out, fdg_outcomes* raw,
                     bool keep_unset, double* transfer_s, double* encode_s) {
    double t1 = now_s();
    const size_t m = list.size();
    const bool want_ops = raw && raw->ops;
    size_t kbytes = (m + 7) & ~(size_t)7;
    size_t packed = kbytes + 6 * 8 * m + (want_ops ? 4 * m : 0);
    size_t need = std::max(packed, 4 * m + 8);
    int rc = ctx->staging(need);
    if (rc) return rc;
    uint8_t* base = static_cast<uint8_t*>(ctx->pinned);
    if (m > 0) {
        FDG_TRY(ctx, b.list.ensure(4 * m), "hipMalloc");
        FDG_TRY(ctx, b.packed.ensure(packed), "hipMalloc");
        // the list goes up through the pinned staging area, then the packed outcomes come
        // back through it (stream order keeps the two uses apart)
        std::memcpy(base, list.data(), 4 * m);
        FDG_TRY(ctx, hipMemcpyAsync(b.list.p, base, 4 * m, hipMemcpyHostToDevice, ctx->stream),
                "list upload");
        uint8_t* dp = static_cast<uint8_t*>(b.packed.p);
        GatherDev g{};
        g.list = b.list.as<uint32_t>();
        g.m = (uint32_t)m;
        g.kind = b.o_kind.as<uint8_t>();
        g.n = b.o_n.as<uint64_t>();
        g.zr = b.o_zr.as<double>();
        g.zi = b.o_zi.as<double>();
        g.dr = b.o_dr.as<double>();
        g.di = b.o_di.as<double>();
        g.dexp = b.o_dexp.as<int64_t>();
        g.ops = b.o_ops.as<uint32_t>();
        g.kind_c = dp;
        g.n_c = reinterpret_cast<uint64_t*>(dp + kbytes);
        g.zr_c = reinterpret_cast<double*>(g.n_c + m);
        g.zi_c = g.zr_c + m;
        g.dr_c = g.zi_c + m;
        g.di_c = g.dr_c + m;
        g.dexp_c = reinterpret_cast<int64_t*>(g.di_c + m);
        g.ops_c = want_ops ? reinterpret_cast<uint32_t*>(g.dexp_c + m) : nullptr;
        FDG_TRY(ctx, launch_gather_outcomes(g, ctx->cus, ctx->stream), "gather");
        FDG_TRY(ctx, hipMemcpyAsync(base, dp, packed, hipMemcpyDeviceToHost, ctx->stream),
                "readback");
    }
    FDG_TRY(ctx, hipStreamSynchronize(ctx->stream), "readback");
    *transfer_s += now_s() - t1;
    const uint8_t* kind = base;
    const uint64_t* nn = reinterpret_cast<const uint64_t*>(base + kbytes);
    const double* zr = reinterpret_cast<const double*>(nn + m);
    const double* zi = zr + m;
    const double* dr = zi + m;
    const double* di = dr + m;
    const int64_t* dexp = reinterpret_cast<const int64_t*>(di + m);
    const uint32_t* ops = want_ops ? reinterpret_cast<const uint32_t*>(dexp + m) : nullptr;
    double t2 = now_s();
    Encoder enc(pl, p->ss);
    double* nu = (p->columns & FDG_COL_NU) ? out->nu : nullptr;
    float* de = (p->columns & FDG_COL_DE) ? out->de : nullptr;
    uint16_t* nm = (p->columns & FDG_COL_NORMAL) ? out->normal : nullptr;
    unsigned nt = std::max(1u, std::min(32u, std::thread::hardware_concurrency()));
    if (!keep_unset) {
        // every sample outside the list reads as an unset one: class 0, zero columns
        run_parts(nt, [&](unsigned w) {
            size_t lo = n * w / nt, hi = n * (w + 1) / nt;
            std::memset(out->class_ + lo, 0, hi - lo);
            if (nu) std::memset(nu + lo, 0, (hi - lo) * 8);
            if (de) std::memset(de + lo, 0, (hi - lo) * 4);
            if (nm) std::memset(nm + lo, 0, (hi - lo) * 2);
            if (raw) {
                if (raw->kind) std::memset(raw->kind + lo, kKindUnset, hi - lo);
                for (void* a : {(void*)raw->n, (void*)raw->zr, (void*)raw->zi, (void*)raw->dr,
                                (void*)raw->di, (void*)raw->dexp})
                    if (a) std::memset(static_cast<uint8_t*>(a) + lo * 8, 0, (hi - lo) * 8);
                if (raw->ops) std::memset(raw->ops + lo, 0, (hi - lo) * 4);
            }
        });
    }
    run_parts(nt, [&](unsigned w) {
        size_t lo = m * w / nt, hi = m * (w + 1) / nt;
        for (size_t t = lo; t < hi; t++) {
            const size_t i = list[t];
            // Outcome::Unresolved carries no iterate (sample.rs:21-40)
            const uint64_t ni = kind[t] == KIND_UNRESOLVED ? 0 : nn[t];
            if (nu) nu[i] = 0.0;
            if (de) de[i] = 0.0f;
            if (nm) nm[i] = 0;
            enc.put(i, kind[t], ni, zr[t], zi[t], dr[t], di[t], dexp[t], out->class_, nu, de,
                    nm);
            if (raw) {
                if (raw->kind) raw->kind[i] = kind[t];
                if (raw->n) raw->n[i] = ni;
                if (raw->zr) raw->zr[i] = zr[t];
                if (raw->zi) raw->zi[i] = zi[t];
                if (raw->dr) raw->dr[i] = dr[t];
                if (raw->di) raw->di[i] = di[t];
                if (raw->dexp) raw->dexp[i] = dexp[t];
                if (raw->ops) raw->ops[i] = ops[t];
            }
        }
    });
    *encode_s = now_s() - t2;
    return FDG_OK;
}

}  // namespace fdg

using namespace fdg;

namespace {

// Launch the BLA table build over the uploaded orbit on the context's stream without
// waiting for it (*built: a table exists, i.e. the orbit has at least 4 points); the level
// count comes from the per-level flags once the build has run (finish_table on the host
// after a synchronisation, or pass_prologue on the device).  The flags are cleared first
// unless the last pass set-up left them clear.
int launch_table(fdg_ctx* ctx, Buffers* bf, uint64_t len, double eps, double dc_max,
                 int64_t dc_exp, bool* built) {
    Buffers& B = *bf;
    if (!(eps > 0.0 && eps <= kEpsLimit && dc_max > 0.0 && std::isfinite(dc_max)))
        return ctx->fail(FDG_EINVAL,
                         "BLA needs 0 < eps <= 2^-24 and a finite dc_max > 0");
    uint64_t steps = len >= 2 ? len - 2 : 0;
    uint64_t stride = next_pow2_host(std::max<uint64_t>(steps, 2));
    B.bla_stride = stride;
    B.bla_steps = steps;
    *built = false;
    if (steps < 2) return FDG_OK;
    FDG_TRY(ctx, B.bla_r.ensure(stride * 8), "hipMalloc");
    // (one entry past the table: the level search's untaken loads may land there)
    FDG_TRY(ctx, B.bla_r2.ensure((stride + 1) * 8), "hipMalloc");
    FDG_TRY(ctx, B.bla_ab.ensure(stride * 32), "hipMalloc");
    FDG_TRY(ctx, B.bla_alpha.ensure(stride * 8), "hipMalloc");
    FDG_TRY(ctx, B.bla_beta.ensure(stride * 8), "hipMalloc");
    void* flags_before = B.bla_valid.p;
    FDG_TRY(ctx, B.bla_valid.ensure(kMaxLevels * 4), "hipMalloc");
    if (B.bla_valid.p != flags_before) B.valid_clean = false;
    if (!B.valid_clean)
        FDG_TRY(ctx, hipMemsetAsync(B.bla_valid.p, 0, kMaxLevels * 4, ctx->stream), "memset");
    B.valid_clean = false;  // the build below sets flags
    FDG_TRY(ctx, hipEventRecord(ctx->ev[0], ctx->stream), "event");
    FDG_TRY(ctx,
            launch_bla_build(B.orbit.as<double>(), B.bla_r.as<double>(),
                             B.bla_r2.as<double>(), B.bla_ab.as<double>(),
                             B.bla_alpha.as<double>(), B.bla_beta.as<double>(),
                             B.bla_valid.as<uint32_t>(), steps, stride, eps, dc_max, dc_exp,
                             kMaxLevels, ctx->stream),
            "BLA build launch");
    FDG_TRY(ctx, hipEventRecord(ctx->ev[1], ctx->stream), "event");
    *built = true;
    return FDG_OK;
}

// After the stream has been synchronised: the table's level count from the per-level
// "any valid block" flags copied to the host, and the build's device time.
int finish_table(fdg_ctx* ctx, const Buffers& B, const uint32_t* valid, uint32_t* levels,
                 double* seconds) {
    float ms = 0;
    FDG_TRY(ctx, hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]), "event timing");
    *seconds = ms * 1e-3;
    // bla.rs:442-445: levels end before the first level with no valid block.
    uint32_t n = 0;
    while (n < kMaxLevels && (B.bla_steps >> (n + 1)) > 0 && valid[n]) n++;
    *levels = n;
    return FDG_OK;
}

// The table built and its level count read back (a synchronous build).
int build_table(fdg_ctx* ctx, Buffers* bf, uint64_t len, double eps, double dc_max,
                int64_t dc_exp, uint32_t* levels, double* seconds) {
    bool built = false;
    *levels = 0;
    int rc = launch_table(ctx, bf, len, eps, dc_max, dc_exp, &built);
    if (rc || !built) return rc;
    uint32_t valid[kMaxLevels];
    FDG_TRY(ctx,
            hipMemcpyAsync(valid, bf->bla_valid.p, sizeof(valid), hipMemcpyDeviceToHost,
                           ctx->stream),
            "readback");
    FDG_TRY(ctx, hipStreamSynchronize(ctx->stream), "BLA build");
    return finish_table(ctx, *bf, valid, levels, seconds);
}

void fill_series(const Series& sr, fdg_series* out) {
    out->n = sr.n;
    out->rho = sr.rho;
    out->points = sr.points;
    out->b_len = (uint32_t)(sr.b.size() / 2);
    if (!sr.b.empty()) std::memcpy(out->b, sr.b.data(), sr.b.size() * sizeof(double));
    out->state_bound = sr.state_bound;
    out->derivative_bound = sr.derivative_bound;
    out->shift_px = sr.shift_px;
    out->built_steps = sr.built_steps;
    out->coefficient_pairs = sr.coefficient_pairs;
    out->growth_steps = sr.growth_steps;
    out->stop = sr.stop;
}

// Predicted frame cost of a probe run, in units of one sample's step: the larger of the
// throughput term (the frame's operations — plain steps, applied blocks and the series
// evaluations at half a step per Horner term — over the steps the chip advances in
// parallel, `parallel`) and the latency term (the longest chain one sample ran, which a
// persistent launch cannot split).  Measured on this kernel: ~3e4 sample-steps advance
// in the time one lane takes one step (128 per CU), and frames with interior samples run
// for as long as their longest sample.
double run_cost(const fdg_stats& st, uint64_t started, uint32_t b_len, double scale,
                double parallel) {
    double ops = (double)st.plain_steps + (double)st.blocks +
                 0.5 * (double)b_len * (double)started;
    return std::max(ops * scale / parallel, (double)st.max_sample_ops);
}
}  // namespace

struct fdg_frame {
    fdg_ctx* ctx = nullptr;
    std::string re, im, width;
    fdg_view view;
    fdg_params params;
    Frame fr;
    int route = 0, form = 0;
    double eps = 0;
    uint32_t count = 0;
    uint64_t len = 0, nucleus = 0;
    uint32_t bits = 0;
    double dc = 0;
    int64_t dc_exp = 0;
    double reference_seconds = 0, transfer_seconds = 0;
    uint64_t reference_period = 0, reference_repeat_at = 0;  // the host orbit's exact repeat
    bool use_bla = false, ran = false;
    bool masked = false;  // a sample mask is set (fdg_frame_mask)
    DevBuf mask;          // 1 byte per sample on the device
    std::vector<double> orbit_re, orbit_im;  // host copy of the reference orbit
    std::vector<uint8_t> mask_host;          // host copy of the sample mask, when set
    bool has_series = false;  // a series start is set (fdg_frame_series), maybe with n = 0
    uint32_t series_n = 0, series_len = 0;
    double series_inv_rho = 0;
    DevBuf series_b;      // its coefficients, (re, im) pairs
    uint64_t run_max_iter = 0;  // the series probe's budget for a run (0: params.max_iter)
    fdg_launch launch{};        // launch tuning (0 fields: the defaults)
    fdg_stats last;
    Buffers buf;
};

namespace {

// The shape a tuned launch is kept under: form, tier, derivative, BLA and the
// power-of-two buckets of the sample count and max_iter.
std::string tune_key(const fdg_frame* f) {
    auto lg = [](uint64_t x) { return x ? 64 - __builtin_clzll(x) : 0; };
    bool deriv = (f->params.columns & (FDG_COL_DE | FDG_COL_NORMAL)) != 0;
    return std::to_string(f->form) + "/" + std::to_string(f->fr.tier) + "/" +
           std::to_string(deriv) + "/" + std::to_string(f->use_bla) + "/" +
           std::to_string(lg(f->count)) + "/" + std::to_string(lg(f->params.max_iter));
}

int check_launch(fdg_ctx* ctx, const fdg_launch& l) {
    if (l.groups_per_cu > 32) return ctx->fail(FDG_EINVAL, "groups_per_cu must be in [1, 32]");
    if (l.claim_div > 64) return ctx->fail(FDG_EINVAL, "claim_div must be in [1, 64]");
    // (the 32-bit work counter overshoots the sample count, below 2^31, by at most one
    // claim per wave after it is exhausted: 2^15 waves x 2^14 stays far from wrapping)
    if (l.claim_cap != 0 && (l.claim_cap % 64 != 0 || l.claim_cap > 16384))
        return ctx->fail(FDG_EINVAL, "claim_cap must be a multiple of 64 in [64, 16384]");
    if (l.refill > 64) return ctx->fail(FDG_EINVAL, "refill must be in [1, 64]");
    return FDG_OK;
}

}  // namespace

extern "C" {

int fdg_abi_version(void) { return FDG_ABI_VERSION; }

int fdg_create(fdg_ctx** out, int device) {
    return guarded<int>(nullptr, [&]() -> int {
        if (!out) return FDG_EINVAL;
        *out = nullptr;
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
            g_create_error = "no HIP device";
            return FDG_EDEVICE;
        }
        if (device < 0 || device >= n) {
            g_create_error = "device index out of range";
            return FDG_EINVAL;
        }
        fdg_ctx* c = new fdg_ctx();
        c->device = device;
        if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&c->stream) != hipSuccess) {
            g_create_error = "cannot open the HIP device";
            delete c;
            return FDG_EDEVICE;
        }
        for (auto& e : c->ev) {
            if (hipEventCreate(&e) != hipSuccess) {
                g_create_error = "cannot create HIP events";
                for (auto& x : c->ev)
                    if (x) (void)hipEventDestroy(x);
                (void)hipStreamDestroy(c->stream);
                delete c;
                return FDG_EDEVICE;
            }
        }
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->cus = prop.multiProcessorCount;
        noalloc_attach(device);  // (the device is current)
        if (hipHostMalloc((void**)&c->small, fdg_ctx::kSmallBytes, hipHostMallocDefault) !=
            hipSuccess) {
            g_create_error = "cannot allocate page-locked host memory";
            for (auto& x : c->ev) (void)hipEventDestroy(x);
            (void)hipStreamDestroy(c->stream);
            delete c;
            return FDG_ENOMEM;
        }
        *out = c;
        return FDG_OK;
    });
}

void fdg_destroy(fdg_ctx* c) {
    guarded_void([&] {
        if (!c) return;
        // Teardown has no caller to report to: each step runs whatever the previous returned.
        (void)hipSetDevice(c->device);
        (void)hipStreamSynchronize(c->stream);
        for (auto& e : c->ev)
            if (e) (void)hipEventDestroy(e);
        if (c->spare) {
            c->spare->release();
            delete c->spare;
        }
        if (c->pinned) (void)hipHostFree(c->pinned);
        if (c->small) (void)hipHostFree(c->small);
        (void)hipStreamDestroy(c->stream);
        delete c;
    });
}

const char* fdg_last_error(const fdg_ctx* c) {
    return c ? c->err.c_str() : g_create_error.c_str();
}

int fdg_enabled(const char* value, int* enabled, int* route) {
    return guarded<int>(nullptr, [&]() -> int {
        std::string v = value ? value : "off";
        if (v == "off") {
            if (enabled) *enabled = 0;
            return FDG_OK;
        }
        if (v == "on" || v == "bla") {
            if (enabled) *enabled = 1;
            if (route) *route = v == "bla" ? FDG_ROUTE_BLA : FDG_ROUTE_PLAIN;
            return FDG_OK;
        }
        return FDG_EINVAL;
    });
}

int fdg_reference(fdg_ctx* ctx_in, const fdg_view* v, const fdg_params* p, double* re,
                  double* im, uint64_t cap, uint64_t* len, uint32_t* bits) {
    return guarded<int>(ctx_in ? &ctx_in->err : nullptr, [&]() -> int {
        // Host-only: a NULL context is allowed (errors then go to fdg_last_error(NULL)).
        fdg_ctx local;
        fdg_ctx* ctx = ctx_in ? ctx_in : &local;
        struct Report {
            fdg_ctx* c;
            bool own;
            ~Report() {
                if (own && !c->err.empty()) g_create_error = c->err;
            }
        } report{ctx, ctx_in == nullptr};
        Frame fr;
        int rc = setup(ctx, v, p, &fr);
        if (rc) return rc;
        Orbit o;
        rc = compute_orbit(ctx, v, p, fr, &o);
        if (rc) return rc;
        if (len) *len = o.len();
        if (bits) *bits = o.bits;
        if (re && im && cap >= o.len()) {
            std::memcpy(re, o.re.data(), o.len() * 8);
            std::memcpy(im, o.im.data(), o.len() * 8);
        }
        return FDG_OK;
    });
}

int fdg_nucleus(fdg_ctx* ctx_in, const char* re, const char* im, uint64_t period,
                uint32_t bits, uint32_t max_iters, char* re_out, char* im_out, size_t cap,
                uint32_t* iterations, double* log2_step) {
    return guarded<int>(ctx_in ? &ctx_in->err : nullptr, [&]() -> int {
        // Host-only: a NULL context is allowed (errors then go to fdg_last_error(NULL)).
        fdg_ctx local;
        fdg_ctx* ctx = ctx_in ? ctx_in : &local;
        struct Report {
            fdg_ctx* c;
            bool own;
            ~Report() {
                if (own && !c->err.empty()) g_create_error = c->err;
            }
        } report{ctx, ctx_in == nullptr};
        if (!re || !im) return ctx->fail(FDG_EINVAL, "null start for the nucleus search");
        if (bits > (1u << 20)) return ctx->fail(FDG_EINVAL, "nucleus search precision too large");
        Nucleus n;
        std::string err;
        bool ok = nucleus_newton(re, im, period, bits, max_iters, &n, &err);
        if (iterations) *iterations = n.iterations;
        if (log2_step) *log2_step = n.log2_step;
        if (!ok) return ctx->fail(FDG_EINVAL, err);
        if (re_out && im_out) {
            if (cap < n.re.size() + 1 || cap < n.im.size() + 1)
                return ctx->fail(FDG_EINVAL, "nucleus output buffers too small");
            std::memcpy(re_out, n.re.c_str(), n.re.size() + 1);
            std::memcpy(im_out, n.im.c_str(), n.im.size() + 1);
        }
        return FDG_OK;
    });
}

int fdg_build_bla(fdg_ctx* ctx, const double* re, const double* im, uint64_t len, double eps,
                  double dc_max, int64_t dc_exp, double* blocks, uint64_t cap_blocks,
                  uint32_t* levels) {
    return guarded<int>(ctx ? &ctx->err : nullptr, [&]() -> int {
        if (!ctx || !re || !im) return FDG_EINVAL;
        FDG_TRY(ctx, hipSetDevice(ctx->device), "hipSetDevice");
        Orbit o;
        o.re.assign(re, re + len);
        o.im.assign(im, im + len);
        double t = 0;
        Buffers B;
        struct Free {
            Buffers* b;
            ~Free() { b->release(); }
        } guard{&B};
        int rc = upload_orbit(ctx, &B.orbit, o, &t);
        if (rc) return rc;
        uint32_t nl = 0;
        rc = build_table(ctx, &B, len, eps, dc_max, dc_exp, &nl, &t);
        if (rc) return rc;
        if (levels) *levels = nl;
        if (!blocks) return FDG_OK;
        uint64_t steps = B.bla_steps, total = 0;
        for (uint32_t i = 0; i < nl; i++) total += steps >> (i + 1);
        if (total > cap_blocks) return ctx->fail(FDG_EINVAL, "block buffer too small");
        uint64_t stride = B.bla_stride;
        std::vector<double> r(stride), ab(4 * stride), al(stride), be(stride);
        if (nl) {
            FDG_TRY(ctx, hipMemcpy(r.data(), B.bla_r.p, stride * 8, hipMemcpyDeviceToHost), "copy");
            FDG_TRY(ctx, hipMemcpy(ab.data(), B.bla_ab.p, stride * 32, hipMemcpyDeviceToHost), "copy");
            FDG_TRY(ctx, hipMemcpy(al.data(), B.bla_alpha.p, stride * 8, hipMemcpyDeviceToHost), "copy");
            FDG_TRY(ctx, hipMemcpy(be.data(), B.bla_beta.p, stride * 8, hipMemcpyDeviceToHost), "copy");
        }
        uint64_t w = 0;
        for (uint32_t i = 0; i < nl; i++) {
            uint64_t off = bla_level_offset(stride, i), c = steps >> (i + 1);
            for (uint64_t k = 0; k < c; k++, w++) {
                double* b = blocks + 7 * w;
                b[0] = ab[4 * (off + k)];
                b[1] = ab[4 * (off + k) + 1];
                b[2] = ab[4 * (off + k) + 2];
                b[3] = ab[4 * (off + k) + 3];
                b[4] = r[off + k];
                b[5] = al[off + k];
                b[6] = be[off + k];
            }
        }
        return FDG_OK;
    });
}

static int frame_open(fdg_ctx* ctx, const fdg_view* v, const fdg_params* p, int route, int form,
                      double bla_eps, const double* sre, const double* sim, uint64_t slen,
                      uint32_t sbits, fdg_frame** out);

int fdg_frame_open(fdg_ctx* ctx, const fdg_view* v, const fdg_params* p, int route, int form,
                   double bla_eps, fdg_frame** out) {
    return guarded<int>(ctx ? &ctx->err : nullptr, [&]() -> int {
        return frame_open(ctx, v, p, route, form, bla_eps, nullptr, nullptr, 0, 0, out);
    });
}

int fdg_frame_open_orbit(fdg_ctx* ctx, const fdg_view* v, const fdg_params* p, int route,
                         int form, double bla_eps, const double* re, const double* im,
                         uint64_t len, uint32_t bits, fdg_frame** out) {
    return guarded<int>(ctx ? &ctx->err : nullptr, [&]() -> int {
        if (!re || !im) return ctx ? ctx->fail(FDG_EINVAL, "no orbit") : FDG_EINVAL;
        return frame_open(ctx, v, p, route, form, bla_eps, re, im, len, bits, out);
    });
}

static int frame_open(fdg_ctx* ctx, const fdg_view* v, const fdg_params* p, int route, int form,
                      double bla_eps, const double* sre, const double* sim, uint64_t slen,
                      uint32_t sbits, fdg_frame** out) {
    if (!ctx || !out) return FDG_EINVAL;
    *out = nullptr;
    FDG_TRY(ctx, hipSetDevice(ctx->device), "hipSetDevice");
    Frame fr;
    int rc = setup(ctx, v, p, &fr);
    if (rc) return rc;
    if (p->columns & (FDG_COL_BOUND | FDG_COL_STRIPE))
        return ctx->fail(FDG_EDECLINED,
                         "bound/stripe columns are CPU-kernel columns (frame.rs:133-144)");
    if (form != FDG_FORM_F64 && form != FDG_FORM_FE32 && form != FDG_FORM_MIXED)
        return ctx->fail(FDG_EINVAL, "unknown delta form");
    if (route != FDG_ROUTE_PLAIN && route != FDG_ROUTE_BLA)
        return ctx->fail(FDG_EINVAL, "unknown route");
    uint64_t count64 = (uint64_t)p->nx * p->ny;
    // Sample indices and the work counter are 32-bit; the counter may run past the count
    // by one chunk per wave.
    if (count64 >= (1ull << 31)) return ctx->fail(FDG_EINVAL, "grid too large");
    // The kernels count iterates in 32 bits.
    if (p->max_iter >= (1ull << 27)) return ctx->fail(FDG_EINVAL, "max_iter must be < 2^27");
    if (route == FDG_ROUTE_BLA) {
        // grid.rs:378-413 bla_checks, and bla.rs:27-33 check
        if (std::isnan(p->escape_radius) || p->escape_radius < 4.0)
            return ctx->fail(FDG_EINVAL, "escape radius is below 4: BLA needs >= 4");
        if (!(bla_eps > 0.0 && bla_eps <= kEpsLimit))
            return ctx->fail(FDG_EINVAL, "BLA needs 0 < eps <= 2^-24");
    }
    fdg_frame* f = new fdg_frame();
    // Owned by `guard` until it is handed out: every early return (and an exception
    // unwinding to the ABI barrier) frees its device buffers and the frame.
    struct Guard {
        fdg_frame* f;
        ~Guard() {
            if (f) {
                f->buf.release();
                delete f;
            }
        }
    } guard{f};
    f->ctx = ctx;
    f->re = v->center_re;
    f->im = v->center_im;
    f->width = v->width;
    f->view = *v;
    f->view.center_re = f->re.c_str();
    f->view.center_im = f->im.c_str();
    f->view.width = f->width.c_str();
    f->params = *p;
    f->fr = fr;
    f->route = route;
    f->form = form;
    f->eps = bla_eps;
    f->count = (uint32_t)count64;
    double t0 = now_s();
    Orbit o;
    if (sre) {
        // grid.rs:458-476: a supplied orbit of the view's precision (exactly 53 bits on the
        // f64 tier, at least the view's on the deep tiers) starting at Z_0 = 0; the header
        // then records its precision
        uint32_t need = fr.tier == TIER_F64 ? 53u : (uint32_t)fr.pl.bits;
        if (sbits != need && (fr.tier == TIER_F64 || sbits < need)) {
            char m[160];
            std::snprintf(m, sizeof(m), "supplied reference has %u bits; this view needs %s%u",
                          sbits, fr.tier == TIER_F64 ? "" : "at least ", need);
            return ctx->fail(FDG_EINVAL, m);
        }
        if (slen < 2 || sre[0] != 0.0 || sim[0] != 0.0 || slen > Orbit::kMaxLen)
            return ctx->fail(FDG_EINVAL,
                             "supplied reference orbit must start at Z_0 = 0 and hold at least 2 "
                             "points");
        o.re.assign(sre, sre + slen);
        o.im.assign(sim, sim + slen);
        o.bits = sbits;
        f->fr.pl.bits = sbits;
    } else {
        rc = compute_orbit(ctx, v, p, fr, &o);
        if (rc) return rc;
    }
    f->reference_seconds = now_s() - t0;
    f->reference_period = o.cycle_period;
    f->reference_repeat_at = o.cycle_at;
    // The loop reads Z_1 from its first step, and the scaled tier's rebase of a delta
    // below 1e-270 continues from Z_1 to Z_2 (sample.rs:422-449): shorter orbits are
    // refused here, where the reference would index past its orbit and panic
    // (sample.rs:367).
    if (o.len() < (fr.tier == TIER_SCALED ? 3u : 2u)) {
        return ctx->fail(FDG_EINVAL,
                         fr.tier == TIER_SCALED
                             ? "reference orbit too short: the scaled tier needs 3 points"
                             : "reference orbit too short: needs 2 points");
    }
    f->len = o.len();
    f->bits = o.bits;
    f->nucleus = o.nucleus_period();
    if (ctx->spare) {
        f->buf.adopt(*ctx->spare);  // the last closed frame's allocations, grown as needed
        delete ctx->spare;
        ctx->spare = nullptr;
    }
    rc = upload_orbit(ctx, &f->buf.orbit, o, &f->transfer_seconds);
    if (!rc) rc = f->buf.outputs(ctx, f->count);
    if (!rc && form != FDG_FORM_F64) {
        // FP32 floatexp orbit, 16 bytes per point, with two zero points of pad past the end;
        // each point carries log2 of its error radius (Reference::error_radius over the
        // orbit as computed: f64 from the rounded centre, or fixed point at its bits), which
        // the FP32 kernel charges into its error bound
        std::vector<double> q(o.len());
        error_radius(o.re.data(), o.im.data(), o.len(),
                     fr.tier == TIER_F64 ? -1 : (int64_t)o.bits, fr.pl.c_re, fr.pl.c_im,
                     q.data());
        uint64_t bytes = 16 * (f->len + 2);
        rc = ctx->hip(f->buf.orbit32.ensure(bytes), "hipMalloc");
        if (!rc) rc = ctx->hip(f->buf.qrad.ensure(8 * o.len()), "hipMalloc");
        if (!rc) rc = ctx->hip(f->buf.cert.ensure(f->count), "hipMalloc");
        if (!rc)
            rc = ctx->hip(hipMemcpyAsync(f->buf.qrad.p, q.data(), 8 * o.len(),
                                         hipMemcpyHostToDevice, ctx->stream),
                          "radius upload");
        if (!rc) rc = ctx->hip(hipMemsetAsync(f->buf.orbit32.p, 0, bytes, ctx->stream), "memset");
        if (!rc)
            rc = ctx->hip(launch_orbit_to_f32(f->buf.orbit.as<double>(), f->buf.qrad.as<double>(),
                                              f->buf.orbit32.as<float>(), f->len, ctx->stream),
                          "orbit conversion");
        if (!rc) rc = ctx->hip(hipStreamSynchronize(ctx->stream), "orbit conversion");
    }
    f->orbit_re = std::move(o.re);
    f->orbit_im = std::move(o.im);
    if (rc) return rc;
    dc_bound(fr.pl, fr.tier, p->nx, p->ny, &f->dc, &f->dc_exp);
    auto tuned = ctx->tuned.find(tune_key(f));
    if (tuned != ctx->tuned.end()) f->launch = tuned->second;
    guard.f = nullptr;
    *out = f;
    return FDG_OK;
}

int fdg_frame_run(fdg_frame* f, fdg_stats* st) {
    return guarded<int>(f && f->ctx ? &f->ctx->err : nullptr, [&]() -> int {
        if (!f) return FDG_EINVAL;
        fdg_ctx* ctx = f->ctx;
        FDG_TRY(ctx, hipSetDevice(ctx->device), "hipSetDevice");
        fdg_stats s;
        std::memset(&s, 0, sizeof(s));
        s.reference_seconds = f->reference_seconds;
        s.reference_period = f->reference_period;
        s.reference_repeat_at = f->reference_repeat_at;
        s.reference_len = f->len;
        s.reference_bits = f->bits;
        s.nucleus_period = f->nucleus;
        s.transfer_seconds = f->transfer_seconds;
        // The table is built on the stream and the kernel reads its level count from the
        // frame descriptor, which one small launch (pass_prologue) stores with the count taken
        // from the build's flags, zeroing the counters in the same launch: no host round trip
        // between the build and the perturbation kernel.  The host learns the count with the
        // statistics at the end.
        bool built = false;
        if (f->route == FDG_ROUTE_BLA) {
            int rc = launch_table(ctx, &f->buf, f->len, f->eps, f->dc, f->dc_exp, &built);
            if (rc) return rc;
        }
        Buffers& b = f->buf;
        FrameDev d;
        std::memset(&d, 0, sizeof(d));
        const Plane& pl = f->fr.pl;
        const fdg_params* p = &f->params;
        d.orbit = b.orbit.as<double>();
        d.len = f->len;
        d.nucleus = f->nucleus;
        d.bla.r2 = b.bla_r2.as<double>();
        d.bla.ab = b.bla_ab.as<double>();
        d.bla.steps = b.bla_steps;
        d.bla.stride = b.bla_stride;
        d.bla.nlevels = 0;  // written on the device from the build's flags (built tables)
        d.c_re = pl.c_re;
        d.c_im = pl.c_im;
        d.h = pl.h();
        d.h_m = pl.h_m;
        d.h_e = pl.h_e;
        d.cos_r = pl.cos_r;
        d.sin_r = pl.sin_r;
        d.half_x = pl.half_x;
        d.half_y = pl.half_y;
        d.nx = p->nx;
        d.ny = p->ny;
        d.count = f->count;
        d.tier = f->fr.tier;
        d.max_iter = f->run_max_iter ? f->run_max_iter : p->max_iter;
        d.r2 = p->escape_radius * p->escape_radius;
        d.mask = f->masked ? f->mask.as<uint8_t>() : nullptr;
        if (f->masked)
            // samples outside the mask keep this "not computed" kind through the pass, which
            // the mixed form's escaped list and the encoder both skip
            FDG_TRY(ctx, hipMemsetAsync(b.o_kind.p, kKindUnset, f->count, ctx->stream), "fill");
        d.kind = b.o_kind.as<uint8_t>();
        d.n = b.o_n.as<uint64_t>();
        d.zr = b.o_zr.as<double>();
        d.zi = b.o_zi.as<double>();
        d.dr = b.o_dr.as<double>();
        d.di = b.o_di.as<double>();
        d.dexp = b.o_dexp.as<int64_t>();
        d.ops = b.o_ops.as<uint32_t>();
        d.next = b.counters.as<uint32_t>();
        d.stats = reinterpret_cast<unsigned long long*>(b.counters.as<uint8_t>() + 64);
        bool series = f->has_series && f->series_n > 0;
        if (series) {
            d.series_b = f->series_b.as<double>();
            d.series_n = f->series_n;
            d.series_len = f->series_len;
            d.series_inv_rho = f->series_inv_rho;
        }
        bool deriv = (p->columns & (FDG_COL_DE | FDG_COL_NORMAL)) != 0;
        const uint32_t per_cu = f->launch.groups_per_cu ? f->launch.groups_per_cu : 8u;
        int grid = (int)std::min<uint64_t>((f->count + 255) / 256, (uint64_t)ctx->cus * per_cu);
        if (grid < 1) grid = 1;
        d.claim_div = f->launch.claim_div;
        d.claim_cap = f->launch.claim_cap;
        d.refill = f->launch.refill;
        static_assert(sizeof(FrameDev) <= 4096, "kernel descriptor slot");
        // counters: words 0-7 the work index, 8-23 the statistics (d.stats), word 24 (bytes
        // 192-195) the table's level count for the host
        uint32_t* levels_dev = reinterpret_cast<uint32_t*>(b.counters.as<uint8_t>() + 192);
        bool mixed = f->form == FDG_FORM_MIXED;
        bool fe32 = f->form == FDG_FORM_FE32 || mixed;
        if (fe32) {
            // FP32 table coefficients (A, B), same padded layout
            if (built) {
                FDG_TRY(ctx, b.ab32.ensure(b.bla_stride * 16), "hipMalloc");
                FDG_TRY(ctx, b.abexp32.ensure(b.bla_stride * 8), "hipMalloc");
                FDG_TRY(ctx,
                        launch_table_to_f32(b.bla_ab.as<double>(), b.ab32.as<float>(),
                                            b.abexp32.as<int32_t>(), b.bla_stride, ctx->stream),
                        "table conversion");
            }
            d.orbit32 = b.orbit32.as<float>();
            d.ab32 = b.ab32.as<float>();
            d.abexp32 = b.abexp32.as<int32_t>();
            d.bla_alpha = b.bla_alpha.as<double>();
            d.bla_beta = b.bla_beta.as<double>();
            d.cert = mixed ? b.cert.as<uint8_t>() : nullptr;
            d.defer_escaped = mixed ? 1u : 0u;
        }
        // The kernel writes every output field of every sample; only the counters start at 0.
        FDG_TRY(ctx,
                launch_pass_prologue(d, b.fdev.as<FrameDev>(),
                                     built ? b.bla_valid.as<uint32_t>() : nullptr, nullptr,
                                     b.bla_steps, kMaxLevels, b.counters.as<unsigned long long>(),
                                     24, levels_dev, true, ctx->stream),
                "pass set-up");
        if (built) b.valid_clean = true;
        FDG_TRY(ctx, hipEventRecord(ctx->ev[2], ctx->stream), "event");
        if (fe32)
            FDG_TRY(ctx,
                    launch_perturb_fe32(b.fdev.as<FrameDev>(), deriv, built, grid,
                                        ctx->stream, mixed),
                    "perturbation launch");
        else if (series)
            FDG_TRY(ctx, launch_perturb_f64_series(b.fdev.as<FrameDev>(), deriv, built, grid,
                                                   ctx->stream),
                    "perturbation launch");
        else
            FDG_TRY(ctx,
                    launch_perturb_f64(b.fdev.as<FrameDev>(), f->fr.tier == TIER_SCALED, deriv,
                                       built, false, grid, ctx->stream),
                    "perturbation launch");
        unsigned long long cnt2[ST_COUNT];
        std::memset(cnt2, 0, sizeof(cnt2));
        if (mixed) {
            // Second pass: the samples the FP32 pass found escaped, recomputed in FP64 (the
            // outcomes of every other sample are the FP32 pass's).
            FDG_TRY(ctx, b.list.ensure((size_t)f->count * 4), "hipMalloc");
            FDG_TRY(ctx, b.counters2.ensure(256), "hipMalloc");
            FDG_TRY(ctx, b.fdev2.ensure(sizeof(FrameDev)), "hipMalloc");
            uint32_t* listed = b.counters2.as<uint32_t>() + 1;
            FrameDev d2 = d;
            d2.mask = nullptr;  // the list holds only computed samples; it is indexed by position
            d2.list = b.list.as<uint32_t>();
            d2.list_count = listed;
            d2.defer_escaped = 0;
            d2.next = b.counters2.as<uint32_t>();
            d2.stats = reinterpret_cast<unsigned long long*>(b.counters2.as<uint8_t>() + 64);
            // words 0 (work index) and 1 (list length) and the statistics start at 0; the
            // level count is the first pass's
            FDG_TRY(ctx,
                    launch_pass_prologue(d2, b.fdev2.as<FrameDev>(), nullptr, levels_dev,
                                         b.bla_steps, kMaxLevels,
                                         b.counters2.as<unsigned long long>(), 32, nullptr, false,
                                         ctx->stream),
                    "pass set-up");
            FDG_TRY(ctx,
                    launch_escaped_list(b.o_kind.as<uint8_t>(), b.cert.as<uint8_t>(),
                                        b.list.as<uint32_t>(), listed, f->count, ctx->stream),
                    "escaped list");
            FDG_TRY(ctx,
                    launch_perturb_f64(b.fdev2.as<FrameDev>(), f->fr.tier == TIER_SCALED, deriv,
                                       built, false, grid, ctx->stream, true),
                    "refinement launch");
            FDG_TRY(ctx,
                    hipMemcpyAsync(ctx->small + fdg_ctx::kSlotCounts2, d2.stats, sizeof(cnt2),
                                   hipMemcpyDeviceToHost, ctx->stream),
                    "readback");
        }
        FDG_TRY(ctx, hipEventRecord(ctx->ev[3], ctx->stream), "event");
        unsigned long long cnt[ST_COUNT];
        // the statistics (counter bytes 64-191) and the level count (bytes 192-195) in one copy
        constexpr size_t kReadback = sizeof(cnt) + 8;
        static_assert(64 + kReadback <= 256 && kReadback <= 1024, "counter slot");
        FDG_TRY(ctx,
                hipMemcpyAsync(ctx->small + fdg_ctx::kSlotCounts, d.stats, kReadback,
                               hipMemcpyDeviceToHost, ctx->stream),
                "readback");
        FDG_TRY(ctx, hipStreamSynchronize(ctx->stream), "perturbation kernel");
        std::memcpy(cnt, ctx->small + fdg_ctx::kSlotCounts, sizeof(cnt));
        if (mixed) std::memcpy(cnt2, ctx->small + fdg_ctx::kSlotCounts2, sizeof(cnt2));
        uint32_t levels = 0;
        if (built) {
            std::memcpy(&levels, ctx->small + fdg_ctx::kSlotCounts + sizeof(cnt), sizeof(levels));
            float ms = 0;
            FDG_TRY(ctx, hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]), "event timing");
            s.bla_seconds = ms * 1e-3;
        }
        s.bla_levels = levels;
        f->use_bla = levels > 0;
        if (built && levels == 0) {
            // A table without a valid block: the kernel ran its BLA form with no level to
            // search, which counts every sample as a fallback and the closed-form ones apart;
            // report what the plain kernel (fd_kernel::render_with) reports.
            cnt[ST_FALLBACK] = cnt2[ST_FALLBACK] = 0;
            cnt[ST_CLOSED] = cnt2[ST_CLOSED] = 0;
        }
        for (int q = 0; q < ST_COUNT; q++)
            cnt[q] = q == ST_MAXOPS ? std::max(cnt[q], cnt2[q]) : cnt[q] + cnt2[q];
        float ms = 0;
        FDG_TRY(ctx, hipEventElapsedTime(&ms, ctx->ev[2], ctx->ev[3]), "event timing");
        s.kernel_seconds = ms * 1e-3;
        s.iterations = cnt[ST_ITERATIONS];
        s.equivalent = cnt[ST_EQUIVALENT];
        s.blocks = cnt[ST_BLOCKS];
        s.skipped = cnt[ST_SKIPPED];
        s.fallback_samples = cnt[ST_FALLBACK];
        s.closed_form_samples = cnt[ST_CLOSED];
        s.plain_steps = cnt[ST_PLAIN];
        s.max_sample_ops = cnt[ST_MAXOPS];
        s.wave_iterations = cnt[ST_WAVE_ITERS];
        s.lane_iterations = cnt[ST_LANE_ITERS];
        s.uncertified_samples = cnt[ST_UNCERT];
        f->last = s;
        f->ran = true;
        if (st) *st = s;
        return FDG_OK;
    });
}

int fdg_frame_read(fdg_frame* f, fdg_columns* out, fdg_outcomes* raw, fdg_stats* st,
                   char* kernel_id, size_t kernel_cap) {
    return guarded<int>(f && f->ctx ? &f->ctx->err : nullptr, [&]() -> int {
        if (!f) return FDG_EINVAL;
        fdg_ctx* ctx = f->ctx;
        if (!f->ran) return ctx->fail(FDG_EINVAL, "frame has not been run");
        const fdg_params* p = &f->params;
        if (!out || !out->class_) return ctx->fail(FDG_EINVAL, "the class column is required");
        if ((p->columns & FDG_COL_NU && !out->nu) || (p->columns & FDG_COL_DE && !out->de) ||
            (p->columns & FDG_COL_NORMAL && !out->normal))
            return ctx->fail(FDG_EINVAL, "a requested column has no buffer");
        FDG_TRY(ctx, hipSetDevice(ctx->device), "hipSetDevice");
        fdg_stats s = f->last;
        // A masked pass reads back only the samples it computed, gathered on the device
        int rc = f->masked
                     ? read_encode_list(ctx, f->buf, mask_to_list(f->mask_host.data(), f->count),
                                        f->count, f->fr.pl, p, out, raw, false,
                                        &s.transfer_seconds, &s.encode_seconds)
                     : read_encode(ctx, f->buf, 0, f->count, f->fr.pl, p, out, raw,
                                   &s.transfer_seconds, &s.encode_seconds, nullptr, nullptr,
                                   nullptr);
        if (rc) return rc;
        if (st) *st = s;
        if (kernel_id && kernel_cap) {
            std::string k = f->fr.pl.kernel(f->fr.tier);
            if (f->use_bla) k += " bla/1";
            if (f->has_series) k += " series/0";  // series.rs:371-372
            k += f->form == FDG_FORM_FE32 ? " gpu-hip-fe32/1"
                 : f->form == FDG_FORM_MIXED ? " gpu-hip-mixed/1" : " gpu-hip-f64/1";
            std::snprintf(kernel_id, kernel_cap, "%s", k.c_str());
        }
        return FDG_OK;
    });
}

int fdg_frame_read_columns(fdg_frame* f, fdg_columns* out, fdg_stats* st, char* kernel_id,
                           size_t kernel_cap) {
    return guarded<int>(f && f->ctx ? &f->ctx->err : nullptr, [&]() -> int {
        if (!f) return FDG_EINVAL;
        fdg_ctx* ctx = f->ctx;
        if (!f->ran) return ctx->fail(FDG_EINVAL, "frame has not been run");
        const fdg_params* p = &f->params;
        if (!out || !out->class_) return ctx->fail(FDG_EINVAL, "the class column is required");
        if ((p->columns & FDG_COL_NU && !out->nu) || (p->columns & FDG_COL_DE && !out->de) ||
            (p->columns & FDG_COL_NORMAL && !out->normal))
            return ctx->fail(FDG_EINVAL, "a requested column has no buffer");
        FDG_TRY(ctx, hipSetDevice(ctx->device), "hipSetDevice");
        fdg_stats s = f->last;
        const size_t n = f->count;
        const bool want_nu = p->columns & FDG_COL_NU, want_de = p->columns & FDG_COL_DE,
                   want_nm = p->columns & FDG_COL_NORMAL;
        Buffers& b = f->buf;
        FDG_TRY(ctx, b.c_cls.ensure(n), "hipMalloc");
        if (want_nu) FDG_TRY(ctx, b.c_nu.ensure(8 * n), "hipMalloc");
        if (want_de) FDG_TRY(ctx, b.c_de.ensure(4 * n), "hipMalloc");
        if (want_nm) FDG_TRY(ctx, b.c_nm.ensure(2 * n), "hipMalloc");
        const Plane& pl = f->fr.pl;
        EncodeDev e{};
        e.kind = b.o_kind.as<uint8_t>();
        e.n = b.o_n.as<uint64_t>();
        e.zr = b.o_zr.as<double>();
        e.zi = b.o_zi.as<double>();
        e.dr = b.o_dr.as<double>();
        e.di = b.o_di.as<double>();
        e.dexp = b.o_dexp.as<int64_t>();
        e.cls = b.c_cls.as<uint8_t>();
        e.nu = want_nu ? b.c_nu.as<double>() : nullptr;
        e.de = want_de ? b.c_de.as<float>() : nullptr;
        e.normal = want_nm ? b.c_nm.as<uint16_t>() : nullptr;
        e.count = (uint32_t)n;
        e.px_m = pl.h_m * (double)p->ss;
        e.px_e = pl.h_e;
        e.cos_r = pl.cos_r;
        e.sin_r = pl.sin_r;
        double t0 = now_s();
        FDG_TRY(ctx, launch_encode_columns(e, ctx->cus, ctx->stream), "column encoding");
        // columns into the pinned staging area, 8-byte aligned each, then into the caller's
        const size_t a1 = (n + 7) & ~(size_t)7, a8 = 8 * n, a4 = (4 * n + 7) & ~(size_t)7;
        const size_t need = a1 + (want_nu ? a8 : 0) + (want_de ? a4 : 0) + (want_nm ? 2 * n : 0);
        int rc = ctx->staging(need);
        if (rc) return rc;
        uint8_t* base = static_cast<uint8_t*>(ctx->pinned);
        uint8_t* h_cls = base;
        double* h_nu = reinterpret_cast<double*>(base + a1);
        float* h_de = reinterpret_cast<float*>(base + a1 + (want_nu ? a8 : 0));
        uint16_t* h_nm = reinterpret_cast<uint16_t*>(base + a1 + (want_nu ? a8 : 0) +
                                                     (want_de ? a4 : 0));
        FDG_TRY(ctx, hipMemcpyAsync(h_cls, e.cls, n, hipMemcpyDeviceToHost, ctx->stream), "readback");
        if (want_nu)
            FDG_TRY(ctx, hipMemcpyAsync(h_nu, e.nu, 8 * n, hipMemcpyDeviceToHost, ctx->stream),
                    "readback");
        if (want_de)
            FDG_TRY(ctx, hipMemcpyAsync(h_de, e.de, 4 * n, hipMemcpyDeviceToHost, ctx->stream),
                    "readback");
        if (want_nm)
            FDG_TRY(ctx, hipMemcpyAsync(h_nm, e.normal, 2 * n, hipMemcpyDeviceToHost, ctx->stream),
                    "readback");
        FDG_TRY(ctx, hipStreamSynchronize(ctx->stream), "column encoding");
        double t1 = now_s();
        unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        if (n < (1u << 16)) nt = 1;
        auto part = [&](unsigned w) {
            size_t lo = n * w / nt, hi = n * (w + 1) / nt;
            std::memcpy(out->class_ + lo, h_cls + lo, hi - lo);
            if (want_nu) std::memcpy(out->nu + lo, h_nu + lo, 8 * (hi - lo));
            if (want_de) std::memcpy(out->de + lo, h_de + lo, 4 * (hi - lo));
            if (want_nm) std::memcpy(out->normal + lo, h_nm + lo, 2 * (hi - lo));
        };
        run_parts(nt, part);
        s.transfer_seconds += t1 - t0;  // device encoding + the 15-byte columns' readback
        s.encode_seconds = now_s() - t1;
        if (st) *st = s;
        if (kernel_id && kernel_cap) {
            std::string k = f->fr.pl.kernel(f->fr.tier);
            if (f->use_bla) k += " bla/1";
            if (f->has_series) k += " series/0";
            k += f->form == FDG_FORM_FE32 ? " gpu-hip-fe32/1"
                 : f->form == FDG_FORM_MIXED ? " gpu-hip-mixed/1" : " gpu-hip-f64/1";
            std::snprintf(kernel_id, kernel_cap, "%s", k.c_str());
        }
        return FDG_OK;
    });
}

int fdg_frame_mask(fdg_frame* f, const uint8_t* mask) {
    return guarded<int>(f && f->ctx ? &f->ctx->err : nullptr, [&]() -> int {
        if (!f) return FDG_EINVAL;
        fdg_ctx* ctx = f->ctx;
        FDG_TRY(ctx, hipSetDevice(ctx->device), "hipSetDevice");
        f->masked = mask != nullptr;
        if (!mask) {
            f->mask_host.clear();
            return FDG_OK;
        }
        f->mask_host.assign(mask, mask + f->count);
        FDG_TRY(ctx, f->mask.ensure(f->count), "hipMalloc");
        FDG_TRY(ctx, hipMemcpyAsync(f->mask.p, mask, f->count, hipMemcpyHostToDevice, ctx->stream),
                "mask upload");
        FDG_TRY(ctx, hipStreamSynchronize(ctx->stream), "mask upload");
        return FDG_OK;
    });
}

int fdg_frame_launch(fdg_frame* f, const fdg_launch* launch) {
    return guarded<int>(f && f->ctx ? &f->ctx->err : nullptr, [&]() -> int {
        if (!f) return FDG_EINVAL;
        fdg_launch l{};
        if (launch) l = *launch;
        int rc = check_launch(f->ctx, l);
        if (rc) return rc;
        f->launch = l;
        return FDG_OK;
    });
}

int fdg_frame_tune(fdg_frame* f, uint32_t reps, fdg_launch* best, double* best_ms,
                   double* default_ms) {
    return guarded<int>(f && f->ctx ? &f->ctx->err : nullptr, [&]() -> int {
        if (!f) return FDG_EINVAL;
        if (reps < 1 || reps > 16) return f->ctx->fail(FDG_EINVAL, "reps must be in [1, 16]");
        const fdg_launch before = f->launch;
        // minimum kernel time over the reps under launch l
        auto time_of = [&](const fdg_launch& l, double* ms) -> int {
            f->launch = l;
            double m = 1e300;
            for (uint32_t r = 0; r < reps; r++) {
                fdg_stats st;
                int rc = fdg_frame_run(f, &st);
                if (rc) return rc;
                m = std::min(m, st.kernel_seconds * 1e3);
            }
            *ms = m;
            return FDG_OK;
        };
        const fdg_launch dflt{};
        double dms = 0.0;
        int rc = time_of(dflt, &dms);
        fdg_launch win = dflt;
        double wms = dms;
        static const uint32_t per_cu[] = {4, 8, 16}, divs[] = {2, 4, 8},
                              caps[] = {256, 1024, 4096}, refills[] = {1, 16};
        // a candidate replaces the default only when it is 1% faster, so that timing noise
        // alone never moves the choice away from the default
        auto consider = [&](const fdg_launch& l) {
            double ms = 0.0;
            rc = time_of(l, &ms);
            if (!rc && ms < (win.groups_per_cu == 0 ? 0.99 * wms : wms)) {
                win = l;
                wms = ms;
            }
        };
        for (uint32_t pc : per_cu)
            for (uint32_t dv : divs)
                for (uint32_t cp : caps)
                    if (!rc) consider(fdg_launch{pc, dv, cp, 0});
        // then partial refills (lanes refilled as they free up) under the best claim policy
        const fdg_launch base = win;
        for (uint32_t rf : refills)
            if (!rc) consider(fdg_launch{base.groups_per_cu, base.claim_div, base.claim_cap, rf});
        if (rc) {
            f->launch = before;
            return rc;
        }
        f->launch = win;
        f->ctx->tuned[tune_key(f)] = win;
        if (best) *best = win;
        if (best_ms) *best_ms = wms;
        if (default_ms) *default_ms = dms;
        return FDG_OK;
    });
}

int fdg_render_masked(fdg_ctx* ctx, const fdg_view* v, const fdg_params* p, int route, int form,
                      double bla_eps, const uint8_t* mask, fdg_columns* out, fdg_outcomes* raw,
                      fdg_stats* st, char* kernel_id, size_t kernel_cap) {
    return guarded<int>(ctx ? &ctx->err : nullptr, [&]() -> int {
        if (!ctx) return FDG_EINVAL;
        if (!out || !out->class_) return ctx->fail(FDG_EINVAL, "the class column is required");
        if (!mask) return ctx->fail(FDG_EINVAL, "sample mask must cover the sample grid");
        fdg_frame* f = nullptr;
        int rc = fdg_frame_open(ctx, v, p, route, form, bla_eps, &f);
        if (rc) return rc;
        rc = fdg_frame_mask(f, mask);
        if (!rc) rc = fdg_frame_run(f, nullptr);
        if (!rc) rc = fdg_frame_read(f, out, raw, st, kernel_id, kernel_cap);
        fdg_frame_close(f);
        return rc;
    });
}

void fdg_frame_close(fdg_frame* f) {
    guarded_void([&] {
        if (!f) return;
        (void)hipSetDevice(f->ctx->device);  // teardown: nothing to report to
        (void)hipStreamSynchronize(f->ctx->stream);
        if (!f->ctx->spare) {
            // kept for the next frame on this context (freed by fdg_destroy)
            Buffers* keep = new (std::nothrow) Buffers();
            if (keep) {
                keep->adopt(f->buf);
                f->ctx->spare = keep;
            }
        }
        f->buf.release();
        f->mask.release();
        f->series_b.release();
        delete f;
    });
}

void fdg_series_config_default(fdg_series_config* cfg) {
    guarded_void([&] {
        if (!cfg) return;
        SeriesConfig c;
        cfg->start_order = c.start_order;
        cfg->max_order = c.max_order;
        cfg->grow_at = c.grow_at;
        cfg->shift_px = c.shift_px;
        cfg->derivative_rel = c.derivative_rel;
    });
}

int fdg_series_build(const double* re, const double* im, uint64_t len, double rho, double pixel,
                     uint64_t max_iter, const fdg_series_config* cfg, fdg_series* out) {
    return guarded<int>(nullptr, [&]() -> int {
        if (!re || !im || !out || len < 1) return FDG_EINVAL;
        SeriesConfig c;
        if (cfg) {
            c.start_order = cfg->start_order;
            c.max_order = cfg->max_order;
            c.grow_at = cfg->grow_at;
            c.shift_px = cfg->shift_px;
            c.derivative_rel = cfg->derivative_rel;
        }
        if (!out->b || out->b_cap < std::max<uint32_t>(c.max_order, 1)) return FDG_EINVAL;
        fill_series(Series::build(re, im, len, rho, pixel, max_iter, c), out);
        return FDG_OK;
    });
}

int fdg_bla_dc_max(const fdg_view* v, const fdg_params* p, double* dc_max) {
    return guarded<int>(nullptr, [&]() -> int {
        if (!v || !p || !dc_max) return FDG_EINVAL;
        if (p->ss == 0 || p->nx % p->ss || p->ny % p->ss) return FDG_EINVAL;
        Plane pl;
        std::string e;
        if (!Plane::make(v->center_re, v->center_im, v->width, v->rotation, p->nx, p->ny, &pl, &e))
            return FDG_EINVAL;
        Tier t = p->tier < 0 ? pl.tier : (Tier)p->tier;
        if (t == TIER_SCALED) return FDG_EINVAL;  // BLA tables need f64 deltas
        int64_t de;
        dc_bound(pl, t, p->nx, p->ny, dc_max, &de);
        return FDG_OK;
    });
}

int fdg_frame_series(fdg_frame* f, const fdg_series* sr) {
    return guarded<int>(f && f->ctx ? &f->ctx->err : nullptr, [&]() -> int {
        if (!f) return FDG_EINVAL;
        fdg_ctx* ctx = f->ctx;
        if (!sr) {
            f->has_series = false;
            return FDG_OK;
        }
        // series.rs:342-366 render_masked's refusals
        if (f->fr.tier == TIER_SCALED)
            return ctx->fail(FDG_EINVAL, "a series start needs f64 deltas and no Bound column");
        if (f->form != FDG_FORM_F64)
            return ctx->fail(FDG_EDECLINED, "a series start is served by the FP64 form");
        if (sr->points != f->len) return ctx->fail(FDG_EINVAL, "series built over another orbit");
        if (std::isnan(f->dc) || f->dc > sr->rho) {
            char m[160];
            std::snprintf(m, sizeof(m), "view reaches |dc| = %e, beyond the series radius %e",
                          f->dc, sr->rho);
            return ctx->fail(FDG_EINVAL, m);
        }
        if (sr->n > 0 && (sr->n + 1 >= f->len || sr->b_len == 0 || !sr->b))
            return ctx->fail(FDG_EINVAL, "series start outside its orbit");
        // (Series::build stops below max_iter; a later start would leave no iterate to run)
        if (sr->n >= f->params.max_iter)
            return ctx->fail(FDG_EINVAL, "series start at or beyond max_iter");
        FDG_TRY(ctx, hipSetDevice(ctx->device), "hipSetDevice");
        if (sr->n > 0) {
            size_t bytes = (size_t)sr->b_len * 2 * sizeof(double);
            FDG_TRY(ctx, f->series_b.ensure(bytes), "hipMalloc");
            FDG_TRY(ctx,
                    hipMemcpyAsync(f->series_b.p, sr->b, bytes, hipMemcpyHostToDevice, ctx->stream),
                    "series upload");
            FDG_TRY(ctx, hipStreamSynchronize(ctx->stream), "series upload");
        }
        f->has_series = true;
        f->series_n = (uint32_t)sr->n;
        f->series_len = sr->b_len;
        f->series_inv_rho = 1.0 / sr->rho;
        return FDG_OK;
    });
}


int fdg_frame_series_prepare(fdg_frame* f, const fdg_series_config* cfg, fdg_series* out,
                             int* admitted) {
    return guarded<int>(f && f->ctx ? &f->ctx->err : nullptr, [&]() -> int {
        if (!f || !out || !admitted) return FDG_EINVAL;
        fdg_ctx* ctx = f->ctx;
        *admitted = 0;
        SeriesConfig c;
        if (cfg) {
            c.start_order = cfg->start_order;
            c.max_order = cfg->max_order;
            c.grow_at = cfg->grow_at;
            c.shift_px = cfg->shift_px;
            c.derivative_rel = cfg->derivative_rel;
        }
        if (!out->b || out->b_cap < std::max<uint32_t>(c.max_order, 1))
            return ctx->fail(FDG_EINVAL, "series coefficient buffer too small");
        out->n = 0;
        out->b_len = 0;
        const fdg_params* p = &f->params;
        uint64_t active = f->count;
        if (f->masked) {
            active = 0;
            for (uint8_t m : f->mask_host) active += m != 0;
        }
        // series.rs:248-261: small frames, the scaled tier and short budgets decline before
        // building (the FP32 forms take no series start)
        int rc = fdg_frame_series(f, nullptr);
        if (rc) return rc;
        if (active < 16384 || f->fr.tier == TIER_SCALED || p->max_iter < 256 ||
            f->form != FDG_FORM_F64)
            return FDG_OK;
        // series.rs:275-284: the series of the probe's budget (max_iter capped at 8192)
        Series sr = Series::build(f->orbit_re.data(), f->orbit_im.data(), f->len, f->dc,
                                  f->fr.pl.h() * (double)p->ss,
                                  std::min<uint64_t>(p->max_iter, 8192), c);
        fill_series(sr, out);
        if (sr.n == 0) return FDG_OK;
        // The probe: every 61st active sample plus the 16 x 16 block at the centre (where a
        // zoom's minibrot and its slow samples sit; the reference's nine probe points include
        // the centre), rendered without and with the start on this frame's route, at the
        // reference probe's budget: max_iter capped at 8192, so a long sample is censored and
        // its series attempt retried, which makes the admission conservative (series.rs:
        // 266-274).  Costs are predicted from counted work, so the decision is the same on
        // every run (the reference times the nine samples on the CPU, series.rs:285-339).
        std::vector<uint8_t> saved = f->mask_host;
        bool was_masked = f->masked;
        std::vector<uint8_t> probe(f->count, 0);
        uint64_t started = 0, seen = 0;
        const uint32_t nx = p->nx, ny = p->ny;
        for (uint32_t i = 0; i < f->count; i++) {
            if (was_masked && !saved[i]) continue;
            uint32_t x = i % nx, y = i / nx;
            bool centre = x + 8 >= nx / 2 && x < nx / 2 + 8 && y + 8 >= ny / 2 && y < ny / 2 + 8;
            if (seen++ % 61 == 0 || centre) {
                probe[i] = 1;
                started++;
            }
        }
        f->run_max_iter = std::min<uint64_t>(p->max_iter, 8192);
        rc = fdg_frame_mask(f, probe.data());
        fdg_stats base, with;
        if (!rc) rc = fdg_frame_run(f, &base);
        if (!rc) rc = fdg_frame_series(f, out);
        if (!rc) rc = fdg_frame_run(f, &with);
        f->run_max_iter = 0;
        int rc2 = fdg_frame_mask(f, was_masked ? saved.data() : nullptr);
        if (rc) return rc;
        if (rc2) return rc2;
        f->ran = false;  // the probe's outcomes are not the frame's
        const double scale = (double)active / (double)std::max<uint64_t>(started, 1);
        const double parallel = 128.0 * ctx->cus;
        double cost0 = run_cost(base, 0, 0, scale, parallel);
        double cost1 = run_cost(with, started, out->b_len, scale, parallel);
        // series.rs:336-339: a predicted net win with a 20% margin
        if (cost1 < 0.8 * cost0) {
            *admitted = 1;
            return FDG_OK;
        }
        return fdg_frame_series(f, nullptr);
    });
}

int fdg_render_series(fdg_ctx* ctx, const fdg_view* v, const fdg_params* p, int route,
                      double bla_eps, const fdg_series* sr, const uint8_t* mask,
                      fdg_columns* out, fdg_outcomes* raw, fdg_stats* st, char* kernel_id,
                      size_t kernel_cap) {
    return guarded<int>(ctx ? &ctx->err : nullptr, [&]() -> int {
        if (!ctx) return FDG_EINVAL;
        if (!out || !out->class_) return ctx->fail(FDG_EINVAL, "the class column is required");
        if (!sr) return ctx->fail(FDG_EINVAL, "no series");
        fdg_frame* f = nullptr;
        int rc = fdg_frame_open(ctx, v, p, route, FDG_FORM_F64, bla_eps, &f);
        if (rc) return rc;
        rc = fdg_frame_series(f, sr);
        if (!rc && mask) rc = fdg_frame_mask(f, mask);
        if (!rc) rc = fdg_frame_run(f, nullptr);
        if (!rc) rc = fdg_frame_read(f, out, raw, st, kernel_id, kernel_cap);
        fdg_frame_close(f);
        return rc;
    });
}

int fdg_jetaa_first(uint32_t w, uint32_t h, uint8_t* mask) {
    return guarded<int>(nullptr, [&]() -> int {
        if (!mask) return FDG_EINVAL;
        jetaa_first(w, h, mask);
        return FDG_OK;
    });
}

int fdg_jetaa_rebuild(fdg_columns* c, uint32_t w, uint32_t h, uint8_t* todo) {
    return guarded<int>(nullptr, [&]() -> int {
        if (!c || !c->class_ || !c->nu || !c->de || !c->normal || !todo) return FDG_EINVAL;
        jetaa_rebuild(c->class_, c->nu, c->de, c->normal, w, h, todo);
        return FDG_OK;
    });
}

int fdg_jetaa_merge(fdg_columns* to, const fdg_columns* from, const uint8_t* mask, uint64_t n) {
    return guarded<int>(nullptr, [&]() -> int {
        if (!to || !from || !mask || !to->class_ || !from->class_) return FDG_EINVAL;
        jetaa_merge(to->class_, to->nu, to->de, to->normal, from->class_, from->nu, from->de,
                    from->normal, mask, n);
        return FDG_OK;
    });
}

int fdg_frame_jet(fdg_frame* f, fdg_columns* out, fdg_stats* st, char* kernel_id,
                  size_t kernel_cap, uint64_t* computed) {
    return guarded<int>(f && f->ctx ? &f->ctx->err : nullptr, [&]() -> int {
        if (!f) return FDG_EINVAL;
        fdg_ctx* ctx = f->ctx;
        const fdg_params* p = &f->params;
        if (p->ss != 2) return ctx->fail(FDG_EINVAL, "jet-AA rebuilds ss 2 frames");
        const uint32_t need = FDG_COL_NU | FDG_COL_DE | FDG_COL_NORMAL;
        if ((p->columns & need) != need || !out || !out->class_ || !out->nu || !out->de ||
            !out->normal)
            return ctx->fail(FDG_EINVAL, "jet-AA needs the nu, de and normal columns");
        const size_t w = p->nx / 2, h = p->ny / 2, n = (size_t)f->count;
        // The passes set masks of their own; a sample mask the caller had set on the frame
        // is put back on every exit (it does not restrict the jet frame).
        const bool had_mask = f->masked;
        const std::vector<uint8_t> saved_mask = had_mask ? f->mask_host : std::vector<uint8_t>();
        auto restore = [&]() { return fdg_frame_mask(f, had_mask ? saved_mask.data() : nullptr); };
        std::vector<uint8_t> first(n), todo(n);
        jetaa_first(w, h, first.data());
        fdg_stats s1, s2;
        std::memset(&s2, 0, sizeof(s2));
        // pass 1: each pixel's top-left sample
        int rc = fdg_frame_mask(f, first.data());
        if (!rc) rc = fdg_frame_run(f, nullptr);
        if (!rc) rc = fdg_frame_read(f, out, nullptr, &s1, kernel_id, kernel_cap);
        if (rc) {
            (void)restore();
            return rc;
        }
        double t = now_s();
        jetaa_rebuild(out->class_, out->nu, out->de, out->normal, w, h, todo.data());
        double rebuild_s = now_s() - t;
        uint64_t rest = 0;
        for (uint8_t v : todo) rest += v;
        if (rest > 0) {
            // pass 2: the rejected predictions, read back packed and encoded straight into
            // their places over the rebuilt frame (jetaa::merge, jetaa.rs:248-261, without
            // a second set of columns: every other sample keeps its pass-1 or rebuilt value)
            rc = fdg_frame_mask(f, todo.data());
            if (!rc) rc = fdg_frame_run(f, &s2);
            if (!rc)
                rc = read_encode_list(ctx, f->buf, mask_to_list(todo.data(), n), n, f->fr.pl, p,
                                      out, nullptr, true, &s2.transfer_seconds,
                                      &s2.encode_seconds);
            if (rc) {
                (void)restore();
                return rc;
            }
        }
        rc = restore();
        if (rc) return rc;
        if (computed) *computed = w * h + rest;
        if (st) {
            fdg_stats s = s1;
            s.iterations += s2.iterations;
            s.equivalent += s2.equivalent;
            s.blocks += s2.blocks;
            s.skipped += s2.skipped;
            s.plain_steps += s2.plain_steps;
            s.fallback_samples += s2.fallback_samples;
            s.closed_form_samples += s2.closed_form_samples;
            s.uncertified_samples += s2.uncertified_samples;
            s.max_sample_ops = std::max(s.max_sample_ops, s2.max_sample_ops);
            s.wave_iterations += s2.wave_iterations;
            s.lane_iterations += s2.lane_iterations;
            s.bla_seconds += s2.bla_seconds;
            s.kernel_seconds += s2.kernel_seconds;
            // (each pass's transfers include the frame's orbit upload once)
            if (rest > 0) s.transfer_seconds += s2.transfer_seconds - f->transfer_seconds;
            s.encode_seconds += s2.encode_seconds + rebuild_s;
            *st = s;
        }
        return FDG_OK;
    });
}

int fdg_render(fdg_ctx* ctx, const fdg_view* v, const fdg_params* p, int route, int form,
               double bla_eps, fdg_columns* out, fdg_outcomes* raw, fdg_stats* st,
               char* kernel_id, size_t kernel_cap) {
    return guarded<int>(ctx ? &ctx->err : nullptr, [&]() -> int {
        if (!ctx) return FDG_EINVAL;
        if (!out || !out->class_) return ctx->fail(FDG_EINVAL, "the class column is required");
        fdg_frame* f = nullptr;
        int rc = fdg_frame_open(ctx, v, p, route, form, bla_eps, &f);
        if (rc) return rc;
        rc = fdg_frame_run(f, nullptr);
        if (!rc) rc = fdg_frame_read(f, out, raw, st, kernel_id, kernel_cap);
        fdg_frame_close(f);
        return rc;
    });
}

int fdg_jacobian(const fdg_view* v, const fdg_params* p, const fdg_outcomes* raw, double* j,
                 int64_t* jexp) {
    return guarded<int>(nullptr, [&]() -> int {
        if (!v || !p || !raw || !raw->kind || !raw->dr || !raw->di || !raw->dexp || !j || !jexp)
            return FDG_EINVAL;
        Plane pl;
        std::string e;
        if (!Plane::make(v->center_re, v->center_im, v->width, v->rotation, p->nx, p->ny, &pl, &e))
            return FDG_EINVAL;
        size_t n = (size_t)p->nx * p->ny;
        for (size_t i = 0; i < n; i++) {
            double* o = j + 4 * i;
            if (raw->kind[i] != FDG_ESCAPED) {
                o[0] = o[1] = o[2] = o[3] = 0.0;
                jexp[i] = 0;
                continue;
            }
            // dz/dc h_m, then the grid's basis: d c / d x = h (cos, sin), d c / d y = h (sin, -cos)
            double ar = raw->dr[i] * pl.h_m, ai = raw->di[i] * pl.h_m;
            double xr = ar * pl.cos_r - ai * pl.sin_r, xi = ar * pl.sin_r + ai * pl.cos_r;
            double yr = ar * pl.sin_r + ai * pl.cos_r, yi = ai * pl.sin_r - ar * pl.cos_r;
            o[0] = xr;  // d Re z / d x
            o[1] = yr;  // d Re z / d y
            o[2] = xi;  // d Im z / d x
            o[3] = yi;  // d Im z / d y
            jexp[i] = raw->dexp[i] + pl.h_e;
        }
        return FDG_OK;
    });
}

int64_t fdg_write_fds(const fdg_view* v, const fdg_params* p, const char* kernel_id,
                      const fdg_columns* c, uint8_t* buf, uint64_t cap) {
    return guarded<int64_t>(nullptr, [&]() -> int64_t {
        if (!v || !p || !c || !c->class_ || !kernel_id) return FDG_EINVAL;
        std::vector<uint8_t> b = write_fds(v->center_re, v->center_im, v->width, v->rotation, p->nx,
                                           p->ny, p->ss, p->max_iter, p->escape_radius, p->columns,
                                           kernel_id, c->class_, c->nu, c->de, c->normal);
        if (buf && cap >= b.size()) std::memcpy(buf, b.data(), b.size());
        return (int64_t)b.size();
    });
}

}  // extern "C"
