// C ABI for films (include/fdg.h fdg_film_*): the keyframes of a zoom at one centre,
// rendered against shared reference orbits as `fd reuse` does (crates/fd-cli/src/
// reuse.rs:1-60,124-142): the f64-tier frames share the 53-bit orbit at the rounded
// centre, every deeper frame the fixed-point orbit computed once at the deepest frame's
// precision, and each frame equals fd_kernel::render_with / render_bla from that orbit
// (grid.rs:95-109,190-210).
//
// Device side: one resident orbit (and its BLA tables) per orbit group; the outcomes of
// all frames in one frame-major array; one persistent work-stealing launch per delta
// arithmetic (f64-tier, fixed-tier Binary64, scaled-tier Binary64Exponent), the three
// captured once into a HIP graph over two streams and replayed.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "internal.hpp"

using namespace fdg;

namespace {

// One shared reference orbit of a film and the BLA tables built over it.
struct FilmOrbit {
    DevBuf orbit, r2, ab;
    uint64_t len = 0, nucleus = 0, steps = 0, stride = 0;
    uint32_t bits = 0;
    void release() {
        orbit.release();
        r2.release();
        ab.release();
    }
};

// Launch groups: the kernels' three arithmetic paths.
enum { G_F64 = 0, G_FIXED = 1, G_SCALED = 2, NGROUPS = 3 };

int group_of(Tier t) { return t == TIER_F64 ? G_F64 : (t == TIER_FIXED ? G_FIXED : G_SCALED); }
// Orbit 0 is the 53-bit f64-tier orbit, orbit 1 the deep fixed-point one (reuse.rs:124-142).
int orbit_of_group(int g) { return g == G_F64 ? 0 : 1; }

}  // namespace

struct fdg_film {
    fdg_ctx* ctx = nullptr;
    fdg_params params;
    int route = 0, form = 0, tables = 0;
    double eps = 0;
    uint32_t nframes = 0, per = 0;
    std::vector<Frame> fr;
    std::vector<int> group;             // launch group of each frame
    std::vector<uint32_t> levels;       // table levels each frame's lookups use
    std::vector<uint32_t> tb;           // first block of each frame's table in its orbit's
    FilmOrbit orb[2];
    uint32_t group_frames[NGROUPS] = {0, 0, 0};
    uint32_t group_levels[NGROUPS] = {0, 0, 0};
    double reference_seconds = 0, transfer_seconds = 0, bla_seconds = 0;
    Buffers buf;                        // outputs of every frame, table-merge scratch
    DevBuf batch_frames, fdev_batch, fdev_each, counters, valid;
    hipStream_t side = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int graph_grid = 0;
    bool ran = false;
    fdg_stats last;
    // the kernel descriptors as uploaded (a mask changes them), and the mask itself
    FrameDev group_desc[NGROUPS];
    std::vector<FrameDev> frame_desc;
    DevBuf mask;
    bool masked = false;

    uint64_t count() const { return (uint64_t)nframes * per; }
    int orbit_of(uint32_t k) const { return orbit_of_group(group[k]); }
    // 32 counter words per launch: the work index, then the statistics from word 8.
    unsigned long long* group_counters(int g) const {
        return counters.as<unsigned long long>() + 32 * g;
    }
    unsigned long long* frame_counters(uint32_t k) const {
        return counters.as<unsigned long long>() + 32 * (NGROUPS + k);
    }
    void release() {  // teardown: each step runs whatever the previous returned
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
        if (fork) (void)hipEventDestroy(fork);
        if (join) (void)hipEventDestroy(join);
        if (side) (void)hipStreamDestroy(side);
        side = nullptr;
        fork = join = nullptr;
        buf.release();
        for (FilmOrbit& o : orb) o.release();
        for (DevBuf* b : {&batch_frames, &fdev_batch, &fdev_each, &counters, &valid, &mask})
            b->release();
    }
};

namespace {

void add_counters(const unsigned long long* c, fdg_stats* s) {
    s->iterations += c[ST_ITERATIONS];
    s->equivalent += c[ST_EQUIVALENT];
    s->blocks += c[ST_BLOCKS];
    s->skipped += c[ST_SKIPPED];
    s->fallback_samples += c[ST_FALLBACK];
    s->closed_form_samples += c[ST_CLOSED];
    s->plain_steps += c[ST_PLAIN];
    s->max_sample_ops = std::max<uint64_t>(s->max_sample_ops, c[ST_MAXOPS]);
    s->wave_iterations += c[ST_WAVE_ITERS];
    s->lane_iterations += c[ST_LANE_ITERS];
}

// The tables of one orbit: one per frame at that frame's dc bound, or one shared at the
// largest bound over the orbit's frames (compared as bla_checks compares, grid.rs:399-405).
// Table t sits at block t * stride of the orbit's r2/ab arrays in the padded level layout;
// the merge scratch (r, alpha, beta) is shared, the builds being ordered on one stream.
int build_orbit_tables(fdg_film* f, int o, std::vector<uint32_t>* valid_base,
                       std::vector<std::vector<uint32_t>>* frames_of_table) {
    fdg_ctx* ctx = f->ctx;
    FilmOrbit& ob = f->orb[o];
    std::vector<uint32_t> members;
    for (uint32_t k = 0; k < f->nframes; k++)
        if (f->orbit_of(k) == o) members.push_back(k);
    if (members.empty() || ob.steps < 2) return FDG_OK;
    bool shared = f->tables == FDG_TABLES_SHARED;
    uint32_t ntab = shared ? 1 : (uint32_t)members.size();
    if ((uint64_t)ntab * ob.stride >= (1ull << 32))
        return ctx->fail(FDG_EINVAL, "film tables exceed 2^32 blocks");
    std::vector<double> dc(ntab);
    std::vector<int64_t> de(ntab);
    for (size_t i = 0; i < members.size(); i++) {
        uint32_t k = members[i], t = shared ? 0 : (uint32_t)i;
        double d;
        int64_t e;
        dc_bound(f->fr[k].pl, f->fr[k].tier, f->params.nx, f->params.ny, &d, &e);
        bool take = i == 0 || !shared ||
                    (e == de[t] ? d > dc[t]
                                : std::log2(d) + (double)e > std::log2(dc[t]) + (double)de[t]);
        if (take) {
            dc[t] = d;
            de[t] = e;
        }
        f->tb[k] = t * (uint32_t)ob.stride;
    }
    // (one entry past the last table: the level search's untaken loads may land there)
    FDG_TRY(ctx, ob.r2.ensure((ntab * ob.stride + 1) * 8), "hipMalloc tables");
    FDG_TRY(ctx, ob.ab.ensure(ntab * ob.stride * 32), "hipMalloc tables");
    Buffers& B = f->buf;
    for (uint32_t t = 0; t < ntab; t++) {
        if (!(dc[t] > 0.0 && std::isfinite(dc[t])))
            return ctx->fail(FDG_EINVAL, "BLA needs a finite dc_max > 0");
        uint32_t slot = (uint32_t)valid_base->size();
        valid_base->push_back(slot * kMaxLevels);
        std::vector<uint32_t> fr;
        for (size_t i = 0; i < members.size(); i++)
            if (shared || i == t) fr.push_back(members[i]);
        frames_of_table->push_back(fr);
        FDG_TRY(ctx,
                launch_bla_build(ob.orbit.as<double>(), B.bla_r.as<double>(),
                                 ob.r2.as<double>() + t * ob.stride,
                                 ob.ab.as<double>() + 4 * t * ob.stride, B.bla_alpha.as<double>(),
                                 B.bla_beta.as<double>(),
                                 f->valid.as<uint32_t>() + slot * kMaxLevels, ob.steps, ob.stride,
                                 f->eps, dc[t], de[t], kMaxLevels, ctx->stream),
                "BLA build launch");
    }
    return FDG_OK;
}

int film_tables(fdg_film* f) {
    fdg_ctx* ctx = f->ctx;
    f->levels.assign(f->nframes, 0);
    f->tb.assign(f->nframes, 0);
    uint64_t scratch = 0, ntab_max = 0;
    for (FilmOrbit& ob : f->orb) {
        if (!ob.len) continue;
        scratch = std::max(scratch, ob.stride);
        ntab_max += f->tables == FDG_TABLES_SHARED ? 1 : f->nframes;
    }
    Buffers& B = f->buf;
    FDG_TRY(ctx, B.bla_r.ensure(scratch * 8), "hipMalloc");
    FDG_TRY(ctx, B.bla_alpha.ensure(scratch * 8), "hipMalloc");
    FDG_TRY(ctx, B.bla_beta.ensure(scratch * 8), "hipMalloc");
    FDG_TRY(ctx, f->valid.ensure(ntab_max * kMaxLevels * 4), "hipMalloc");
    FDG_TRY(ctx, hipMemsetAsync(f->valid.p, 0, ntab_max * kMaxLevels * 4, ctx->stream), "memset");
    std::vector<uint32_t> valid_base;
    std::vector<std::vector<uint32_t>> frames_of_table;
    std::vector<int> orbit_of_table;
    FDG_TRY(ctx, hipEventRecord(ctx->ev[0], ctx->stream), "event");
    for (int o = 0; o < 2; o++) {
        int rc = build_orbit_tables(f, o, &valid_base, &frames_of_table);
        if (rc) return rc;
        orbit_of_table.resize(valid_base.size(), o);
    }
    FDG_TRY(ctx, hipEventRecord(ctx->ev[1], ctx->stream), "event");
    std::vector<uint32_t> valid(valid_base.size() * kMaxLevels);
    if (!valid.empty())
        FDG_TRY(ctx,
                hipMemcpyAsync(valid.data(), f->valid.p, valid.size() * 4,
                               hipMemcpyDeviceToHost, ctx->stream),
                "readback");
    FDG_TRY(ctx, hipStreamSynchronize(ctx->stream), "BLA build");
    float ms = 0;
    FDG_TRY(ctx, hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]), "event timing");
    f->bla_seconds = ms * 1e-3;
    for (size_t t = 0; t < valid_base.size(); t++) {
        FilmOrbit& ob = f->orb[orbit_of_table[t]];
        uint32_t n = 0;
        while (n < kMaxLevels && (ob.steps >> (n + 1)) > 0 && valid[valid_base[t] + n]) n++;
        // A lookup walks up from level 0 and stops at the first radius it fails, so
        // clearing level n ends every walk in this table where Bla::find's would (levels
        // end before the first level without a valid block, bla.rs:442-445) when the
        // launch's level count is that of a deeper table.
        uint64_t c = ob.steps >> (n + 1);
        uint32_t k0 = frames_of_table[t][0];
        if (c > 0)
            FDG_TRY(ctx,
                    launch_zero_words(reinterpret_cast<unsigned long long*>(
                                          ob.r2.as<double>() + f->tb[k0] +
                                          bla_level_offset(ob.stride, n)),
                                      (uint32_t)c, ctx->stream),
                    "table clear");
        for (uint32_t k : frames_of_table[t]) f->levels[k] = n;
    }
    FDG_TRY(ctx, hipStreamSynchronize(ctx->stream), "table clear");
    return FDG_OK;
}

FrameDev orbit_frame_dev(const fdg_film* f, int o) {
    FrameDev d;
    std::memset(&d, 0, sizeof(d));
    const FilmOrbit& ob = f->orb[o];
    const Buffers& b = f->buf;
    const fdg_params* p = &f->params;
    d.orbit = ob.orbit.as<double>();
    d.len = ob.len;
    d.nucleus = ob.nucleus;
    d.bla.r2 = ob.r2.as<double>();
    d.bla.ab = ob.ab.as<double>();
    d.bla.steps = ob.steps;
    d.bla.stride = ob.stride;
    d.nx = p->nx;
    d.ny = p->ny;
    d.max_iter = p->max_iter;
    d.r2 = p->escape_radius * p->escape_radius;
    d.kind = b.o_kind.as<uint8_t>();
    d.n = b.o_n.as<uint64_t>();
    d.zr = b.o_zr.as<double>();
    d.zi = b.o_zi.as<double>();
    d.dr = b.o_dr.as<double>();
    d.di = b.o_di.as<double>();
    d.dexp = b.o_dexp.as<int64_t>();
    d.ops = b.o_ops.as<uint32_t>();
    return d;
}

bool deriv_of(const fdg_film* f) {
    return (f->params.columns & (FDG_COL_DE | FDG_COL_NORMAL)) != 0;
}

// The graph: clear the launch counters with a fill kernel (no memset nodes), then the
// fixed-tier batch on the main stream and the scaled-tier and f64-tier batches on the
// side stream, joined back.
int film_capture(fdg_film* f, int grid) {
    fdg_ctx* ctx = f->ctx;
    // a previous capture (another grid) is replaced; its teardown has nothing to report
    if (f->exec) (void)hipGraphExecDestroy(f->exec);
    if (f->graph) (void)hipGraphDestroy(f->graph);
    f->exec = nullptr;
    f->graph = nullptr;
    bool deriv = deriv_of(f);
    const FrameDev* fd = f->fdev_batch.as<FrameDev>();
    FDG_TRY(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal),
            "graph capture");
    hipError_t e = launch_zero_words(f->group_counters(0), 32 * NGROUPS, ctx->stream);
    if (e == hipSuccess) e = hipEventRecord(f->fork, ctx->stream);
    if (e == hipSuccess) e = hipStreamWaitEvent(f->side, f->fork, 0);
    for (int g : {G_SCALED, G_F64})
        if (e == hipSuccess && f->group_frames[g])
            e = launch_perturb_f64(fd + g, g == G_SCALED, deriv, f->group_levels[g] > 0, true,
                                   grid, f->side);
    if (e == hipSuccess && f->group_frames[G_FIXED])
        e = launch_perturb_f64(fd + G_FIXED, false, deriv, f->group_levels[G_FIXED] > 0, true,
                               grid, ctx->stream);
    if (e == hipSuccess) e = hipEventRecord(f->join, f->side);
    if (e == hipSuccess) e = hipStreamWaitEvent(ctx->stream, f->join, 0);
    hipGraph_t g = nullptr;
    hipError_t e2 = hipStreamEndCapture(ctx->stream, &g);
    if (e != hipSuccess || e2 != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);  // the capture's own error is the one reported
        return ctx->hip(e != hipSuccess ? e : e2, "graph capture");
    }
    f->graph = g;
    FDG_TRY(ctx, hipGraphInstantiate(&f->exec, g, nullptr, nullptr, 0), "graph instantiate");
    f->graph_grid = grid;
    return FDG_OK;
}

}  // namespace

extern "C" {

int fdg_film_open(fdg_ctx* ctx, const fdg_view* views, uint32_t nframes, const fdg_params* p,
                  int route, int form, double bla_eps, int tables, fdg_film** out) {
    return guarded<int>(ctx ? &ctx->err : nullptr, [&]() -> int {
        if (!ctx || !out) return FDG_EINVAL;
        *out = nullptr;
        if (!views || nframes == 0) return ctx->fail(FDG_EINVAL, "a film needs at least one view");
        if (!p) return ctx->fail(FDG_EINVAL, "null params");
        FDG_TRY(ctx, hipSetDevice(ctx->device), "hipSetDevice");
        if (p->columns & (FDG_COL_BOUND | FDG_COL_STRIPE))
            return ctx->fail(FDG_EDECLINED,
                             "bound/stripe columns are CPU-kernel columns (frame.rs:133-144)");
        if (form != FDG_FORM_F64) return ctx->fail(FDG_EINVAL, "unknown delta form");
        if (route != FDG_ROUTE_PLAIN && route != FDG_ROUTE_BLA)
            return ctx->fail(FDG_EINVAL, "unknown route");
        if (tables != FDG_TABLES_PER_FRAME && tables != FDG_TABLES_SHARED)
            return ctx->fail(FDG_EINVAL, "unknown table mode");
        if (p->max_iter >= (1ull << 27)) return ctx->fail(FDG_EINVAL, "max_iter must be < 2^27");
        uint64_t per = (uint64_t)p->nx * p->ny;
        if (per * nframes >= (1ull << 31)) return ctx->fail(FDG_EINVAL, "film too large");
        if (route == FDG_ROUTE_BLA) {
            if (std::isnan(p->escape_radius) || p->escape_radius < 4.0)
                return ctx->fail(FDG_EINVAL, "escape radius is below 4: BLA needs >= 4");
            if (!(bla_eps > 0.0 && bla_eps <= kEpsLimit))
                return ctx->fail(FDG_EINVAL, "BLA needs 0 < eps <= 2^-24");
        }
        std::vector<Frame> fr(nframes);
        int lead[2] = {-1, -1};  // per orbit: the frame it is computed for (deepest, earliest)
        for (uint32_t k = 0; k < nframes; k++) {
            int rc = setup(ctx, &views[k], p, &fr[k]);
            if (rc) return rc;
            if (!same_decimal(views[k].center_re, views[0].center_re) ||
                !same_decimal(views[k].center_im, views[0].center_im))
                return ctx->fail(FDG_EINVAL, "film frames must share one centre");
            int o = orbit_of_group(group_of(fr[k].tier));
            if (lead[o] < 0 || fr[k].pl.bits > fr[lead[o]].pl.bits) lead[o] = (int)k;
        }
        fdg_film* f = new fdg_film();
        struct Guard {
            fdg_film* f;
            bool keep = false;
            ~Guard() {
                if (!keep) {
                    f->release();
                    delete f;
                }
            }
        } guard{f};
        f->ctx = ctx;
        f->params = *p;
        f->route = route;
        f->form = form;
        f->tables = tables;
        f->eps = bla_eps;
        f->nframes = nframes;
        f->per = (uint32_t)per;
        f->fr = fr;
        for (uint32_t k = 0; k < nframes; k++) f->group.push_back(group_of(fr[k].tier));
        for (int o = 0; o < 2; o++) {
            if (lead[o] < 0) continue;
            double t0 = now_s();
            Orbit orbit;
            int rc = compute_orbit(ctx, &views[lead[o]], p, fr[lead[o]], &orbit);
            if (rc) return rc;
            // as a single frame refuses them: the loop reads Z_1, and the scaled tier's
            // rebase from a delta below 1e-270 continues to Z_2
            if (orbit.len() < (o == 1 ? 3u : 2u))
                return ctx->fail(FDG_EINVAL, "reference orbit too short for the film's frames");
            f->reference_seconds += now_s() - t0;
            FilmOrbit& ob = f->orb[o];
            ob.len = orbit.len();
            ob.bits = orbit.bits;
            ob.nucleus = orbit.nucleus_period();
            ob.steps = ob.len >= 2 ? ob.len - 2 : 0;
            ob.stride = next_pow2_host(std::max<uint64_t>(ob.steps, 2));
            // grid.rs:474-476: the plane records the supplied orbit's precision
            for (uint32_t k = 0; k < nframes; k++)
                if (f->orbit_of(k) == o && o == 1) f->fr[k].pl.bits = orbit.bits;
            rc = upload_orbit(ctx, &ob.orbit, orbit, &f->transfer_seconds);
            if (rc) return rc;
        }
        int rc = f->buf.outputs(ctx, f->count());
        if (rc) return rc;
        f->levels.assign(nframes, 0);
        f->tb.assign(nframes, 0);
        if (route == FDG_ROUTE_BLA) {
            rc = film_tables(f);
            if (rc) return rc;
        }
        // batch groups, frames in order within each; one BatchFrameDev array
        std::vector<BatchFrameDev> bd;
        uint32_t first[NGROUPS];
        for (int g = 0; g < NGROUPS; g++) {
            first[g] = (uint32_t)bd.size();
            for (uint32_t k = 0; k < nframes; k++) {
                if (f->group[k] != g) continue;
                const Plane& pl = f->fr[k].pl;
                BatchFrameDev b;
                b.h = pl.h();
                b.h_m = pl.h_m;
                b.h_e = pl.h_e;
                b.cos_r = pl.cos_r;
                b.sin_r = pl.sin_r;
                b.half_x = pl.half_x;
                b.half_y = pl.half_y;
                b.out0 = k * f->per;
                b.tb = f->tb[k];
                bd.push_back(b);
                f->group_frames[g]++;
                f->group_levels[g] = std::max(f->group_levels[g], f->levels[k]);
            }
        }
        FDG_TRY(ctx, f->batch_frames.ensure(bd.size() * sizeof(BatchFrameDev)), "hipMalloc");
        FDG_TRY(ctx, f->fdev_batch.ensure(NGROUPS * sizeof(FrameDev)), "hipMalloc");
        FDG_TRY(ctx, f->fdev_each.ensure(nframes * sizeof(FrameDev)), "hipMalloc");
        FDG_TRY(ctx, f->counters.ensure((size_t)(NGROUPS + nframes) * 256), "hipMalloc");
        FrameDev gd[NGROUPS];
        for (int g = 0; g < NGROUPS; g++) {
            gd[g] = orbit_frame_dev(f, orbit_of_group(g));
            gd[g].tier = g;  // G_F64 / G_FIXED / G_SCALED are the tiers' values
            gd[g].frames = f->batch_frames.as<BatchFrameDev>() + first[g];
            gd[g].nframes = f->group_frames[g];
            gd[g].per_frame = f->per;
            gd[g].count = f->group_frames[g] * f->per;
            gd[g].bla.nlevels = f->group_levels[g];
            gd[g].next = reinterpret_cast<uint32_t*>(f->group_counters(g));
            gd[g].stats = f->group_counters(g) + 8;
            if (f->group_frames[g]) {
                const Plane& pl = f->fr[bd[first[g]].out0 / f->per].pl;
                gd[g].c_re = pl.c_re;  // the rounded centre, the same for every frame
                gd[g].c_im = pl.c_im;
            }
        }
        std::vector<FrameDev> each(nframes);
        for (uint32_t k = 0; k < nframes; k++) {
            const Plane& pl = f->fr[k].pl;
            FrameDev d = orbit_frame_dev(f, f->orbit_of(k));
            uint64_t off = (uint64_t)k * f->per;
            d.bla.r2 += f->tb[k];
            d.bla.ab += 4 * (uint64_t)f->tb[k];
            d.bla.nlevels = f->levels[k];
            d.c_re = pl.c_re;
            d.c_im = pl.c_im;
            d.h = pl.h();
            d.h_m = pl.h_m;
            d.h_e = pl.h_e;
            d.cos_r = pl.cos_r;
            d.sin_r = pl.sin_r;
            d.half_x = pl.half_x;
            d.half_y = pl.half_y;
            d.count = f->per;
            d.tier = f->fr[k].tier;
            d.kind += off;
            d.n += off;
            d.zr += off;
            d.zi += off;
            d.dr += off;
            d.di += off;
            d.dexp += off;
            d.ops += off;
            d.next = reinterpret_cast<uint32_t*>(f->frame_counters(k));
            d.stats = f->frame_counters(k) + 8;
            each[k] = d;
        }
        FDG_TRY(ctx,
                hipMemcpy(f->batch_frames.p, bd.data(), bd.size() * sizeof(BatchFrameDev),
                          hipMemcpyHostToDevice),
                "upload");
        FDG_TRY(ctx, hipMemcpy(f->fdev_batch.p, gd, sizeof(gd), hipMemcpyHostToDevice), "upload");
        std::memcpy(f->group_desc, gd, sizeof(gd));
        f->frame_desc = each;
        FDG_TRY(ctx,
                hipMemcpy(f->fdev_each.p, each.data(), nframes * sizeof(FrameDev),
                          hipMemcpyHostToDevice),
                "upload");
        FDG_TRY(ctx, hipStreamCreate(&f->side), "side stream");
        FDG_TRY(ctx, hipEventCreateWithFlags(&f->fork, hipEventDisableTiming), "event");
        FDG_TRY(ctx, hipEventCreateWithFlags(&f->join, hipEventDisableTiming), "event");
        guard.keep = true;
        *out = f;
        return FDG_OK;
    });
}

int fdg_film_run(fdg_film* f, int launch, int grid, fdg_stats* st) {
    return guarded<int>(f && f->ctx ? &f->ctx->err : nullptr, [&]() -> int {
        if (!f) return FDG_EINVAL;
        fdg_ctx* ctx = f->ctx;
        FDG_TRY(ctx, hipSetDevice(ctx->device), "hipSetDevice");
        if (launch != FDG_LAUNCH_BATCH && launch != FDG_LAUNCH_PER_FRAME)
            return ctx->fail(FDG_EINVAL, "unknown launch mode");
        if (grid < 0 || grid > 65536) return ctx->fail(FDG_EINVAL, "grid must be in [0, 65536]");
        if (grid == 0) grid = ctx->cus * 8;
        fdg_stats s;
        std::memset(&s, 0, sizeof(s));
        int o = f->orb[1].len ? 1 : 0;
        s.reference_seconds = f->reference_seconds;
        s.reference_len = f->orb[o].len;
        s.reference_bits = f->orb[o].bits;
        s.nucleus_period = f->orb[o].nucleus;
        s.transfer_seconds = f->transfer_seconds;
        s.bla_seconds = f->bla_seconds;
        s.bla_levels = *std::max_element(f->levels.begin(), f->levels.end());
        std::vector<unsigned long long> cnt;
        if (f->masked)
            // samples outside the mask keep this "not computed" kind through the pass (a fill
            // on the stream, outside the graph)
            FDG_TRY(ctx, hipMemsetAsync(f->buf.o_kind.p, kKindUnset, f->count(), ctx->stream),
                    "fill");
        if (launch == FDG_LAUNCH_BATCH) {
            if (!f->exec || f->graph_grid != grid) {
                int rc = film_capture(f, grid);
                if (rc) return rc;
            }
            FDG_TRY(ctx, hipEventRecord(ctx->ev[2], ctx->stream), "event");
            FDG_TRY(ctx, hipGraphLaunch(f->exec, ctx->stream), "graph launch");
            FDG_TRY(ctx, hipEventRecord(ctx->ev[3], ctx->stream), "event");
            cnt.resize(32 * NGROUPS);
            FDG_TRY(ctx,
                    hipMemcpyAsync(cnt.data(), f->group_counters(0), cnt.size() * 8,
                                   hipMemcpyDeviceToHost, ctx->stream),
                    "readback");
            FDG_TRY(ctx, hipStreamSynchronize(ctx->stream), "film kernels");
            for (int g = 0; g < NGROUPS; g++) add_counters(cnt.data() + 32 * g + 8, &s);
        } else {
            bool deriv = deriv_of(f);
            FDG_TRY(ctx, launch_zero_words(f->frame_counters(0), 32 * f->nframes, ctx->stream),
                    "counter clear");
            FDG_TRY(ctx, hipEventRecord(ctx->ev[2], ctx->stream), "event");
            int g = (int)std::min<uint64_t>((uint64_t)grid, (f->per + 255) / 256);
            for (uint32_t k = 0; k < f->nframes; k++)
                FDG_TRY(ctx,
                        launch_perturb_f64(f->fdev_each.as<FrameDev>() + k,
                                           f->fr[k].tier == TIER_SCALED, deriv, f->levels[k] > 0,
                                           false, std::max(g, 1), ctx->stream),
                        "perturbation launch");
            FDG_TRY(ctx, hipEventRecord(ctx->ev[3], ctx->stream), "event");
            cnt.resize(32 * (size_t)f->nframes);
            FDG_TRY(ctx,
                    hipMemcpyAsync(cnt.data(), f->frame_counters(0), cnt.size() * 8,
                                   hipMemcpyDeviceToHost, ctx->stream),
                    "readback");
            FDG_TRY(ctx, hipStreamSynchronize(ctx->stream), "film kernels");
            for (uint32_t k = 0; k < f->nframes; k++) add_counters(cnt.data() + 32 * k + 8, &s);
        }
        float ms = 0;
        FDG_TRY(ctx, hipEventElapsedTime(&ms, ctx->ev[2], ctx->ev[3]), "event timing");
        s.kernel_seconds = ms * 1e-3;
        f->last = s;
        f->ran = true;
        if (st) *st = s;
        return FDG_OK;
    });
}

int fdg_film_read(fdg_film* f, uint32_t k, fdg_columns* out, fdg_outcomes* raw, fdg_stats* st,
                  char* kernel_id, size_t kernel_cap) {
    return guarded<int>(f && f->ctx ? &f->ctx->err : nullptr, [&]() -> int {
        if (!f) return FDG_EINVAL;
        fdg_ctx* ctx = f->ctx;
        if (!f->ran) return ctx->fail(FDG_EINVAL, "film has not been run");
        if (k >= f->nframes) return ctx->fail(FDG_EINVAL, "frame index out of range");
        const fdg_params* p = &f->params;
        if (!out || !out->class_) return ctx->fail(FDG_EINVAL, "the class column is required");
        if ((p->columns & FDG_COL_NU && !out->nu) || (p->columns & FDG_COL_DE && !out->de) ||
            (p->columns & FDG_COL_NORMAL && !out->normal))
            return ctx->fail(FDG_EINVAL, "a requested column has no buffer");
        FDG_TRY(ctx, hipSetDevice(ctx->device), "hipSetDevice");
        fdg_stats s = f->last;
        s.transfer_seconds = f->transfer_seconds;
        std::vector<uint64_t> nn;
        std::vector<uint8_t> kind;
        std::vector<uint32_t> ops;
        int rc = read_encode(ctx, f->buf, (size_t)k * f->per, f->per, f->fr[k].pl, p, out, raw,
                             &s.transfer_seconds, &s.encode_seconds, &nn, &kind, &ops);
        if (rc) return rc;
        const FilmOrbit& ob = f->orb[f->orbit_of(k)];
        s.reference_len = ob.len;
        s.reference_bits = ob.bits;
        s.nucleus_period = ob.nucleus;
        s.bla_levels = f->levels[k];
        // This frame's Stats (grid.rs:640-650): iterations counts a block as one iterate and
        // stops at max_iter; equivalent counts every step.  n holds every sample's stop
        // iterate, so spent - skipped + blocks = spent - n + (plain steps + blocks).
        bool bla = f->levels[k] > 0;
        uint64_t its = 0, eq = 0, mx = 0;
        for (size_t i = 0; i < f->per; i++) {
            uint64_t spent = kind[i] == KIND_UNRESOLVED ? p->max_iter : nn[i];
            eq += spent;
            its += bla ? spent - nn[i] + ops[i] : spent;
            mx = std::max<uint64_t>(mx, ops[i]);
        }
        s.iterations = its;
        s.equivalent = eq;
        s.max_sample_ops = mx;
        // launch-wide counters are not split by frame
        s.blocks = s.skipped = s.plain_steps = s.fallback_samples = s.closed_form_samples = 0;
        s.wave_iterations = s.lane_iterations = 0;
        if (st) *st = s;
        if (kernel_id && kernel_cap) {
            std::string kid = f->fr[k].pl.kernel(f->fr[k].tier);
            if (f->levels[k] > 0) kid += " bla/1";
            kid += " gpu-hip-f64/1";
            std::snprintf(kernel_id, kernel_cap, "%s", kid.c_str());
        }
        return FDG_OK;
    });
}

int fdg_film_mask(fdg_film* f, const uint8_t* masks) {
    return guarded<int>(f && f->ctx ? &f->ctx->err : nullptr, [&]() -> int {
        if (!f) return FDG_EINVAL;
        fdg_ctx* ctx = f->ctx;
        FDG_TRY(ctx, hipSetDevice(ctx->device), "hipSetDevice");
        const size_t n = f->count();
        if (masks) {
            FDG_TRY(ctx, f->mask.ensure(n), "hipMalloc");
            FDG_TRY(ctx, hipMemcpy(f->mask.p, masks, n, hipMemcpyHostToDevice), "mask upload");
        }
        f->masked = masks != nullptr;
        // the batch kernels index the mask by output sample, the per-frame ones by the
        // frame's own sample
        FrameDev gd[NGROUPS];
        std::memcpy(gd, f->group_desc, sizeof(gd));
        std::vector<FrameDev> each = f->frame_desc;
        const uint8_t* m = f->masked ? f->mask.as<uint8_t>() : nullptr;
        for (auto& d : gd) d.mask = m;
        for (uint32_t k = 0; k < f->nframes; k++) each[k].mask = m ? m + (size_t)k * f->per : nullptr;
        FDG_TRY(ctx, hipMemcpy(f->fdev_batch.p, gd, sizeof(gd), hipMemcpyHostToDevice), "upload");
        FDG_TRY(ctx,
                hipMemcpy(f->fdev_each.p, each.data(), each.size() * sizeof(FrameDev),
                          hipMemcpyHostToDevice),
                "upload");
        return FDG_OK;
    });
}

int fdg_film_jet(fdg_film* f, fdg_columns* outs, uint64_t* computed) {
    return guarded<int>(f && f->ctx ? &f->ctx->err : nullptr, [&]() -> int {
        if (!f) return FDG_EINVAL;
        fdg_ctx* ctx = f->ctx;
        const fdg_params* p = &f->params;
        if (p->ss != 2) return ctx->fail(FDG_EINVAL, "jet-AA rebuilds ss 2 frames");
        const uint32_t need = FDG_COL_NU | FDG_COL_DE | FDG_COL_NORMAL;
        if ((p->columns & need) != need || !outs)
            return ctx->fail(FDG_EINVAL, "jet-AA needs the nu, de and normal columns");
        for (uint32_t k = 0; k < f->nframes; k++)
            if (!outs[k].class_ || !outs[k].nu || !outs[k].de || !outs[k].normal)
                return ctx->fail(FDG_EINVAL, "jet-AA needs the nu, de and normal columns");
        const size_t w = p->nx / 2, h = p->ny / 2, per = f->per, n = f->count();
        // pass 1: every frame's top-left samples, in one batched launch
        std::vector<uint8_t> masks(n);
        jetaa_first(w, h, masks.data());
        for (uint32_t k = 1; k < f->nframes; k++)
            std::memcpy(masks.data() + (size_t)k * per, masks.data(), per);
        int rc = fdg_film_mask(f, masks.data());
        if (!rc) rc = fdg_film_run(f, FDG_LAUNCH_BATCH, 0, nullptr);
        for (uint32_t k = 0; k < f->nframes && !rc; k++)
            rc = fdg_film_read(f, k, &outs[k], nullptr, nullptr, nullptr, 0);
        if (rc) {
            (void)fdg_film_mask(f, nullptr);
            return rc;
        }
        // the rebuilds, then pass 2: every frame's rejected samples in one launch
        uint64_t rest = 0;
        for (uint32_t k = 0; k < f->nframes; k++) {
            uint8_t* todo = masks.data() + (size_t)k * per;
            jetaa_rebuild(outs[k].class_, outs[k].nu, outs[k].de, outs[k].normal, w, h, todo);
            for (size_t i = 0; i < per; i++) rest += todo[i];
        }
        if (rest > 0) {
            rc = fdg_film_mask(f, masks.data());
            if (!rc) rc = fdg_film_run(f, FDG_LAUNCH_BATCH, 0, nullptr);
            std::vector<uint8_t> c2(per);
            std::vector<double> nu2(per);
            std::vector<float> de2(per);
            std::vector<uint16_t> nm2(per);
            fdg_columns second{c2.data(), nu2.data(), de2.data(), nm2.data()};
            for (uint32_t k = 0; k < f->nframes && !rc; k++) {
                rc = fdg_film_read(f, k, &second, nullptr, nullptr, nullptr, 0);
                if (!rc)
                    jetaa_merge(outs[k].class_, outs[k].nu, outs[k].de, outs[k].normal, c2.data(),
                                nu2.data(), de2.data(), nm2.data(), masks.data() + (size_t)k * per,
                                per);
            }
            if (rc) {
                (void)fdg_film_mask(f, nullptr);
                return rc;
            }
        }
        if (computed) *computed = (uint64_t)f->nframes * w * h + rest;
        return fdg_film_mask(f, nullptr);
    });
}

void fdg_film_close(fdg_film* f) {
    guarded_void([&] {
        if (!f) return;
        (void)hipSetDevice(f->ctx->device);  // teardown: nothing to report to
        (void)hipStreamSynchronize(f->ctx->stream);
        if (f->side) (void)hipStreamSynchronize(f->side);
        f->release();
        delete f;
    });
}

}  // extern "C"
