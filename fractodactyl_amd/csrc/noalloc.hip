// fdg_render_orbit_noalloc: a frame over a supplied orbit for callers that cannot allocate
// device memory.  Everything the pass needs beyond the caller's inputs and outputs — the
// orbit's device copy, the BLA table, the counters and the kernel descriptor — lives in
// one file-scope device array, `scratch`, of a fixed size.  The runtime gives each GPU
// its own copy of the array at its own address: fdg_create looks that address up once
// per device (hipGetSymbolAddress, with the device current) and every host use goes
// through it, never through the array's name (a host-side stand-in).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>

#include "internal.hpp"

namespace fdg {
namespace {

// Layout of the scratch array (byte offsets, each 256-aligned).
constexpr uint64_t kMaxPoints = FDG_NOALLOC_MAX_POINTS;
constexpr uint64_t kMaxStride = kMaxPoints - 2;  // next power of two >= points - 2
static_assert((kMaxStride & (kMaxStride - 1)) == 0, "table stride is a power of two");
constexpr uint64_t align256(uint64_t x) { return (x + 255) & ~(uint64_t)255; }
constexpr uint64_t kOffDesc = 0;                               // FrameDev
constexpr uint64_t kOffCounters = 512;                         // work index, statistics, levels
constexpr uint64_t kOffValid = kOffCounters + 256;             // per-level build flags
constexpr uint64_t kOffStageRe = kOffValid + 256;              // the caller's orbit, re then im
constexpr uint64_t kOffStageIm = kOffStageRe + align256(8 * kMaxPoints);
constexpr uint64_t kOffOrbit = kOffStageIm + align256(8 * kMaxPoints);  // (re, im) + 4 zero points
constexpr uint64_t kOffR2 = kOffOrbit + align256(16 * (kMaxPoints + 4));
constexpr uint64_t kOffAb = kOffR2 + 8 * kMaxStride;  // (r2's entry past the end lies in ab)
constexpr uint64_t kOffR = kOffAb + 32 * kMaxStride;
constexpr uint64_t kOffAlpha = kOffR + 8 * kMaxStride;
constexpr uint64_t kOffBeta = kOffAlpha + 8 * kMaxStride;
constexpr uint64_t kScratchBytes = kOffBeta + 8 * kMaxStride;
static_assert(kScratchBytes <= FDG_NOALLOC_SCRATCH_BYTES, "scratch size stated in fdg.h");
static_assert(sizeof(FrameDev) <= kOffCounters, "descriptor slot");
static_assert(kMaxLevels * 4 <= 256, "flag slot");

__device__ __align__(256) uint8_t scratch[FDG_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_ptr[kMaxDevices];      // the array's device address on each GPU
bool g_looked_up[kMaxDevices];  // the lookup ran (g_ptr null: it failed)

// The orbit interleaved as the kernels read it, (re, im) per point, followed by four zero
// points (the kernels' next-step prefetch reads up to two points past the end).
__global__ void interleave_orbit(const double* re, const double* im, double* orbit,
                                 uint32_t len) {
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < len) {
        orbit[2 * k] = re[k];
        orbit[2 * k + 1] = im[k];
    } else if (k < len + 4) {
        orbit[2 * k] = 0.0;
        orbit[2 * k + 1] = 0.0;
    }
}

__global__ void zero_flags(uint32_t* p, uint32_t n) {
    uint32_t i = threadIdx.x;
    if (i < n) p[i] = 0u;
}

}  // namespace

// fdg_create's set-up: the scratch array's address on `device`, looked up once per device
// with that device current.  A failed lookup is remembered; the entry point then reports it.
void noalloc_attach(int device) {
    if (device < 0 || device >= kMaxDevices) return;
    std::lock_guard<std::mutex> lock(g_mu);
    if (g_looked_up[device]) return;
    void* ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    g_ptr[device] = e == hipSuccess ? ptr : nullptr;
    g_looked_up[device] = true;
}

static uint8_t* scratch_of(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_mu);
    return static_cast<uint8_t*>(g_ptr[device]);
}

}  // namespace fdg

using namespace fdg;

extern "C" int fdg_render_orbit_noalloc(fdg_ctx* ctx, const fdg_view* v, const fdg_params* p,
                                        int route, double bla_eps, const double* re,
                                        const double* im, uint64_t len, uint32_t bits,
                                        fdg_outcomes* out, fdg_stats* st) {
    return guarded<int>(ctx ? &ctx->err : nullptr, [&]() -> int {
        if (!ctx) return FDG_EINVAL;
        if (!re || !im || !out || !out->kind || !out->n || !out->zr || !out->zi || !out->dr ||
            !out->di || !out->dexp || !out->ops)
            return ctx->fail(FDG_EINVAL, "an orbit and every outcome array are required");
        FDG_TRY(ctx, hipSetDevice(ctx->device), "hipSetDevice");
        uint8_t* base = scratch_of(ctx->device);
        if (!base) return ctx->fail(FDG_EDEVICE, "no scratch array address on this device");
        // The checks of fdg_frame_open_orbit (grid.rs:458-476, 779-799; bla.rs:27-33).
        Frame fr;
        int rc = setup(ctx, v, p, &fr);
        if (rc) return rc;
        if (p->columns & (FDG_COL_BOUND | FDG_COL_STRIPE))
            return ctx->fail(FDG_EDECLINED,
                             "bound/stripe columns are CPU-kernel columns (frame.rs:133-144)");
        if (route != FDG_ROUTE_PLAIN && route != FDG_ROUTE_BLA)
            return ctx->fail(FDG_EINVAL, "unknown route");
        const uint64_t count64 = (uint64_t)p->nx * p->ny;
        if (count64 >= (1ull << 31)) return ctx->fail(FDG_EINVAL, "grid too large");
        if (p->max_iter >= (1ull << 27)) return ctx->fail(FDG_EINVAL, "max_iter must be < 2^27");
        if (route == FDG_ROUTE_BLA) {
            if (std::isnan(p->escape_radius) || p->escape_radius < 4.0)
                return ctx->fail(FDG_EINVAL, "escape radius is below 4: BLA needs >= 4");
            if (!(bla_eps > 0.0 && bla_eps <= kEpsLimit))
                return ctx->fail(FDG_EINVAL, "BLA needs 0 < eps <= 2^-24");
        }
        const uint32_t need = fr.tier == TIER_F64 ? 53u : (uint32_t)fr.pl.bits;
        if (bits != need && (fr.tier == TIER_F64 || bits < need))
            return ctx->fail(FDG_EINVAL, "supplied reference is below the view's precision");
        if (len < (fr.tier == TIER_SCALED ? 3u : 2u) || re[0] != 0.0 || im[0] != 0.0)
            return ctx->fail(FDG_EINVAL,
                             "supplied reference orbit must start at Z_0 = 0 and hold at least "
                             "2 points (3 on the scaled tier)");
        if (len > kMaxPoints)
            return ctx->fail(FDG_EINVAL,
                             "orbit longer than the no-allocation scratch holds "
                             "(FDG_NOALLOC_MAX_POINTS)");
        fr.pl.bits = bits;
        const uint32_t count = (uint32_t)count64;
        hipStream_t s = ctx->stream;
        double t0 = now_s();
        const uint64_t nucleus = Orbit::nucleus_period_of(re, im, len);
        const double host_s = now_s() - t0;

        double* orbit = reinterpret_cast<double*>(base + kOffOrbit);
        double* stage_re = reinterpret_cast<double*>(base + kOffStageRe);
        double* stage_im = reinterpret_cast<double*>(base + kOffStageIm);
        unsigned long long* counters = reinterpret_cast<unsigned long long*>(base + kOffCounters);
        uint32_t* valid = reinterpret_cast<uint32_t*>(base + kOffValid);
        uint32_t* levels_dev = reinterpret_cast<uint32_t*>(base + kOffCounters + 192);
        FrameDev* desc = reinterpret_cast<FrameDev*>(base + kOffDesc);
        double t1 = now_s();
        FDG_TRY(ctx, hipMemcpyAsync(stage_re, re, 8 * len, hipMemcpyHostToDevice, s), "upload");
        FDG_TRY(ctx, hipMemcpyAsync(stage_im, im, 8 * len, hipMemcpyHostToDevice, s), "upload");
        hipLaunchKernelGGL(interleave_orbit, dim3((unsigned)((len + 4 + 255) / 256)), dim3(256),
                           0, s, stage_re, stage_im, orbit, (uint32_t)len);
        FDG_TRY(ctx, hipGetLastError(), "orbit layout");

        const uint64_t steps = len - 2;
        const uint64_t stride = next_pow2_host(std::max<uint64_t>(steps, 2));
        double dc = 0;
        int64_t dc_exp = 0;
        dc_bound(fr.pl, fr.tier, p->nx, p->ny, &dc, &dc_exp);
        const bool built = route == FDG_ROUTE_BLA && steps >= 2;
        if (route == FDG_ROUTE_BLA && !(dc > 0.0 && std::isfinite(dc)))
            return ctx->fail(FDG_EINVAL, "BLA needs a finite dc_max > 0");
        if (built) {
            // flags cleared by a fill kernel (no memset node, should a caller capture this)
            hipLaunchKernelGGL(zero_flags, dim3(1), dim3(64), 0, s, valid, kMaxLevels);
            FDG_TRY(ctx, hipGetLastError(), "flag clear");
            FDG_TRY(ctx, hipEventRecord(ctx->ev[0], s), "event");
            FDG_TRY(ctx,
                    launch_bla_build(orbit, reinterpret_cast<double*>(base + kOffR),
                                     reinterpret_cast<double*>(base + kOffR2),
                                     reinterpret_cast<double*>(base + kOffAb),
                                     reinterpret_cast<double*>(base + kOffAlpha),
                                     reinterpret_cast<double*>(base + kOffBeta), valid, steps,
                                     stride, bla_eps, dc, dc_exp, kMaxLevels, s),
                    "BLA build launch");
            FDG_TRY(ctx, hipEventRecord(ctx->ev[1], s), "event");
        }
        FrameDev d;
        std::memset(&d, 0, sizeof(d));
        const Plane& pl = fr.pl;
        d.orbit = orbit;
        d.len = len;
        d.nucleus = nucleus;
        d.bla.r2 = reinterpret_cast<const double*>(base + kOffR2);
        d.bla.ab = reinterpret_cast<const double*>(base + kOffAb);
        d.bla.steps = built ? steps : 0;
        d.bla.stride = stride;
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
        d.count = count;
        d.tier = fr.tier;
        d.max_iter = p->max_iter;
        d.r2 = p->escape_radius * p->escape_radius;
        d.kind = out->kind;
        d.n = out->n;
        d.zr = out->zr;
        d.zi = out->zi;
        d.dr = out->dr;
        d.di = out->di;
        d.dexp = out->dexp;
        d.ops = out->ops;
        d.next = reinterpret_cast<uint32_t*>(counters);
        d.stats = counters + 8;
        FDG_TRY(ctx,
                launch_pass_prologue(d, desc, built ? valid : nullptr, nullptr, steps, kMaxLevels,
                                     counters, 24, levels_dev, false, s),
                "pass set-up");
        const bool deriv = (p->columns & (FDG_COL_DE | FDG_COL_NORMAL)) != 0;
        int grid = (int)std::min<uint64_t>((count + 255) / 256, (uint64_t)ctx->cus * 8);
        if (grid < 1) grid = 1;
        FDG_TRY(ctx, hipEventRecord(ctx->ev[2], s), "event");
        FDG_TRY(ctx,
                launch_perturb_f64(desc, fr.tier == TIER_SCALED, deriv, built, false, grid, s),
                "perturbation launch");
        FDG_TRY(ctx, hipEventRecord(ctx->ev[3], s), "event");
        unsigned long long cnt[ST_COUNT];
        constexpr size_t kReadback = sizeof(cnt) + 8;  // statistics and the level count
        FDG_TRY(ctx,
                hipMemcpyAsync(ctx->small + fdg_ctx::kSlotCounts, counters + 8, kReadback,
                               hipMemcpyDeviceToHost, s),
                "readback");
        FDG_TRY(ctx, hipStreamSynchronize(s), "perturbation kernel");
        const double wall_s = now_s() - t1;
        std::memcpy(cnt, ctx->small + fdg_ctx::kSlotCounts, sizeof(cnt));
        uint32_t levels = 0;
        if (built) std::memcpy(&levels, ctx->small + fdg_ctx::kSlotCounts + sizeof(cnt), 4);
        if (built && levels == 0) cnt[ST_FALLBACK] = cnt[ST_CLOSED] = 0;  // as fdg_frame_run
        if (st) {
            fdg_stats r;
            std::memset(&r, 0, sizeof(r));
            r.reference_len = len;
            r.reference_bits = bits;
            r.bla_levels = levels;
            r.nucleus_period = nucleus;
            r.iterations = cnt[ST_ITERATIONS];
            r.equivalent = cnt[ST_EQUIVALENT];
            r.blocks = cnt[ST_BLOCKS];
            r.skipped = cnt[ST_SKIPPED];
            r.plain_steps = cnt[ST_PLAIN];
            r.fallback_samples = cnt[ST_FALLBACK];
            r.closed_form_samples = cnt[ST_CLOSED];
            r.max_sample_ops = cnt[ST_MAXOPS];
            r.wave_iterations = cnt[ST_WAVE_ITERS];
            r.lane_iterations = cnt[ST_LANE_ITERS];
            r.reference_seconds = host_s;
            float ms = 0;
            if (built) {
                FDG_TRY(ctx, hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]), "event timing");
                r.bla_seconds = ms * 1e-3;
            }
            FDG_TRY(ctx, hipEventElapsedTime(&ms, ctx->ev[2], ctx->ev[3]), "event timing");
            r.kernel_seconds = ms * 1e-3;
            r.transfer_seconds = wall_s - r.kernel_seconds - r.bla_seconds;
            *st = r;
        }
        return FDG_OK;
    });
}
