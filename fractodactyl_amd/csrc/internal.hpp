// Host-side internals shared by the C ABI's translation units (api.cpp: contexts, single
// frames; film.cpp: batches of frames sharing an orbit).
#pragma once
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "abi_guard.hpp"
#include "fdg.h"
#include "host.hpp"
#include "kernels.hpp"

namespace fdg {

hipError_t launch_bla_build(const double* orbit, double* r, double* r2, double* ab, double* alpha,
                            double* beta, uint32_t* valid, uint64_t steps, uint64_t stride,
                            double eps, double dc_max, int64_t dc_exp, uint32_t max_levels,
                            hipStream_t stream);
hipError_t launch_perturb_f64(const FrameDev* f, bool ext, bool deriv, bool bla, bool batch,
                              int grid, hipStream_t stream, bool list = false);
hipError_t launch_zero_words(unsigned long long* p, uint32_t n, hipStream_t stream);
// A single frame's pass set-up in one launch: the descriptor d stored at dst with the
// level count (from the build's flags `valid`, else *levels_in), nzero (<= 64) counter
// words zeroed, the count copied to *levels_out, and with clear the flags reset.
hipError_t launch_pass_prologue(const FrameDev& d, FrameDev* dst, uint32_t* valid,
                                const uint32_t* levels_in, uint64_t steps, uint32_t max_levels,
                                unsigned long long* counters, uint32_t nzero,
                                uint32_t* levels_out, bool clear, hipStream_t stream);
// The orbit as FP32 mantissa pairs with exponents; with q (nullable, one radius per point)
// each point also carries log2 of its error radius.
hipError_t launch_orbit_to_f32(const double* orbit, const double* q, float* o32, uint64_t points,
                               hipStream_t stream);
hipError_t launch_table_to_f32(const double* ab, float* ab32, int32_t* abexp, uint64_t blocks,
                               hipStream_t stream);
// The mixed form's refinement list: the computed samples the FP32 pass did not certify
// (cert[i] == 0; with cert null, the escaped ones).
hipError_t launch_escaped_list(const uint8_t* kind, const uint8_t* cert, uint32_t* list,
                               uint32_t* count, uint32_t n, hipStream_t stream);
// The FP64 kernel on the ordinary tiers with every sample starting from a series state
// (FrameDev::series_*); unresolved series starts retry from zero in the same launch.
hipError_t launch_perturb_f64_series(const FrameDev* f, bool deriv, bool bla, int grid,
                                     hipStream_t stream);
// Store::put on the device for e.count samples (encode.hip).
hipError_t launch_encode_columns(const EncodeDev& e, int cus, hipStream_t stream);
hipError_t launch_gather_outcomes(const GatherDev& g, int cus, hipStream_t stream);
// cert: the mixed form's certifying pass (error bound per sample, FrameDev::cert written).
hipError_t launch_perturb_fe32(const FrameDev* f, bool deriv, bool bla, int grid,
                               hipStream_t stream, bool cert = false);

// The no-allocation entry point's set-up (noalloc.hip): looks up the scratch array's address
// on `device` once, with that device current; fdg_create calls it.
void noalloc_attach(int device);

constexpr uint32_t kMaxLevels = 48;
// Outcome kind of a sample outside a sample mask (not computed; its columns stay unset).
constexpr uint8_t kKindUnset = 0xff;
constexpr double kEpsLimit = 1.0 / 16777216.0;  // bla.rs:288-292 EPS_LIMIT = 2^-24


// A grow-only device allocation.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    hipError_t ensure(size_t n) {
        if (n <= bytes) return hipSuccess;
        if (p) (void)hipFree(p);  // the allocation below reports any device failure
        p = nullptr;
        bytes = 0;
        hipError_t e = hipMalloc(&p, n);
        if (e == hipSuccess) bytes = n;
        return e;
    }
    template <class T>
    T* as() const {
        return static_cast<T*>(p);
    }
    void release() {
        if (p) (void)hipFree(p);  // teardown: nothing to report to
        p = nullptr;
        bytes = 0;
    }
};

inline double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

}  // namespace fdg

namespace fdg {
struct Buffers;
}

struct fdg_ctx {
    int device = 0;
    // The device buffers of the last frame closed on this context, handed to the next frame
    // opened (an end-to-end caller that opens and closes a frame per render reuses them
    // instead of allocating ~53 bytes per sample each time); freed with the context.
    fdg::Buffers* spare = nullptr;
    int cus = 256;
    hipStream_t stream = nullptr;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    std::string err;
    // the autotuner's choices by workload shape (fdg_frame_tune), applied to frames opened
    // later with the same shape
    std::map<std::string, fdg_launch> tuned;
    void* pinned = nullptr;  // grow-only page-locked staging for outcome readback
    size_t pinned_bytes = 0;
    // Page-locked slots for a frame pass's small readbacks (statistics and the table's level
    // count): copies into pageable memory are staged by the runtime and cost more than
    // these few hundred bytes.
    uint8_t* small = nullptr;
    static constexpr size_t kSmallBytes = 4096;
    static constexpr size_t kSlotCounts = 0, kSlotCounts2 = 1024;

    int fail(int code, const std::string& m) {
        err = m;
        return code;
    }
    // At least n bytes of pinned staging (hipHostMalloc'd: page-locked from the start).
    int staging(size_t n) {
        if (n <= pinned_bytes) return 0;
        if (pinned) (void)hipHostFree(pinned);  // the allocation below reports failures
        pinned = nullptr;
        pinned_bytes = 0;
        hipError_t e = hipHostMalloc(&pinned, n, hipHostMallocDefault);
        if (e != hipSuccess) return hip(e, "hipHostMalloc staging");
        pinned_bytes = n;
        return 0;
    }
    int hip(hipError_t e, const char* what) {
        if (e == hipSuccess) return FDG_OK;
        err = std::string(what) + ": " + hipGetErrorString(e);
        return e == hipErrorOutOfMemory ? FDG_ENOMEM : FDG_EDEVICE;
    }
};

#define FDG_TRY(ctx, call, what)                       \
    do {                                               \
        int rc_ = (ctx)->hip((call), (what));          \
        if (rc_ != FDG_OK) return rc_;                 \
    } while (0)

namespace fdg {

// Device residency of one frame: the orbit, its BLA table and the per-sample outcomes.
struct Buffers {
    DevBuf orbit, bla_r, bla_r2, bla_ab, bla_alpha, bla_beta, bla_valid;
    DevBuf o_kind, o_n, o_zr, o_zi, o_dr, o_di, o_dexp, o_ops, counters, fdev;
    DevBuf orbit32, ab32, abexp32;  // FP32 copies for the FE32 and mixed forms
    DevBuf qrad, cert;              // FP32 forms: orbit error radii, per-sample certificates
    DevBuf list, counters2, fdev2;  // the mixed form's FP64 refinement pass
    DevBuf c_cls, c_nu, c_de, c_nm;  // columns encoded on the device
    DevBuf packed;                  // a masked pass's computed samples, gathered for readback
    uint64_t bla_stride = 0, bla_steps = 0;
    // bla_valid holds zeros (a pass set-up cleared the flags after reading them), so the
    // next build needs no clearing of its own
    bool valid_clean = false;
    void release() {
        valid_clean = false;
        for (DevBuf* b : {&orbit, &bla_r, &bla_r2, &bla_ab, &bla_alpha, &bla_beta, &bla_valid,
                          &o_kind, &o_n, &o_zr, &o_zi, &o_dr, &o_di, &o_dexp, &o_ops,
                          &counters, &fdev, &orbit32, &ab32, &abexp32, &qrad, &cert, &list, &counters2,
                          &fdev2,
                          &c_cls, &c_nu, &c_de, &c_nm, &packed})
            b->release();
    }
    // Per-sample outcome arrays for n samples, the counters and one FrameDev.
    int outputs(fdg_ctx* ctx, size_t n);
    // Take over `from`'s allocations (from then owns none); the table-build state starts
    // afresh.
    void adopt(Buffers& from) {
        *this = from;
        from = Buffers();
        bla_stride = bla_steps = 0;
        valid_clean = false;
    }
};

struct Frame {
    Plane pl;
    Tier tier;
};

// grid.rs:779-799 setup: plane, tier (and its validity at the depth).
int setup(fdg_ctx* ctx, const fdg_view* v, const fdg_params* p, Frame* fr);
// grid.rs:802-815 compute: the frame's own reference orbit.
int compute_orbit(fdg_ctx* ctx, const fdg_view* v, const fdg_params* p, const Frame& fr,
                  Orbit* o);
uint64_t next_pow2_host(uint64_t x);
// Upload an orbit interleaved (re, im) with a 64-byte pad into *dst.
int upload_orbit(fdg_ctx* ctx, DevBuf* dst, const Orbit& o, double* seconds);
// orbit.rs:228-247 `same`: two decimal strings name the same number.
bool same_decimal(const std::string& a, const std::string& b);
// Copy n outcomes from offset off of B's arrays, encode the requested columns (Store::put)
// and fill raw (Unresolved samples report n = 0).  ops_out (nullable) receives the
// per-sample operation counts and n_out (nullable) the stop iterates, before zeroing.
int read_encode(fdg_ctx* ctx, const Buffers& B, size_t off, size_t n, const Plane& pl,
                const fdg_params* p, fdg_columns* out, fdg_outcomes* raw, double* transfer_s,
                double* encode_s, std::vector<uint64_t>* n_out, std::vector<uint8_t>* kind_out,
                std::vector<uint32_t>* ops_out);
// The set samples of a mask, in order.
std::vector<uint32_t> mask_to_list(const uint8_t* mask, size_t n);
// read_encode for a masked pass: only the samples list[0 .. m) (in order) are gathered on
// the device, read back packed and encoded into their places; every other sample reads as
// unset (class 0, zero columns, raw kind 0xff and zeros), or is left as it is with
// keep_unset (a second pass merged over a first one's columns).
int read_encode_list(fdg_ctx* ctx, Buffers& b, const std::vector<uint32_t>& list, size_t n,
                     const Plane& pl, const fdg_params* p, fdg_columns* out, fdg_outcomes* raw,
                     bool keep_unset, double* transfer_s, double* encode_s);

}  // namespace fdg
