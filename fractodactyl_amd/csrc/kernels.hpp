// Kernel-argument layouts shared by the host launch code and the HIP kernels.
#pragma once
#include <cstdint>

namespace fdg {

// BLA table in HBM.  Level i (block length 2 << i, i = 0 .. nlevels-1) holds
// steps >> (i + 1) blocks, stored from offset stride - (stride >> i) where stride is the
// next power of two >= steps; block t of level i starts at reference index 1 + t 2^(i+1)
// (crates/fd-kernel/src/bla.rs:399-419).  Arrays are structure-of-arrays so the radius
// walk touches only r2.
struct BlaDev {
    const double* r2;       // squared validity radius (0 when never valid)
    const double* ab;       // 4 doubles per block: a.re, a.im, b.re, b.im
    uint64_t steps;         // reference points - 2
    uint64_t stride;        // next power of two >= steps (level-offset base)
    uint32_t nlevels;       // levels actually used by lookups
};

__host__ __device__ inline uint64_t bla_level_offset(uint64_t stride, uint32_t i) {
    return stride - (stride >> i);
}

// Statistics accumulated by one perturbation launch (grid.rs:40-70 Stats / BlaStats).
enum StatSlot : int {
    ST_ITERATIONS = 0,  // Stats::iterations: blocks count as one iterate
    ST_EQUIVALENT = 1,  // sum of Outcome::iterations (BLA-skipped steps included)
    ST_BLOCKS = 2,
    ST_SKIPPED = 3,
    ST_FALLBACK = 4,    // samples that applied no block (excluding closed-form)
    ST_CLOSED = 5,      // closed-form main-cardioid / period-2 samples
    ST_PLAIN = 6,       // plain perturbation steps executed
    ST_MAXOPS = 7,      // most operations (plain steps + blocks) one sample executed
    ST_WAVE_ITERS = 8,  // loop iterations summed over waves
    ST_LANE_ITERS = 9,  // active lanes summed over those iterations (utilisation)
    ST_UNCERT = 10,     // FP32 forms: non-escaped samples whose error bound reached the
                        // escape radius (not certified)
    ST_COUNT = 16
};

// Geometry of one frame of a batch whose frames share the reference orbit (the keyframes
// of a zoom film at one centre; grid.rs:95-101 "one deep orbit serves every shallower
// frame at its centre").  Frames of a batch have the same grid, so sample g of the batch
// is sample g % per_frame of the batch's frame g / per_frame.
struct BatchFrameDev {
    double h, h_m;
    int64_t h_e;
    double cos_r, sin_r, half_x, half_y;
    uint32_t out0;          // output index of the frame's first sample
    uint32_t tb;            // first block of the frame's BLA table (0: the shared table)
};

// One frame (or one batch of frames) for the perturbation kernels.
struct FrameDev {
    const double* orbit;    // 2 doubles per point: Z_m re, im
    uint64_t len;           // reference points
    uint64_t nucleus;       // Reference::nucleus_period, 0 = none
    BlaDev bla;
    // sample geometry (view.rs:117-128)
    double c_re, c_im, h, h_m;
    int64_t h_e;
    double cos_r, sin_r, half_x, half_y;
    uint32_t nx, ny, count;
    int32_t tier;           // 0 F64, 1 Fixed, 2 Scaled
    uint64_t max_iter;
    double r2;              // escape radius squared
    const uint8_t* mask;    // nullable: compute only mask[i] != 0
    // outputs (structure of arrays, count each)
    uint8_t* kind;
    uint64_t* n;                    // iterate reached (also for Unresolved: the stop point)
    double* zr;
    double* zi;
    double* dr;
    double* di;
    int64_t* dexp;
    uint32_t* ops;                  // executed plain steps + applied blocks per sample
    uint32_t* next;                 // work-stealing counter (zeroed by the host)
    unsigned long long* stats;      // ST_COUNT counters (zeroed by the host)
    // FP32 delta form only: the orbit as float pairs and the table coefficients
    // (A, B) as float quadruples, in the same layouts as orbit / bla.ab
    const float* orbit32;
    const float* ab32;
    const int32_t* abexp32;         // per block: the exponents of A and B
    uint32_t defer_escaped;         // FP32 pass of the mixed form: the samples it does not
                                    // certify are recomputed in FP64, so they add no
                                    // statistics here
    uint8_t* cert;                  // FP32 forms (nullable): per sample 1 = certified (an
                                    // error bound shows the exact iteration stays below the
                                    // escape radius), 0 = escaped or not certified
    const double* bla_alpha;        // FP32 forms: the table's state envelopes (alpha, beta),
    const double* bla_beta;         // same layout as bla.r2
    const uint32_t* list;           // nullable: compute only samples list[0 .. *list_count)
    const uint32_t* list_count;
    // Series start (the series kernels only, ordinary tiers): every sample starts at
    // iterate series_n with delta sum b[k] x^k, x = dc / rho (series.rs:236-257), b[1..]
    // as series_len (re, im) pairs.
    const double* series_b;
    uint32_t series_n, series_len;
    double series_inv_rho;          // 1 / rho, as Series::eval computes it
    // batches only (the single-frame kernels read the geometry above)
    const BatchFrameDev* frames;    // nframes entries
    uint32_t nframes, per_frame;    // count = nframes * per_frame
    // claim policy (0 = default): a wave's claim is the work left / (claim_div x waves),
    // at most claim_cap samples (a multiple of 64), at least its idle lanes; a wave refills
    // once `refill` (1-64, default 64) of its lanes are idle
    uint32_t claim_div, claim_cap, refill;
};

// The sample (row-major index in an nx x ny grid) that claimed index `my` stands for.
// Claims hand out whole groups of 64 indices; a group is an 8 x 8 tile of the grid (tiles in
// row-major order) when both sides are multiples of 8, else 64 consecutive samples of the
// row-major scan.  A wave refilled all at once then starts its lanes on neighbouring
// samples, whose orbits stay close (the same blocks at the same orbit indices) for most of
// their iterations.
__device__ inline uint32_t sample_of(uint32_t my, uint32_t nx, uint32_t ny) {
    if ((nx | ny) & 7u) return my;
    const uint32_t t = my >> 6, q = my & 63u, tiles_x = nx >> 3;
    return ((t / tiles_x) * 8u + (q >> 3)) * nx + (t % tiles_x) * 8u + (q & 7u);
}

// One frame pass's outcomes and the columns the device encoder writes from them
// (Store::put, store.rs:66-153); nu / de / normal are null when not requested.
struct EncodeDev {
    const uint8_t* kind;
    const uint64_t* n;
    const double *zr, *zi, *dr, *di;
    const int64_t* dexp;
    uint8_t* cls;
    double* nu;
    float* de;
    uint16_t* normal;
    uint32_t count;
    double px_m;            // output pixel size h_m ss, times 2^px_e
    int64_t px_e;
    double cos_r, sin_r;
};

// A masked pass's computed samples (list[0 .. m), sample indices) gathered from the
// outcome arrays into packed ones for the readback.
struct GatherDev {
    const uint32_t* list;
    uint32_t m;
    const uint8_t* kind;
    const uint64_t* n;
    const double *zr, *zi, *dr, *di;
    const int64_t* dexp;
    const uint32_t* ops;            // (read when ops_c is set)
    uint8_t* kind_c;
    uint64_t* n_c;
    double *zr_c, *zi_c, *dr_c, *di_c;
    int64_t* dexp_c;
    uint32_t* ops_c;                // nullable
};

}  // namespace fdg
