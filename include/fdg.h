/*
 * fdg.h - C ABI of the MI355X deep-zoom perturbation/BLA path.
 *
 * The drop-in boundary for fractodactyl's GPU perturbation route: the Rust host
 * (crates/fd-cli/src/perturb_gpu.rs, perturb_gpu/frame.rs) binds these entry points
 * through a thin FFI crate instead of dispatching WGSL through wgpu.  Plain C types
 * only; no exception crosses the boundary; every call returns 0 on success or a
 * negative FDG_E* code with a message from fdg_last_error().
 *
 * Threading: a context owns one HIP device, one stream and its resident buffers; use a
 * context from one host thread at a time.  Several contexts may coexist.
 *
 * Exceptions: every entry point catches what the C++ runtime may throw inside it — a
 * host allocation that fails returns FDG_ENOMEM, any other exception FDG_EDEVICE — and
 * keeps the message for fdg_last_error (of the context, or of the calling thread when the
 * call has none).
 */
#ifndef FDG_H
#define FDG_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FDG_ABI_VERSION 1

enum {
    FDG_OK = 0,
    FDG_EINVAL = -1,    /* invalid request (the reference's Err(String) cases) */
    FDG_EDECLINED = -2, /* a request this route does not serve (Bound/Stripe columns) */
    FDG_EDEVICE = -3,   /* HIP runtime / device failure */
    FDG_ENOMEM = -4
};

/* Columns: fd_samples::ColumnSet bits (crates/fd-samples/src/column.rs:9-23). */
enum { FDG_COL_CLASS = 1, FDG_COL_NU = 2, FDG_COL_DE = 4, FDG_COL_NORMAL = 8,
       FDG_COL_BOUND = 16, FDG_COL_STRIPE = 32 };
/* Tiers: fd_kernel::Tier (crates/fd-kernel/src/view.rs:6-14); -1 = cheapest valid. */
enum { FDG_TIER_AUTO = -1, FDG_TIER_F64 = 0, FDG_TIER_FIXED = 1, FDG_TIER_SCALED = 2 };
/* Routes: `--gpu-perturb on|bla` (crates/fd-cli/src/perturb_gpu/frame.rs:62-84). */
enum { FDG_ROUTE_PLAIN = 0, FDG_ROUTE_BLA = 1 };
/* Delta arithmetic.  F64 reproduces the reference CPU kernel bit for bit (Binary64 on
 * the f64/fx tiers, Binary64Exponent on the scaled tier).  FE32 ("floatexp") iterates
 * FP32 mantissas with carried int32 exponents on every tier, with FP64 only in the BLA
 * validity test and the host encoding: close to, not equal to, the reference's results;
 * it has no Brent/Newton periodicity confirmation, so samples the reference proves
 * Interior that way end Unresolved (the counterpart of the reference GPU route's
 * float-float CE words, crates/fd-cli/src/perturb_gpu/frame.rs:15-117).  MIXED runs the
 * FE32 pass, then recomputes in F64 every sample it found escaped: escaped samples equal
 * the F64 form's, the others are the FE32 pass's (the reverse of the reference route's
 * admit-escaped / restart-the-rest split, frame.rs:338-341). */
enum { FDG_FORM_F64 = 0, FDG_FORM_FE32 = 1, FDG_FORM_MIXED = 2 };
/* Outcome kinds (crates/fd-kernel/src/sample.rs:21-40). */
enum { FDG_ESCAPED = 0, FDG_INTERIOR = 1, FDG_UNRESOLVED = 2 };

typedef struct fdg_ctx fdg_ctx;

/* fd_samples::View (crates/fd-samples/src/header.rs:263-273): exact decimal strings. */
typedef struct {
    const char *center_re;
    const char *center_im;
    const char *width;
    double rotation;
} fdg_view;

/* fd_kernel::Params (crates/fd-kernel/src/grid.rs:17-38); worker threads and stripe
 * parameters have no meaning here. */
typedef struct {
    uint32_t nx, ny, ss;
    uint64_t max_iter;  /* below 2^27: orbit and table offsets are formed in 32 bits */
    double escape_radius;
    uint32_t columns;
    int32_t tier;
} fdg_params;

/* fd_samples::Samples columns (crates/fd-samples/src/samples.rs:6-21), caller-allocated
 * with nx*ny elements each; a column not requested may be NULL. */
typedef struct {
    uint8_t *class_;
    double *nu;
    float *de;
    uint16_t *normal;
} fdg_columns;

/* Raw per-sample outcomes before quantisation (sample.rs:21-40); any member may be NULL.
 * z and dz/dc = (dr, di) 2^dexp are meaningful for escaped samples only; n is 0 for
 * unresolved ones. */
typedef struct {
    uint8_t *kind;
    uint64_t *n;
    double *zr, *zi, *dr, *di;
    int64_t *dexp;
    uint32_t *ops; /* operations executed: plain steps + applied BLA blocks */
} fdg_outcomes;

/* Work counters (grid.rs:40-70 Stats + BlaStats) and timings. */
typedef struct {
    uint64_t reference_len;
    uint32_t reference_bits;
    uint32_t bla_levels;
    uint64_t nucleus_period;
    uint64_t iterations;          /* Stats::iterations (an applied block counts once) */
    uint64_t equivalent;          /* plain-equivalent iterations (skipped steps counted) */
    uint64_t blocks, skipped, plain_steps;
    uint64_t fallback_samples, closed_form_samples;
    uint64_t max_sample_ops;      /* most operations executed by one sample */
    uint64_t wave_iterations;     /* kernel loop iterations summed over wavefronts */
    uint64_t lane_iterations;     /* active lanes summed over those iterations */
    double reference_seconds;     /* host arbitrary-precision orbit */
    double bla_seconds;           /* GPU table build (device time) */
    double kernel_seconds;        /* GPU perturbation kernel (device time) */
    double transfer_seconds;      /* uploads + outcome readback */
    double encode_seconds;        /* Store::put into the columns (host) */
    uint64_t uncertified_samples; /* mixed form: samples whose FP32 error bound reached
                                     the escape radius before they ended (the FP32 pass
                                     stops them there; they are recomputed in FP64 with
                                     the escaped ones) */
    uint64_t reference_period;    /* the host fixed-point orbit's state repeated exactly
                                     with this period at point reference_repeat_at, so its
                                     later points were copied, not computed (a nucleus
                                     centre, or an attracting cycle the truncated orbit
                                     settled on); 0: no exact repeat */
    uint64_t reference_repeat_at;
} fdg_stats;

int fdg_abi_version(void);

/* Open a context on HIP device `device`.  No counterpart in the reference beyond the
 * wgpu device open in crates/fd-shade/src/gpu/testkit.rs:12-47. */
int fdg_create(fdg_ctx **out, int device);
void fdg_destroy(fdg_ctx *ctx);
/* Message for the last non-zero status on this context, or, with ctx NULL, of the calling
 * thread's last failed call without a context (fdg_create, host-only helpers). */
const char *fdg_last_error(const fdg_ctx *ctx);

/* perturb_gpu::enabled (frame.rs:62-68): "on" -> 1 route plain, "bla" -> 1 route BLA,
 * "off"/NULL -> 0; anything else is FDG_EINVAL. */
int fdg_enabled(const char *value, int *enabled, int *route);

/* perturb_gpu::render / render_bla (frame.rs:70-84, render_inner :119-355).  Renders every
 * sample of the frame on the GPU: the reference orbit (the exact-centre fixed-point
 * orbit of Reference::from_fixed, or Reference::new on the f64 tier) is computed on the
 * host; route FDG_ROUTE_BLA builds a Bla table on the GPU over it with `bla_eps` for the
 * frame's |dc| bound (as `fd control --bla per-frame` does, control.rs:356-366) and
 * applies its blocks.  Writes the requested columns, exactly fd_kernel::render_with /
 * render_bla's values for form FDG_FORM_F64, optionally the raw outcomes, and the header
 * kernel id (plane kernel, " bla/1" when blocks were used, then " gpu-hip-f64/1", or
 * " gpu-hip-fe32/1" / " gpu-hip-mixed/1" for forms FE32 / MIXED). */
int fdg_render(fdg_ctx *ctx, const fdg_view *view, const fdg_params *params, int route, int form,
               double bla_eps, fdg_columns *out, fdg_outcomes *raw, fdg_stats *stats,
               char *kernel_id, size_t kernel_cap);

/* The same frame split so that its inputs stay resident in HBM between passes, for
 * callers that render a view repeatedly (benchmarks, progressive refinement, films): no
 * separate counterpart in the reference, where render_inner (frame.rs:119-355) does all
 * four steps in one call.  fdg_frame_open computes the reference orbit on the host and
 * uploads it; fdg_frame_run builds the BLA table (route BLA) and runs the perturbation
 * kernel, leaving the outcomes on the device; fdg_frame_read copies them back and encodes
 * the columns (as fdg_render); fdg_frame_close frees the frame's device memory. */
typedef struct fdg_frame fdg_frame;
int fdg_frame_open(fdg_ctx *ctx, const fdg_view *view, const fdg_params *params, int route,
                   int form, double bla_eps, fdg_frame **out);
int fdg_frame_run(fdg_frame *frame, fdg_stats *stats);
int fdg_frame_read(fdg_frame *frame, fdg_columns *out, fdg_outcomes *raw, fdg_stats *stats,
                   char *kernel_id, size_t kernel_cap);
void fdg_frame_close(fdg_frame *frame);

/* fdg_frame_read with Store::put (crates/fd-kernel/src/store.rs:66-153) on the GPU: the
 * outcomes are encoded into the columns in HBM and only the 15 bytes per sample of the
 * columns are read back (instead of 53 of raw outcomes).  The class column equals
 * fdg_frame_read's; nu, de and normal are computed with the device's log2 / atan2, so they
 * may differ from the host encoding in the last bit (nu, de) or by one quantum (normal) —
 * as the reference's own columns differ between platforms' libms.  Non-escaped samples
 * read nu 0, de 0, normal 0. */
int fdg_frame_read_columns(fdg_frame *frame, fdg_columns *out, fdg_stats *stats,
                           char *kernel_id, size_t kernel_cap);

/* fd_kernel::render_with_masked / render_bla_masked (crates/fd-kernel/src/grid.rs:173-209):
 * compute only the samples whose mask entry (one byte per sample, row-major over the
 * sample grid, non-zero = compute) is set; the others are left unset — class 0, nu 0, de 0,
 * normal 0 — and their raw outcome reads kind 255 with zeros.  The statistics count the
 * computed samples only.  fdg_frame_mask sets the mask for the following runs of a
 * resident frame (NULL clears it), so the two passes of a jet-AA frame
 * (crates/fd-cli/src/jetaa.rs:1-45) share one orbit and table upload; fdg_render_masked is
 * the one-call form. */
int fdg_frame_mask(fdg_frame *frame, const uint8_t *mask);

/* Launch tuning of a frame's persistent perturbation kernel.  No counterpart in the
 * reference: its CPU renderer has no launch, and its wgpu route a fixed workgroup shape.
 * Outputs and sample statistics do not depend on it (each sample is computed alone, in
 * any order); only the wave statistics (wave_iterations, lane_iterations) do.  Every
 * field 0 means its default. */
typedef struct fdg_launch {
    uint32_t groups_per_cu; /* persistent workgroups of 256 lanes per CU, 1-32; default 8 */
    uint32_t claim_div;     /* a wave claims the work left / (claim_div x waves), 1-64;
                               default 4 */
    uint32_t claim_cap;     /* at most this many samples per claim, a multiple of 64 in
                               [64, 16384]; default 1024 */
    uint32_t refill;        /* a wave takes new samples once this many of its 64 lanes are
                               idle, 1-64; default 64 (whole waves: a wave's samples, an
                               8 x 8 tile of the grid, run in lockstep) */
} fdg_launch;

/* Set the launch tuning of the frame's following runs (NULL: the defaults). */
int fdg_frame_launch(fdg_frame *frame, const fdg_launch *launch);

/* The autotuner: times `reps` (1-16) runs of the frame under each candidate launch
 * (groups per CU 4, 8, 16 x claim divisor 2, 4, 8 x claim cap 256, 1024, 4096) by the
 * perturbation kernel's event time, keeps the fastest (its minimum over the reps) for the
 * frame's following runs and for frames the context opens later with the same shape
 * (form, tier, derivative, BLA, and the power-of-two buckets of the sample count and
 * max_iter), and reports it in `best` (nullable) with its time and the default launch's
 * (best_ms, default_ms: minimum kernel milliseconds; nullable).  The frame's outputs after
 * it are those of its last run. */
int fdg_frame_tune(fdg_frame *frame, uint32_t reps, fdg_launch *best, double *best_ms,
                   double *default_ms);
int fdg_render_masked(fdg_ctx *ctx, const fdg_view *view, const fdg_params *params, int route,
                      int form, double bla_eps, const uint8_t *mask, fdg_columns *out,
                      fdg_outcomes *raw, fdg_stats *stats, char *kernel_id, size_t kernel_cap);

/* The parameter-series initializer (crates/fd-kernel/src/series.rs; fd_kernel::Series,
 * render_series at grid.rs:218-226).  With dc = rho x, |x| <= 1 over the frame, the delta
 * after n steps is S_n(x) = sum b[k] x^k; fdg_series_build runs Series::build (the
 * growing-order recurrence, its state/derivative majorants and admission guard, stopping
 * 32 steps after the last admitted one) on the host over an orbit — the one fdg_reference
 * returns for the frame — and writes the admitted start: n (0 = none admitted), the
 * coefficients b[1..b_len] as (re, im) pairs into the caller's b (b_cap >= max_order pairs)
 * and the bounds, bit for bit as the reference computes them.  rho is the frame's |dc|
 * bound (fdg_bla_dc_max, fd_kernel::bla_dc_max at grid.rs:432-440; refused on the scaled
 * tier) and pixel the output pixel size in c.  cfg NULL = SeriesConfig::default
 * (series.rs:43-55).
 *
 * fdg_frame_series sets a series start for the following runs of a resident frame (NULL
 * clears it); fdg_render_series is the one-call form (mask nullable), the counterpart of
 * Series::render_masked (series.rs:342-374) with its refusals: the scaled tier, a series
 * over another orbit (points), a view reaching beyond rho; forms other than FP64 are
 * declined.  Every sample starts at iterate n from its series state; an Unresolved one
 * retries from zero in the same launch (grid.rs:623-649), and the statistics count the
 * work as the reference does (the skipped prefix is not work).  Kernel id: " series/0"
 * after " bla/1". */
typedef struct {
    uint32_t start_order, max_order;
    double grow_at, shift_px, derivative_rel;
} fdg_series_config;
enum { FDG_SERIES_STOP_REFERENCE = 0, FDG_SERIES_STOP_RADIUS, FDG_SERIES_STOP_NONFINITE,
       FDG_SERIES_STOP_ESCAPE, FDG_SERIES_STOP_BUDGET, FDG_SERIES_STOP_STALL };
typedef struct {
    uint64_t n;            /* iterates every sample skips */
    double rho;            /* normalising radius */
    uint64_t points;       /* orbit points it was built over */
    double *b;             /* caller-owned: b[1..b_len] as (re, im) pairs */
    uint32_t b_cap, b_len;
    double state_bound, derivative_bound, shift_px;
    uint64_t built_steps, coefficient_pairs, growth_steps;
    int32_t stop;          /* FDG_SERIES_STOP_* */
} fdg_series;
void fdg_series_config_default(fdg_series_config *cfg);
int fdg_series_build(const double *re, const double *im, uint64_t len, double rho, double pixel,
                     uint64_t max_iter, const fdg_series_config *cfg, fdg_series *out);
int fdg_bla_dc_max(const fdg_view *view, const fdg_params *params, double *dc_max);
int fdg_frame_series(fdg_frame *frame, const fdg_series *series);
/* Series::prepare (series.rs:229-340) on a resident FP64 frame: declines (no series, *admitted
 * 0) below 16,384 active samples, on the scaled tier and below 256 iterations; otherwise
 * builds the series over the frame's orbit for max_iter capped at 8192 into *out, renders
 * a probe (every 61st active sample and the 16 x 16 centre block, max_iter capped at 8192
 * as the reference's probe) without and with it on the frame's route, and sets it on the
 * frame (*admitted 1) when the series arm's predicted cost is below 0.8x: the larger of
 * the frame's operations (scaled from the probe) over the chip's parallel step rate and
 * the longest sample's chain of operations.  The reference decides by timing nine samples
 * on the CPU; predicting from counted work makes the decision reproducible.  The probe
 * runs overwrite the frame's outcomes: run the frame afterwards. */
int fdg_frame_series_prepare(fdg_frame *frame, const fdg_series_config *cfg, fdg_series *out,
                             int *admitted);
int fdg_render_series(fdg_ctx *ctx, const fdg_view *view, const fdg_params *params, int route,
                      double bla_eps, const fdg_series *series, const uint8_t *mask,
                      fdg_columns *out, fdg_outcomes *raw, fdg_stats *stats, char *kernel_id,
                      size_t kernel_cap);

/* Jet-AA for ss 2 film frames (crates/fd-cli/src/jetaa.rs; film.rs:1359-1457
 * render_frame_jet).  On a 2w x 2h sample grid: fdg_jetaa_first writes the pass-1 mask
 * (each pixel's top-left sample, jetaa.rs:38-44); fdg_jetaa_rebuild predicts every pixel's
 * other three samples from the pass-1 columns (nu, de and normal required), writes the
 * accepted predictions into the columns and sets todo (one byte per sample) where pass 2
 * must compute (jetaa.rs:123-245); fdg_jetaa_merge copies the masked samples of a second
 * pass (jetaa.rs:248-261).  All three are host-only and equal the reference's bytes.
 * fdg_frame_jet runs the whole jet-AA frame on a resident ss 2 frame: pass 1 on the GPU,
 * the rebuild on the host, pass 2 on the GPU for the rejected samples only, merged into
 * out; *computed receives the samples the GPU computed (w h + rejected), stats the sum of
 * both passes (encode_seconds includes the rebuild).  A sample mask set on the frame does
 * not restrict the jet frame and is set again afterwards;
 * a series start set on the frame applies to both passes. */
int fdg_jetaa_first(uint32_t w, uint32_t h, uint8_t *mask);
int fdg_jetaa_rebuild(fdg_columns *columns, uint32_t w, uint32_t h, uint8_t *todo);
int fdg_jetaa_merge(fdg_columns *to, const fdg_columns *from, const uint8_t *mask, uint64_t n);
int fdg_frame_jet(fdg_frame *frame, fdg_columns *out, fdg_stats *stats, char *kernel_id,
                  size_t kernel_cap, uint64_t *computed);

/* A resident frame over a supplied reference orbit (fd_kernel::render_with / render_bla with
 * Some((orbit, bits)), crates/fd-kernel/src/grid.rs:95-109,446-476): re/im of len points
 * computed at `bits` (exactly 53 on the f64 tier, at least the view's precision on the
 * deep tiers, refused otherwise with the reference's message), starting at Z_0 = 0.  With
 * the view's own precision the frame equals one whose orbit was computed; the header
 * records `bits`.  This is how a stored orbit (below) is reused instead of recomputed. */
int fdg_frame_open_orbit(fdg_ctx *ctx, const fdg_view *view, const fdg_params *params,
                         int route, int form, double bla_eps, const double *re,
                         const double *im, uint64_t len, uint32_t bits, fdg_frame **out);

/* Canonical atlas chunks (docs/spec/ATLAS.md:127-333; crates/fd-atlas/src/{chunk,slab,
 * manifest,bla,sha256}.rs): a 32-byte header stating the numeric contract, the payload
 * little-endian with canonical floats, zero-padded to 8 bytes; a chunk's id is the SHA-256
 * of its bytes.  Host-only.  Writers return the byte count (the size needed when buf is
 * NULL or cap too small) or FDG_EINVAL for content the reference refuses; readers parse,
 * re-encode and refuse non-canonical bytes (FDG_EINVAL).
 *   orbit slab (kind 1): points start .. start+len-1 of an orbit cut into slabs of `size`
 *     (start a multiple of size, 1 <= len <= size, finite), at `bits` precision;
 *   orbit manifest (kind 9): an orbit bound to its exact decimal centre, its slabs' ids in
 *     order (32 bytes each);
 *   BLA table (kind 2): the id of the orbit manifest, eps, dc_max, orbit points and the
 *     levels 1.. of Bla::build, blocks in fdg_build_bla's layout (7 doubles each). */
int fdg_chunk_id(const uint8_t *bytes, uint64_t n, uint8_t id[32]);
int64_t fdg_chunk_orbit_slab(const double *re, const double *im, uint64_t start, uint32_t len,
                             uint32_t size, uint32_t bits, uint8_t *buf, uint64_t cap);
int fdg_chunk_orbit_slab_read(const uint8_t *bytes, uint64_t n, uint64_t *start, uint32_t *len,
                              uint32_t *size, uint32_t *bits, double *re, double *im,
                              uint64_t cap);
int64_t fdg_chunk_orbit_manifest(const char *center_re, const char *center_im, uint32_t bits,
                                 const uint8_t *slab_ids, uint64_t n, uint8_t *buf,
                                 uint64_t cap);
int fdg_chunk_orbit_manifest_read(const uint8_t *bytes, uint64_t n, char *center_re,
                                  char *center_im, uint64_t str_cap, uint32_t *bits,
                                  uint8_t *slab_ids, uint64_t ids_cap, uint64_t *count);
int64_t fdg_chunk_bla(const uint8_t orbit_id[32], double eps, double dc_max, uint64_t points,
                      uint32_t levels, const double *blocks, uint8_t *buf, uint64_t cap);
int fdg_chunk_bla_read(const uint8_t *bytes, uint64_t n, uint8_t orbit_id[32], double *eps,
                       double *dc_max, uint64_t *points, uint32_t *levels, double *blocks,
                       uint64_t cap_blocks);

/* Films: a batch of frames at one centre (the keyframes of a zoom) rendered against ONE
 * reference orbit, as the reference renders a frame with a supplied deeper orbit:
 * fd_kernel::render_with / render_bla with `Some((orbit, bits))` (crates/fd-kernel/src/
 * grid.rs:95-109,190-210; "one deep orbit serves every shallower frame at its centre",
 * and the header records the precision used).  Each frame equals that render bit for bit.
 *
 * fdg_film_open: `views` are nframes views whose centres name the same decimal number
 * (orbit.rs:228-247 `same`); `params` is shared (grid, max_iter, escape radius, columns).
 * The orbit is computed once at the deepest frame's precision and stays in HBM.  Route BLA
 * builds the tables on the GPU at open: FDG_TABLES_PER_FRAME one table per frame at that
 * frame's dc bound (as `fd control --bla per-frame`), FDG_TABLES_SHARED one table at the
 * largest dc bound over the frames (grid.rs:378-413 bla_checks accepts it for every frame).
 * An f64-tier frame is refused (its orbit must be the 53-bit one at the rounded centre,
 * grid.rs:461-465).
 * fdg_film_run: FDG_LAUNCH_BATCH renders every frame in one persistent work-stealing launch
 * per delta arithmetic (the Binary64 frames and the Binary64Exponent frames, on two
 * streams), captured once into a HIP graph and replayed; FDG_LAUNCH_PER_FRAME launches the
 * single-frame kernel once per frame, in order.  grid = persistent workgroups per launch,
 * 0 for the default.  stats aggregates the film.
 * fdg_film_read: frame k's columns / raw outcomes / stats, as fdg_frame_read. */
typedef struct fdg_film fdg_film;
enum { FDG_TABLES_PER_FRAME = 0, FDG_TABLES_SHARED = 1 };
enum { FDG_LAUNCH_BATCH = 0, FDG_LAUNCH_PER_FRAME = 1 };
int fdg_film_open(fdg_ctx *ctx, const fdg_view *views, uint32_t nframes,
                  const fdg_params *params, int route, int form, double bla_eps, int tables,
                  fdg_film **out);
int fdg_film_run(fdg_film *film, int launch, int grid, fdg_stats *stats);
int fdg_film_read(fdg_film *film, uint32_t frame, fdg_columns *out, fdg_outcomes *raw,
                  fdg_stats *stats, char *kernel_id, size_t kernel_cap);
void fdg_film_close(fdg_film *film);
/* fdg_film_mask: compute only the samples whose byte is set in the following runs — masks
 * holds nframes x nx ny bytes, frame-major (NULL clears); others are left unset as in
 * fdg_frame_mask.  fdg_film_jet: jet-AA for every frame of an ss 2 film (columns nu, de,
 * normal; outs holds nframes column sets): each pass is one batched persistent launch over
 * all frames (the graph of fdg_film_run) — pass 1 every frame's top-left samples, the host
 * rebuilds, pass 2 every frame's rejected samples — merged as fdg_frame_jet merges;
 * *computed receives the samples computed over the film.  The mask is cleared after. */
int fdg_film_mask(fdg_film *film, const uint8_t *masks);
int fdg_film_jet(fdg_film *film, fdg_columns *outs, uint64_t *computed);

/* fd_kernel::reference (grid.rs:84-87): the orbit the renderer uses and its precision in
 * bits.  With cap too small, only *len and *bits are written.  Host-only: ctx may be NULL. */
int fdg_reference(fdg_ctx *ctx, const fdg_view *view, const fdg_params *params, double *re,
                  double *im, uint64_t cap, uint64_t *len, uint32_t *bits);

/* Newton on the period-P nucleus equation z_P(c) = 0 from the start (re, im), host-only
 * (ctx may be NULL): the iteration of the reference's nucleus finder
 * (crates/fd-cli/src/zone_build.rs:215-233), with z in the reference orbit's truncated
 * fixed point at `bits` fraction bits (fd_fixed::orbit, orbit.rs:14-43) and the derivative
 * in floating point with a carried exponent.  Stops once a step is below 2^-(bits - 8)
 * and writes the nucleus as decimals with ceil(bits log10 2) + 4 fraction digits into
 * re_out / im_out (cap bytes each, NUL-terminated), the Newton steps taken and log2 of the
 * last step.  FDG_EINVAL for a bad start, an orbit that escapes, a vanishing derivative, or
 * no convergence within max_iters (iterations and log2_step are written then too). */
int fdg_nucleus(fdg_ctx *ctx, const char *re, const char *im, uint64_t period, uint32_t bits,
                uint32_t max_iters, char *re_out, char *im_out, size_t cap, uint32_t *iterations,
                double *log2_step);

/* Bla::build_scaled (crates/fd-kernel/src/bla.rs:161-181) computed on the GPU as a
 * level-by-level tree reduction.  Blocks of levels 1.. are written level-major, 7 doubles
 * each (a.re a.im b.re b.im r alpha beta, Block::to_array order, bla.rs:321-325), level j
 * holding (len-2)>>j blocks; *levels receives the stored level count. */
int fdg_build_bla(fdg_ctx *ctx, const double *re, const double *im, uint64_t len, double eps,
                  double dc_max, int64_t dc_exp, double *blocks, uint64_t cap_blocks,
                  uint32_t *levels);

/* The screen-space Jacobian of z at escape for jet-AA: d(z.re, z.im)/d(x, y) where (x, y)
 * is the sample's grid position (view.rs:117-122 unit_offset times h).  z_n(c) is
 * holomorphic, so it is the rotation-scaling matrix of the carried dz/dc composed with the
 * grid's rotation: J = [[dr, -di], [di, dr]] h [[cos, sin], [sin, -cos]].  Written per
 * sample as j[4*i..4*i+3] = J00, J01, J10, J11 times 2^jexp[i] (escaped samples only; 0
 * elsewhere).  Host-only.  No counterpart in the reference, whose jet-AA reconstructs
 * derivatives from neighbour stencils (crates/fd-cli/src/jetaa.rs:1-28). */
int fdg_jacobian(const fdg_view *view, const fdg_params *params, const fdg_outcomes *raw,
                 double *j, int64_t *jexp);

/* fd_samples::write (crates/fd-samples/src/write.rs:8-57): canonical .fds bytes of the
 * columns.  Returns the byte count (the size needed when cap is too small), or < 0. */
int64_t fdg_write_fds(const fdg_view *view, const fdg_params *params, const char *kernel_id,
                      const fdg_columns *columns, uint8_t *buf, uint64_t cap);

/* Certified error control of the reference's GPU route (host only).
 * fdg_error_radius = Reference::error_radius (crates/fd-kernel/src/reference.rs:122-148):
 *   q[m] >= |A_m - Z_m| for the stored orbit Z (len points) and the exact orbit A of the
 *   exact centre; bits < 0 for an f64 orbit from (c_re, c_im), the correctly rounded
 *   centre, else the fixed-point fraction bits it was computed with.
 * fdg_guarded_bla = GuardedBla::build (crates/fd-kernel/src/bla/guarded.rs:103-134): BLA
 *   levels 1.. whose blocks carry outward-inflated state envelopes and derivative
 *   envelopes over the orbit's radii q; written level-major, 9 doubles per block (a.re
 *   a.im b.re b.im r alpha beta dalpha dbeta), level j holding (len-2) >> j blocks;
 *   *levels receives the level count (the table ends before the first level without a
 *   valid block).  Refusals as the reference's: eps outside (0, 2^-24], dc_max not finite
 *   and positive, a non-finite or negative radius.
 * fdg_external_escape_ok = external_escape_ok (crates/fd-kernel/src/store.rs:155-198):
 *   1 when an externally computed escaped state with radii ez (z units) and ed
 *   (derivative mantissa units) is admitted for output pixel px_m 2^px_e (displacement
 *   <= 5e-4 px, DE within 1e-3, angle within 0.1 degree, escape proven), else 0.
 * fdg_overlay_escaped = overlay_escaped (crates/fd-kernel/src/grid.rs:111-171): writes the
 *   states whose `present` byte is set through the column encoder (Store::put), leaving
 *   the other samples' columns as they are; refuses Bound/Stripe columns and invalid
 *   states (n beyond max_iter, non-finite z or dz/dc, |z| within the escape radius,
 *   dz/dc = 0). */
typedef struct {
    uint64_t n;
    double zr, zi;      /* z at escape */
    double dr, di;      /* dz/dc mantissa */
    int64_t dexp;       /* dz/dc = (dr, di) 2^dexp */
} fdg_escaped_state;
int fdg_error_radius(const double *re, const double *im, uint64_t len, int64_t bits, double c_re,
                     double c_im, double *q);
int fdg_guarded_bla(const double *re, const double *im, const double *q, uint64_t len, double eps,
                    double dc_max, double *blocks, uint64_t cap_blocks, uint32_t *levels);
int fdg_external_escape_ok(const fdg_escaped_state *state, double ez, double ed, double px_m,
                           int64_t px_e);
int fdg_overlay_escaped(const fdg_view *view, const fdg_params *params, fdg_columns *out,
                        const fdg_escaped_state *states, const uint8_t *present);

/* No counterpart in the reference: an entry point for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The FP64 form of fdg_frame_open_orbit + fdg_frame_run in one synchronous
 * call: the same view, params, route, eps and supplied orbit (re, im: len host points at
 * `bits`, the same refusals), with the raw outcomes written by the kernel straight into
 * the caller's DEVICE arrays (every member of out non-NULL, nx*ny elements each).  Every
 * other buffer — the orbit's device copy, the BLA table, the counters and the kernel
 * descriptor — sits in one static device array of FDG_NOALLOC_SCRATCH_BYTES per GPU, so
 * the call allocates no device or page-locked memory; orbits longer than
 * FDG_NOALLOC_MAX_POINTS (2^17 + 2: a table of 2^17 steps) are FDG_EINVAL.  Outcomes are
 * the kernel's: an Unresolved sample's n is the iterate it stopped at (fdg_frame_read
 * reports 0, as Outcome::Unresolved carries none).  Every call on a GPU shares that GPU's
 * array and leaves its contents there: calls that overlap on one GPU (several contexts or
 * threads) race, so the caller serialises them. */
#define FDG_NOALLOC_MAX_POINTS ((1u << 17) + 2u)
#define FDG_NOALLOC_SCRATCH_BYTES (13u << 20)
int fdg_render_orbit_noalloc(fdg_ctx *ctx, const fdg_view *view, const fdg_params *params,
                             int route, double bla_eps, const double *re, const double *im,
                             uint64_t len, uint32_t bits, fdg_outcomes *out, fdg_stats *stats);

#ifdef __cplusplus
}
#endif
#endif
