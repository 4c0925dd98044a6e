/*
 * fdo.h - the CPU oracle: a plain-C restatement of fractodactyl's CPU
 * perturbation/BLA sample kernel, used ONLY as test infrastructure (parity
 * checker for tests/, __graft_entry__.smoke() and the cpu_baseline leg of
 * bench.py).  Nothing in the product links or calls it.
 *
 * Every function follows a reference function; the citation is given at the
 * definition in fdo.c.  The algorithms are restated with the reference's exact
 * floating-point operation order (built with -ffp-contract=off, no fast-math)
 * so that results are bit-identical; this is pinned against the reference's
 * frozen KERN-05 frames (crates/fd-kernel/fixtures/KERN-05.bundle) and its
 * FNV byte hashes (crates/fd-kernel/src/sample.rs:714-752).
 */
#ifndef FDO_H
#define FDO_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Column bits, crates/fd-samples/src/column.rs:9-23 */
enum { FDO_COL_CLASS = 1, FDO_COL_NU = 2, FDO_COL_DE = 4, FDO_COL_NORMAL = 8, FDO_COL_BOUND = 16 };
/* Tiers, crates/fd-kernel/src/view.rs:6-14 */
enum { FDO_TIER_AUTO = -1, FDO_TIER_F64 = 0, FDO_TIER_FIXED = 1, FDO_TIER_SCALED = 2 };
/* Outcome kinds, crates/fd-kernel/src/sample.rs:21-40 */
enum { FDO_ESCAPED = 0, FDO_INTERIOR = 1, FDO_UNRESOLVED = 2 };

typedef struct {
    double c_re, c_im, h_m;
    int64_t h_e;
    int32_t tier;
    uint64_t bits;
    double cos_r, sin_r, half_x, half_y;
} fdo_plane;

typedef struct {
    uint64_t reference_len;
    uint64_t iterations;       /* Stats::iterations (blocks count as one) */
    uint64_t equivalent;       /* sum of Outcome::iterations (skipped steps included) */
    uint64_t blocks, skipped, fallback_samples, closed_form_samples;
    uint32_t bla_levels;
    double reference_seconds, bla_seconds, render_seconds;
} fdo_stats;

/* Raw per-sample outcomes (nullable members are skipped). */
typedef struct {
    uint8_t *kind;
    uint64_t *n;
    double *zr, *zi, *dr, *di;
    int64_t *dexp;
} fdo_outcomes;

/* Encoded columns, caller-allocated nx*ny each (nullable when not requested). */
typedef struct {
    uint8_t *class_;
    double *nu;
    float *de;
    uint16_t *normal;
} fdo_columns;

double fdo_exp2i(int64_t k);

/* Plane::new. 0 ok, -1 error (message in err). */
int fdo_plane_new(const char *re, const char *im, const char *width, double rotation,
                  uint32_t nx, uint32_t ny, fdo_plane *out, char *err, size_t errcap);

/* Reference orbit for a tier (grid.rs compute). Returns the length, or -1. */
int64_t fdo_reference(const char *re, const char *im, const char *width, double rotation,
                      uint32_t nx, uint32_t ny, uint64_t max_iter, double escape_radius,
                      int tier, double *out_re, double *out_im, uint64_t cap,
                      uint32_t *bits, char *err, size_t errcap);

/* fd_fixed::orbit at `bits` fraction bits (Reference::from_fixed). Returns length. */
int64_t fdo_reference_fixed(const char *re, const char *im, uint64_t bits, uint64_t max_iter,
                            double *out_re, double *out_im, uint64_t cap);

/* Reference::nucleus_period: p, or 0 for None. */
uint64_t fdo_nucleus_period(const double *re, const double *im, uint64_t len);

/* Bla::build_scaled. Blocks of all levels, level-major, 7 doubles each
 * (a.re a.im b.re b.im r alpha beta).  Level j (1-based) has (len-2)>>j blocks.
 * Returns the number of levels, or -1 (bad eps/dc_max). */
int fdo_bla_build(const double *re, const double *im, uint64_t len, double eps, double dc_max,
                  int64_t dc_exp, double *out, uint64_t cap_blocks);

/* A whole frame (grid.rs render_with / render_bla with the table built as
 * fd control --bla per-frame does).  bla: 0 off, 1 on (eps as given).
 * kernel_id receives the header kernel string. Returns 0 or -1. */
int fdo_render(const char *re, const char *im, const char *width, double rotation,
               uint32_t nx, uint32_t ny, uint32_t ss, uint64_t max_iter, double escape_radius,
               uint32_t columns, int tier, int bla, double bla_eps, int threads,
               const uint8_t *mask, fdo_outcomes *oc, fdo_columns *cols, fdo_stats *stats,
               char *kernel_id, size_t kcap, char *err, size_t errcap);

/* fdo_render with a supplied reference orbit (grid.rs:95-109 render_with, :190-210
 * render_bla): orbit_re/orbit_im/orbit_len of precision orbit_bits (NULL: computed as
 * fdo_render does), and for bla a table built at table_dc 2^table_dc_exp (table_dc <= 0:
 * the frame's own bound).  Refusals as grid.rs:461-471 and bla_checks (:378-413). */
int fdo_render_with(const char *re, const char *im, const char *width, double rotation,
                    uint32_t nx, uint32_t ny, uint32_t ss, uint64_t max_iter,
                    double escape_radius, uint32_t columns, int tier, int bla, double bla_eps,
                    const double *orbit_re, const double *orbit_im, uint64_t orbit_len,
                    uint32_t orbit_bits, double table_dc, int64_t table_dc_exp, int threads,
                    const uint8_t *mask, fdo_outcomes *oc, fdo_columns *cols, fdo_stats *stats,
                    char *kernel_id, size_t kcap, char *err, size_t errcap);

/* One sample through the kernel (sample.rs iterate) on a given orbit, for unit tests.
 * extended selects Binary64Exponent; offset_exp is its exponent.  Returns kind. */
int fdo_sample(const double *re, const double *im, uint64_t len, const double *bla_blocks,
               int bla_levels, int have_c, double c_re, double c_im, double ar, double ai,
               int extended, int64_t offset_exp, int derivative, uint64_t max_iter, double r2,
               uint64_t nucleus, uint64_t *n_out, double *z_out, double *d_out,
               int64_t *dexp_out, uint64_t *blocks_out, uint64_t *skipped_out);
void fdo_sample_many(const double *re, const double *im, uint64_t len, const double *bla_blocks,
                     int bla_levels, double c_re, double c_im, int have_c, const double *ar,
                     const double *ai, uint64_t count, int extended, int64_t offset_exp,
                     int derivative, uint64_t max_iter, double r2, uint64_t nucleus,
                     uint8_t *kind_out, uint64_t *n_out, double *z_out, double *d_out,
                     int64_t *dexp_out);

/* fd_samples::write of the given columns.  Returns bytes written, or needed size when
 * buf is too small (call with cap 0 to size). */
uint64_t fdo_write_fds(uint8_t *buf, uint64_t cap, const char *re, const char *im,
                       const char *width, double rotation, uint32_t nx, uint32_t ny, uint32_t ss,
                       uint64_t max_iter, double escape_radius, uint32_t columns,
                       const char *kernel_id, const fdo_columns *cols);

/* series.rs:28-55 SeriesConfig (defaults: 8, 64, 1e-4, 1e-8, 1e-6). */
typedef struct {
    uint32_t start_order, max_order;
    double grow_at, shift_px, derivative_rel;
} fdo_series_config;

/* series.rs:57-82 Series; b (caller-owned, b_cap >= max_order pairs) receives b[1..b_len]
 * as re, im pairs. */
enum { FDO_STOP_REFERENCE = 0, FDO_STOP_RADIUS, FDO_STOP_NONFINITE, FDO_STOP_ESCAPE,
       FDO_STOP_BUDGET, FDO_STOP_STALL };
typedef struct {
    uint64_t n;
    double rho;
    uint64_t points;
    double *b;
    uint32_t b_cap, b_len;
    double state_bound, derivative_bound, shift_px;
    uint64_t built_steps, coefficient_pairs, growth_steps;
    int32_t stop;
} fdo_series;

/* Series::build (stalled_steps 32) / build_stalled over an orbit.  0 ok, -1 b_cap short. */
int fdo_series_build(const double *re, const double *im, uint64_t len, double rho, double pixel,
                     uint64_t max_iter, const fdo_series_config *cfg, uint64_t stalled_steps,
                     fdo_series *out);
/* Series::eval: (delta re, im, dz/dc re, im) at offset (ar, ai). */
void fdo_series_eval(const fdo_series *s, int derivative, double ar, double ai, double out[4]);
/* render_series over the frame's own orbit (fdo_render's with a series start). */
int fdo_render_series(const char *re, const char *im, const char *width, double rotation,
                      uint32_t nx, uint32_t ny, uint32_t ss, uint64_t max_iter,
                      double escape_radius, uint32_t columns, int tier, int bla, double bla_eps,
                      const fdo_series *series, int threads, const uint8_t *mask,
                      fdo_outcomes *oc, fdo_columns *cols, fdo_stats *stats, char *kernel_id,
                      size_t kcap, char *err, size_t errcap);

/* jetaa.rs first / rebuild / merge on an ss 2 grid of w x h pixels. */
void fdo_jetaa_first(uint64_t w, uint64_t h, uint8_t *mask);
void fdo_jetaa_rebuild(uint8_t *cls, double *nu, float *de, uint16_t *normal, uint64_t w,
                       uint64_t h, uint8_t *todo);
void fdo_jetaa_merge(uint8_t *cls, double *nu, float *de, uint16_t *normal, const uint8_t *fcls,
                     const double *fnu, const float *fde, const uint16_t *fnormal,
                     const uint8_t *mask, uint64_t n);

int fdo_bla_find_many(const double *re, const double *im, uint64_t len, double eps, double dc_max,
                      int64_t dc_exp, const uint64_t *ms, uint64_t nm, double d2, uint64_t budget,
                      int32_t *out_level);
int fdo_fixed_parse(const char *s, uint64_t limbs, uint64_t *out);

#ifdef __cplusplus
}
#endif
/* certified error control of the GPU route: reference.rs error_radius, bla/guarded.rs
 * GuardedBla::build, store.rs external_escape_ok */
void fdo_error_radius(const double *re, const double *im, uint64_t len, int64_t bits,
                      double c_re, double c_im, double *q);
int fdo_guarded_build(const double *re, const double *im, const double *q, uint64_t len,
                      double eps, double dc_max, double *out, uint64_t cap);
int fdo_external_escape_ok(uint64_t n, double zr, double zi, double dr, double di, int64_t dexp,
                           double ez, double ed, double px_m, int64_t px_e);

/* direct iteration of z^2 + c at `bits` bits with dz/dc (test hook for the direct fixture) */
int64_t fdo_direct(const char *re, const char *im, uint64_t bits, uint64_t max_iter, double r2,
                   double *out, int64_t *de);

/* sample.rs:1119-1122 zero_step on explicit operands (test hook) */
void fdo_zero_step(double wr, double wi, int64_t e, double ar, double ai, int64_t ae,
                   double *out, int64_t *oe);

#endif
