/*
 * s1ifg.h — C ABI of the MI355X interferogram-formation library (libs1ifg.so).
 *
 * The reference (WWSAR/sentinel_gpu_processor) has no in-process interface
 * for this path: Python launches standalone executables and exchanges argv,
 * stdin/stdout, files and an SQLite parameter table with them.  Each entry
 * point below replaces one stage of those executables; the comment on each
 * names the reference code it replaces (file:line in the reference tree).
 *
 * Conventions
 *  - Every pointer to pixel data is a DEVICE pointer (hipMalloc / torch
 *    tensor storage) unless the comment says host.  Small parameter blocks
 *    (orbit, burst metadata, grids, polynomials) are HOST pointers.
 *  - The caller allocates every input and output.  A context owns only its
 *    stream binding and a workspace it grows on demand.
 *  - Return value: 0 on success, a negative S1IFG_E* code on failure; the
 *    message is available from s1ifg_last_error() on the calling thread.
 *    Nothing in the library exits or aborts the process.
 *  - Every entry point enqueues work on the context stream and returns
 *    without synchronising; s1ifg_ctx_sync() waits.
 *  - Complex data are interleaved float pairs (re, im), row-major.
 *  - One context per host thread (or per stream); contexts are independent.
 */
#ifndef S1IFG_H
#define S1IFG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define S1IFG_ABI_VERSION 1

enum {
  S1IFG_OK = 0,
  S1IFG_EINVAL = -1, /* bad argument (sizes, null pointers, metadata) */
  S1IFG_EHIP = -2,   /* HIP runtime error */
  S1IFG_ENOMEM = -3, /* device allocation failed */
  S1IFG_ERANGE = -4, /* size beyond what the kernels support */
  S1IFG_ESTATE = -5  /* object used with the wrong context / device */
};

enum { S1IFG_INTERP_BILINEAR = 0, S1IFG_INTERP_SINC = 1 };
enum { S1IFG_GOLD_FP32 = 0, S1IFG_GOLD_FP16_SPLIT = 1 };
enum { S1IFG_DEM_F32 = 0, S1IFG_DEM_I16 = 1 };
/* coherence flags */
enum {
  S1IFG_COH_FILL_CLAMP = 0, /* gamma > 1 -> 1 */
  S1IFG_COH_FILL_NAN = 1,   /* gamma > 1 -> NaN (mask) */
  S1IFG_COH_FILL_HASH = 2   /* gamma > 1 -> seeded uniform [0,1) */
};

typedef struct { float re, im; } s1ifg_cf32;

typedef struct s1ifg_ctx s1ifg_ctx;
typedef struct s1ifg_geom s1ifg_geom;

/* Orbit state vectors, uniformly spaced in time (HOST pointers).
 * Replaces the orbit text file read by read_orbit_ascii (orbit.cu:147-168). */
typedef struct {
  int32_t n;          /* number of state vectors, >= 4 */
  int32_t pad_;
  const double *t;    /* [n] seconds */
  const double *pos;  /* [n*3] ECEF metres */
  const double *vel;  /* [n*3] ECEF metres/second */
} s1ifg_orbit;

/* Burst metadata.  Replaces the SQLite `file` table keys read at
 * geo2rdr.cu:333-376 plus the FM-rate / Doppler-centroid files read by
 * read_polynomials (sario.cu:80-103).  fm/dc are HOST arrays of records
 * (t, t0, c0, c1, c2). */
typedef struct {
  int32_t lines;               /* linesPerBurst */
  int32_t samples;             /* samplesPerBurst */
  int32_t first_valid_line;    /* firstValidLine{k} */
  int32_t last_valid_line;     /* lastValidLine{k} */
  int32_t first_valid_sample;  /* firstValidSample{k} */
  int32_t last_valid_sample;   /* lastValidSample{k} */
  double t_start;              /* azimuthTimeSeconds{k} */
  double prf;                  /* prf (reproject uses 1/prf) */
  double azimuth_time_interval;/* azimuthTimeInterval (deramp eta) */
  double slant_range_time;     /* slantRangeTime, two-way seconds */
  double range_sampling_rate;  /* rangeSamplingRate, Hz */
  double wavelength;           /* wvl, metres */
  double radar_frequency;      /* radarFrequency, Hz */
  double azimuth_steering_rate;/* azimuthSteeringRate, deg/s */
  int32_t n_fm;
  int32_t n_dc;
  const double *fm;            /* [n_fm*5] */
  const double *dc;            /* [n_dc*5] */
} s1ifg_burst;

/* Target lat/lon grid (the .rsc keys Y_FIRST, X_FIRST, Y_STEP, X_STEP,
 * FILE_LENGTH, WIDTH; sario.hpp:399-408).  Pixel (r, c) is at
 * (lat0 + r*dlat, lon0 + c*dlon). */
typedef struct {
  int32_t nrow, ncol;
  double lat0, lon0, dlat, dlon;
} s1ifg_grid;

/* DEM on its own lat/lon grid (DEVICE data); heights are bilinearly
 * interpolated at target pixels, with indices clamped to the DEM.  When the
 * DEM grid equals the target grid this is the reference's h = dem[i]
 * (geo2rdr.cu:199). */
typedef struct {
  const void *data;
  int32_t dtype;  /* S1IFG_DEM_F32 or S1IFG_DEM_I16 */
  int32_t nrow, ncol;
  int32_t pad_;
  double lat0, lon0, dlat, dlon;
} s1ifg_dem;

/* ---------------------------------------------------------------- context */
int s1ifg_abi_version(void);
/* Thread-local message of the last failing call on this thread. */
const char *s1ifg_last_error(void);
/* Replaces set_gpu / parse_gpu_arg (gpu_device.hpp:22-79). */
int s1ifg_ctx_create(int device, s1ifg_ctx **out);
void s1ifg_ctx_destroy(s1ifg_ctx *ctx);
/* Bind a HIP stream (hipStream_t as void*); NULL = the null stream. */
int s1ifg_ctx_set_stream(s1ifg_ctx *ctx, void *stream);
int s1ifg_ctx_sync(s1ifg_ctx *ctx);
/* Validation hook (no reference counterpart): a DEVICE array of
 * S1IFG_TILE_STATS uint64 counters to which later s1ifg_pair / s1ifg_pair_raw
 * calls with sinc interpolation on this context add the classes of their
 * 16 x 16-cell tiles: [0] tiles, [1] burst-tiles skipped by the tile's
 * corner test, [2] burst-tiles interpolated from the LDS-staged tap window,
 * [3] burst-tiles whose window exceeds the LDS budget (taps read from
 * global memory), [4] burst-tiles not skipped without any valid pixel, [5]
 * burst-tiles of s1ifg_pair_raw interpolated from a sheared LDS window (the
 * rectangular one exceeds the budget, the sheared one fits).  NULL turns
 * counting off (the default).  The caller zeroes the array. */
#define S1IFG_TILE_STATS 6
int s1ifg_ctx_set_tile_stats(s1ifg_ctx *ctx, uint64_t *counts);
/* Workspace currently held by the context, in bytes. */
int s1ifg_ctx_workspace(s1ifg_ctx *ctx, int64_t *bytes);

/* ----------------------------------------------------- burst geometry */
/* Host FP64 set-up of one burst: orbit interpolation polynomials, ks, FM/DC
 * coefficients at the burst centre, kt/etaref per range sample, the burst
 * mid-time orbit state.  Replaces geo2rdr.cu:386-400 and 465-549 and the
 * device copies at 444-461, 545-549, 596-600. */
int s1ifg_geom_create(s1ifg_ctx *ctx, const s1ifg_orbit *orbit,
                      const s1ifg_burst *burst, s1ifg_geom **out);
void s1ifg_geom_destroy(s1ifg_geom *g);
/* Scalars of the set-up for inspection: out[0]=ks, out[1..4]=fm (t0,c0,c1,c2),
 * out[5..8]=dc (t0,c0,c1,c2), out[9]=etac0, out[10]=tmid, out[11..13]=xmid,
 * out[14..16]=vmid.  n >= 17. */
int s1ifg_geom_query(const s1ifg_geom *g, double *out, int n);
/* The per-range-sample deramp vectors of the set-up, kt[j] and etaref[j]
 * (geo2rdr.cu:533-542), as the kernels read them, into HOST arrays of n =
 * samples doubles (either may be NULL).  Synchronous (validation only). */
int s1ifg_geom_vectors(const s1ifg_geom *g, double *kt, double *etaref, int n);
/* Burst footprint (latmin, latmax, lonmin, lonmax) for heights in
 * [hmin, hmax]; replaces bounds()/rah2ll (bounds.cu:237-317). HOST output. */
int s1ifg_geom_bounds(const s1ifg_geom *g, double hmin, double hmax,
                      double latlon_out[4]);
/* Orbit position/velocity/acceleration at n times (DEVICE arrays):
 * replaces intp_orbit (orbit.cu:224-245) for validation. */
int s1ifg_orbit_eval(s1ifg_ctx *ctx, const s1ifg_geom *g, const double *t,
                     int64_t n, double *pos, double *vel, double *acc);

/* --------------------------------------------------------- burst stages */
/* A8 deramp: out = in * exp(-i*pi*kt[j]*(eta[i]-etaref[j])^2); in may equal
 * out.  Replaces deramp_burst (geo2rdr.cu:138-158). */
int s1ifg_deramp(s1ifg_ctx *ctx, const s1ifg_geom *g, const s1ifg_cf32 *in,
                 s1ifg_cf32 *out);
/* A8 deramp of a burst in the Sentinel-1 SLC measurement format: in is
 * interleaved CInt16 (re, im int16 per sample, lines x samples, DEVICE),
 * out complex float (must not overlap in).  The int16 -> float conversion
 * is exact, so out equals s1ifg_deramp of the converted burst bit for bit;
 * it reads 4 bytes per sample instead of 8.  Replaces the CInt16 -> cf32
 * conversion of the TIFF stream (sentinel_scene.py:164-207, GDAL
 * ReadAsArray) plus deramp_burst on the streamed burst (geo2rdr.cu:552-564,
 * 138-158). */
int s1ifg_deramp_ci16(s1ifg_ctx *ctx, const s1ifg_geom *g, const int16_t *in,
                      s1ifg_cf32 *out);

/* A5/A9 geo2rdr offsets on the target grid (validation mode: the fused
 * kernels never materialise them).  dh: optional per-pixel DEM error (float,
 * DEVICE) added to the DEM height.  offpoly: optional HOST double[6] offset
 * polynomial (see s1ifg_geocode).  rgoff/azoff/rng: DEVICE double
 * [nrow*ncol]; any may be NULL.  Replaces the Newton part of reproject
 * (geo2rdr.cu:195-207; orbitrangetime orbit.cu:247-279). */
int s1ifg_geo2rdr(s1ifg_ctx *ctx, const s1ifg_geom *g, const s1ifg_grid *grid,
                  const s1ifg_dem *dem, const float *dh, const double *offpoly,
                  double *rgoff, double *azoff, double *rng);

/* A9 geocode one deramped burst onto the target grid: geo2rdr, validity
 * window, interpolation (bilinear as the reference, or 8-tap windowed sinc),
 * reramp and range flattening; invalid pixels are 0.  offpoly (HOST double[6],
 * nullable) shifts the sampling position by
 *   d_rg = p0 + p1*(rgoff/samples) + p2*(azoff/lines)
 *   d_az = p3 + p4*(rgoff/samples) + p5*(azoff/lines)  pixels.
 * Replaces reproject (geo2rdr.cu:170-248). */
int s1ifg_geocode(s1ifg_ctx *ctx, const s1ifg_geom *g, const s1ifg_grid *grid,
                  const s1ifg_dem *dem, const float *dh, const double *offpoly,
                  const s1ifg_cf32 *deramped, int interp, s1ifg_cf32 *out);

/* A10 first and last-exclusive non-zero rows of an image into DEVICE
 * int32[2] ({0,0} when all zero).  Replaces
 * find_first_last_nonzero_rows_kernel + find_nonzero_rows
 * (geo2rdr.cu:86-128, 273-299). */
int s1ifg_valid_rows(s1ifg_ctx *ctx, const s1ifg_cf32 *img, int nrow, int ncol,
                     int32_t *first_last);

/* ------------------------------------------------------ pair products */
/* A12/A13/A16 fused crossmul + look means: ifg_ml = mean over rl x cl of
 * conj(ref)*sec; pow_ref/pow_sec (nullable) = mean |ref|^2, |sec|^2 on the
 * same cells.  Strides in elements.  Output is (nrow/rl) x (ncol/cl) with row
 * stride ncol/cl.  Replaces conj_mul, cpx_col_look, cpx_row_look
 * (sario.cu:199-286) as launched by launch_async_pipeline
 * (crossmul_daemon.cu:921-975). */
int s1ifg_crossmul_looks(s1ifg_ctx *ctx, const s1ifg_cf32 *ref, int64_t ld_ref,
                         const s1ifg_cf32 *sec, int64_t ld_sec, int nrow,
                         int ncol, int rl, int cl, s1ifg_cf32 *ifg_ml,
                         float *pow_ref, float *pow_sec);

/* A13 complex look means of one image (column mean over cl, then row mean
 * over rl) into a (nrow/rl) x (ncol/cl) output.  Replaces cpx_col_look /
 * cpx_row_look (sario.cu:241-286) and cpxlooks (sario.py:397-413). */
int s1ifg_looks(s1ifg_ctx *ctx, const s1ifg_cf32 *in, int64_t ld, int nrow, int ncol,
                int rl, int cl, s1ifg_cf32 *out);

/* A13 phase of n complex values: out = atan2(im, re).  Replaces point_angle
 * (sario.cu:231-239) used by the daemon's --out-float mode. */
int s1ifg_phase(s1ifg_ctx *ctx, const s1ifg_cf32 *in, int64_t n, float *out);

/* A16 amplitude looks of one image: amp = sqrt(mean over rl x cl of |s|^2);
 * at rl = cl = 1 the reference returns sqrt(|s|) (sario.py:416-418) and so
 * does this call.  Replaces multilook/powlooks (coherence.py:14-55,
 * sario.py:397-420). */
int s1ifg_amplitude(s1ifg_ctx *ctx, const s1ifg_cf32 *slc, int64_t ld,
                    int nrow, int ncol, int rl, int cl, float *amp);

/* A17 coherence on the multilooked grid.  win = 1: the reference look-cell
 * estimator |ifg|/(amp1+1e-12)/(amp2+1e-12) with amp = sqrt(pow)
 * (coherence.py:142-144; the reference's unseeded random fill for gamma > 1
 * is replaced by the S1IFG_COH_FILL_* rule).  win odd >= 3: windowed
 * estimator |sum ifg| / sqrt(sum pow_ref * sum pow_sec) over win x win,
 * clipped at the borders (no reference counterpart). */
int s1ifg_coherence(s1ifg_ctx *ctx, const s1ifg_cf32 *ifg_ml,
                    const float *pow_ref, const float *pow_sec, int nrow,
                    int ncol, int win, int fill, uint32_t seed, float *coh);

/* A18 Goldstein adaptive filter (patch n_win with step n_win/4, alpha), with
 * a deterministic gather-form overlap-add.  mode: S1IFG_GOLD_FP32 (radix-2
 * FFT in FP32) or S1IFG_GOLD_FP16_SPLIT (DFT-by-GEMM on MFMA, hi/lo split
 * FP16 operands, FP32 accumulation; patches whose components are all +0 or
 * -0 filter to zeros and are passed over).  Pixels whose weight sum is 0
 * are 0.  n_win must be 32; nrow, ncol >= n_win.  Replaces
 * goldstein_filter_cuda_host and its kernels (goldstein.cu:87-408). */
int s1ifg_goldstein(s1ifg_ctx *ctx, const s1ifg_cf32 *in, int nrow, int ncol,
                    int n_win, float alpha, int mode, s1ifg_cf32 *out);

/* Fused burst pair: geo2rdr for both orbits, interpolation of both deramped
 * bursts, reramp + flattening, crossmul and look means, in one pass over the
 * target grid (the geocoded SLCs are never written).  offpoly applies to the
 * secondary.  Outputs as s1ifg_crossmul_looks over the whole target grid.
 * Replaces the geo2rdr -> .gslc -> crossmul_daemon chain
 * (geo2rdr.cu:170-248, crossmul_daemon.cu:921-975). */
int s1ifg_pair(s1ifg_ctx *ctx, const s1ifg_geom *g_ref, const s1ifg_geom *g_sec,
               const s1ifg_grid *grid, const s1ifg_dem *dem, const float *dh,
               const double *offpoly_sec, const s1ifg_cf32 *ref_deramped,
               const s1ifg_cf32 *sec_deramped, int interp, int rl, int cl,
               s1ifg_cf32 *ifg_ml, float *pow_ref, float *pow_sec);

/* s1ifg_pair on the RAW bursts: each burst sample is deramped (the phase of
 * deramp_burst, geo2rdr.cu:138-158, in FP64, with the hardware sine and
 * cosine) as the kernel reads it, so no deramped copy is written or read
 * back.  Replaces the deramp_burst -> reproject -> crossmul chain of
 * geo2rdr.cu:138-248 and crossmul_daemon.cu:921-975 for one pair in one
 * pass.  Same arguments and outputs as s1ifg_pair, the bursts raw. */
int s1ifg_pair_raw(s1ifg_ctx *ctx, const s1ifg_geom *g_ref, const s1ifg_geom *g_sec,
                   const s1ifg_grid *grid, const s1ifg_dem *dem, const float *dh,
                   const double *offpoly_sec, const s1ifg_cf32 *ref_slc, const s1ifg_cf32 *sec_slc,
                   int interp, int rl, int cl, s1ifg_cf32 *ifg_ml, float *pow_ref,
                   float *pow_sec);

/* s1ifg_pair without any device allocation: the same kernels and results,
 * with the grid tables (the call's only scratch memory) in a fixed 2 MiB
 * device array of the library instead of the context workspace.  No
 * counterpart in the reference: it is for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Limit: 32 * nrow + 24 * ncol bytes (each rounded up to 256)
 * <= 2 MiB, e.g. 16384 x 65536; larger grids fail with S1IFG_ERANGE.  Every
 * call on a device shares that device's array: calls that overlap on one
 * device race (serialise them, e.g. on one stream). */
int s1ifg_pair_noalloc(s1ifg_ctx *ctx, const s1ifg_geom *g_ref, const s1ifg_geom *g_sec,
                       const s1ifg_grid *grid, const s1ifg_dem *dem, const float *dh,
                       const double *offpoly_sec, const s1ifg_cf32 *ref_deramped,
                       const s1ifg_cf32 *sec_deramped, int interp, int rl, int cl,
                       s1ifg_cf32 *ifg_ml, float *pow_ref, float *pow_sec);

/* ------------------------------------------------------------ adjoints */
/* Adjoint of s1ifg_pair (no reference counterpart).  Given the gradients of
 * a real loss with respect to the pair's outputs on the multilooked grid --
 * g_ifg = dL/dRe ifg + i dL/dIm ifg, g_pow_ref, g_pow_sec (DEVICE; any may
 * be NULL = 0) -- writes dL/d offpoly_sec (DEVICE double[6], may be NULL)
 * and dL/d dh per target pixel (DEVICE float nrow*ncol, may be NULL; 0
 * outside the look cells).  The forward is recomputed (geometry, sinc taps
 * and their derivatives), nothing is stored.  Same arguments as s1ifg_pair
 * for the forward part; offpoly_sec NULL = no offset polynomial (its
 * gradient is still returned, at p = 0, if g_offpoly is given). */
int s1ifg_pair_backward(s1ifg_ctx *ctx, const s1ifg_geom *g_ref, const s1ifg_geom *g_sec,
                        const s1ifg_grid *grid, const s1ifg_dem *dem, const float *dh,
                        const double *offpoly_sec, const s1ifg_cf32 *ref_deramped,
                        const s1ifg_cf32 *sec_deramped, int interp, int rl, int cl,
                        const s1ifg_cf32 *g_ifg, const float *g_pow_ref,
                        const float *g_pow_sec, double *g_offpoly, float *g_dh);

/* Adjoint of s1ifg_coherence with fill CLAMP / NAN / HASH (no gradient
 * where the coherence was filled or is 0).  win = 1 needs all three
 * outputs; win >= 3 writes the non-NULL ones. */
int s1ifg_coherence_backward(s1ifg_ctx *ctx, const s1ifg_cf32 *ifg_ml,
                             const float *pow_ref, const float *pow_sec, int nrow,
                             int ncol, int win, const float *g_coh,
                             s1ifg_cf32 *g_ifg, float *g_pow_ref, float *g_pow_sec);

/* ------------------------------------------------------------ batches */
/* One burst pair of a batch (all pointers DEVICE except the geometries).
 * The raw bursts are deramped into ref_deramped / sec_deramped, which must
 * not alias the raw bursts (a batch may be run many times).  A NULL raw
 * burst means its deramped buffer is an input.  Pairs may share bursts (a
 * stack of sequential pairs): a burst given by the same raw / deramped
 * buffers in several tasks is deramped once.  The products are on the
 * multilooked grid (grid.nrow/rl x grid.ncol/cl); filt and coh may be NULL
 * to skip the Goldstein filter / coherence of that pair. */
typedef struct {
  const s1ifg_geom *g_ref, *g_sec;
  s1ifg_grid grid;
  s1ifg_dem dem;
  const float *dh;           /* per target pixel height error, or NULL */
  double offpoly_sec[6];     /* secondary offset polynomial */
  int32_t use_offpoly;
  int32_t pad_;
  const s1ifg_cf32 *ref_slc, *sec_slc;
  s1ifg_cf32 *ref_deramped, *sec_deramped;
  s1ifg_cf32 *ifg_ml;
  float *pow_ref, *pow_sec;
  s1ifg_cf32 *filt;
  float *coh;
} s1ifg_pair_task;

/* Processing parameters shared by the pairs of a batch, and the launch
 * shape (0 = built-in default; an autotuner measures and picks them). */
typedef struct {
  int32_t interp, rl, cl;
  float alpha;
  int32_t gold_mode, coh_win, coh_fill;
  uint32_t coh_seed;
  int32_t pair_block_x;      /* pair tile width in multilooked pixels: 16..256 */
  int32_t pair_blocks_per_cu;/* persistent grid of the pair kernel */
  int32_t gold_blocks_per_cu;/* persistent grid of the Goldstein patch pass */
  int32_t stream_blocks_per_cu; /* deramp / overlap-add / coherence */
  int32_t raw;               /* != 0: no deramp pass; the pair kernel reads
                              * ref_slc / sec_slc and deramps the samples as
                              * it stages them (as s1ifg_pair_raw; sinc
                              * interpolation with rl * cl <= 4 only), and
                              * ref_deramped / sec_deramped may be NULL */
} s1ifg_batch_opts;

/* A *_blocks_per_cu value: one block per work item (a grid as large as the
 * work) instead of persistent blocks; 0 selects the default. */
#define S1IFG_BLOCK_PER_ITEM (-1)

typedef struct s1ifg_batch s1ifg_batch;

/* All bursts of an IW pair (or of several pairs of a stack) in one launch
 * sequence of persistent kernels: deramp of every burst, the fused pair
 * kernel, the Goldstein patch pass, its overlap-add, coherence.  Each pair's
 * products are bit-identical to s1ifg_deramp x2 + s1ifg_pair +
 * s1ifg_goldstein + s1ifg_coherence on that pair alone.  The batch owns its
 * scratch memory (tables, descriptors, Goldstein patches).  Replaces the
 * per-burst loops of geo2rdr (geo2rdr.cu:552-664) and the crossmul daemon's
 * lanes (crossmul_daemon.cu:921-1117) for the same pairs. */
int s1ifg_batch_create(s1ifg_ctx *ctx, const s1ifg_pair_task *tasks, int n,
                       const s1ifg_batch_opts *opts, s1ifg_batch **out);
void s1ifg_batch_destroy(s1ifg_batch *b);
/* Enqueue the batch on the context stream.  use_graph != 0: the launch
 * sequence is captured once into a HIP graph (on the batch's own stream)
 * and replayed as one graph launch afterwards. */
int s1ifg_batch_run(s1ifg_batch *b, int use_graph);
/* Device bytes held by the batch. */
int s1ifg_batch_bytes(const s1ifg_batch *b, int64_t *bytes);
/* The launch shape in use (after defaults are filled in). */
int s1ifg_batch_get_opts(const s1ifg_batch *b, s1ifg_batch_opts *opts);
/* Change the launch shape (the pair_block_x / *_blocks_per_cu fields of
 * opts; the processing parameters are kept).  Waits for the batch's pending
 * work; the next graph run captures a new graph.  Results do not depend on
 * the launch shape. */
int s1ifg_batch_set_launch(s1ifg_batch *b, const s1ifg_batch_opts *opts);

/* ---- Stitching (interfere.py:81-180) ---- */

/* One burst interferogram of a subswath: device data of (bottom - top) rows
 * x (right - left) columns, row-major, cf32 (or f32 phase when out_float),
 * with its bounds on the full image grid. */
typedef struct {
  const void *data;
  int32_t left, top, right, bottom;
  int32_t estimate;  /* 1: this burst's ids differ from the previous burst's
                        (interfere.py:91-95, 112-117): estimate the offset */
  int32_t pad_;
} s1ifg_stitch_burst;

/* Mosaic of one subswath's bursts, already sorted by top + bottom
 * (sario.py:240-245), into `mosaic` ((bottom - top) x (right - left) of the
 * given subswath bounds, the first / last bursts' box, sario.py:230-237).
 * Burst 0 is resampled into the bounds; each later burst, which must lie
 * inside them, gets the running phase offset (angle of the mean phase
 * difference over its overlap with the mosaic, plus the median of the
 * residual angles) and fills the pixels still without data
 * (interfere.py:97-143).  offsets: device array of n doubles receiving the
 * offset applied to each burst (0 for burst 0), or NULL.  Replaces
 * stitch_subswath (interfere.py:81-150) for in-memory bursts. */
int s1ifg_stitch_subswath(s1ifg_ctx *ctx, const s1ifg_stitch_burst *bursts, int n,
                          int out_float, void *mosaic, int left, int top, int right,
                          int bottom, double *offsets);
/* dst[r, c] = src[r, c] where src has data (real part, or value, non-zero):
 * one subswath mosaic into the full image at dst (row stride ld_dst
 * elements).  Replaces the merge loop of stitch (interfere.py:170-179). */
int s1ifg_merge_subswath(s1ifg_ctx *ctx, const void *src, int nrow, int ncol, int out_float,
                         void *dst, int64_t ld_dst);

/* ---- PS phase similarity (phase_similarity.cu) ----
 * Stacks are nifg interferograms of nrow x ncol complex floats, image after
 * image ([nifg][nrow][ncol], DEVICE); candidate masks are int32 [nrow][ncol]
 * (non-zero = candidate), DEVICE.  Neighbours of a pixel are explored in the
 * reference's order (s1ifg_ps_scan_offsets) and the first nneigh candidates
 * inside the image are used; similarities are float sums over the
 * interferograms in order, divided by nifg.  1 <= nneigh <= 128 (0 is taken
 * as 1, as the reference's loop does), rdmax <= 2048. */

/* The neighbour offsets (row, column) in exploration order: rings
 * rdmin+1 .. rdmax-1 of the reference's gap-filled midpoint circle.
 * Writes min(count, cap) pairs to the HOST array offsets (may be NULL with
 * cap 0) and returns count (>= 0), or a negative S1IFG_E* code.  Replaces
 * scan_array (phase_similarity.cu:295-366). */
int s1ifg_ps_scan_offsets(int rdmin, int rdmax, int32_t *offsets, int cap);
/* Unit phasors: z / |z|, 0 where |z| <= 1e-8; n complex elements, in may
 * equal out.  Replaces normalize (phase_similarity.cu:116-134). */
int s1ifg_ps_normalize(s1ifg_ctx *ctx, const s1ifg_cf32 *in, s1ifg_cf32 *out, int64_t n);
/* ps[k] = 1 where sim[k] >= threshold, else 0 when overwrite != 0 (else
 * unchanged); n elements.  Replaces update_ps (phase_similarity.cu:145-157). */
int s1ifg_ps_update(s1ifg_ctx *ctx, int32_t *ps, const float *sim, float threshold,
                    int overwrite, int64_t n);
/* Median similarity of each candidate to its first nneigh candidate
 * neighbours (element count/2 of the sorted similarities; 0 without
 * neighbours); 0 for pixels that are not candidates.  stack_norm: an
 * already normalised stack.  Replaces median_similarity<N>
 * (phase_similarity.cu:174-233). */
int s1ifg_ps_median_similarity(s1ifg_ctx *ctx, const s1ifg_cf32 *stack_norm, int nifg,
                               int nrow, int ncol, const int32_t *ps, int nneigh, int rdmin,
                               int rdmax, float *sim_median);
/* 1 for candidates; for the other pixels the largest similarity to their
 * first nneigh candidate neighbours (0 without any).  Replaces
 * maximum_similarity (phase_similarity.cu:235-283). */
int s1ifg_ps_maximum_similarity(s1ifg_ctx *ctx, const s1ifg_cf32 *stack_norm, int nifg,
                                int nrow, int ncol, const int32_t *ps, int nneigh, int rdmin,
                                int rdmax, float *sim_max);
/* The executable's whole sequence on one resident stack: normalise,
 * median similarity of the candidates, ps <- (median >= med_sim_th)
 * (updated in place), maximum similarity with the updated candidates.
 * stack is read only (the normalised copy lives in the workspace).
 * Replaces similarity() (phase_similarity.cu:375-548) with the whole image
 * resident (no row batches). */
int s1ifg_ps_similarity(s1ifg_ctx *ctx, const s1ifg_cf32 *stack, int nifg, int nrow,
                        int ncol, int32_t *ps, int nneigh, int rdmin, int rdmax,
                        float med_sim_th, float *med_sim, float *max_sim);

#ifdef __cplusplus
}
#endif
#endif /* S1IFG_H */
