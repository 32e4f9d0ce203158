/*
 * gisspa.h -- C ABI of the MI355X template-matching search library (libgisspa.so).
 *
 * The reference has no plugin or FFI: its boundary is the C++ engine pair
 *   SearchNorm   (include/norm.cuh:12-55,   src/norm.cu:48-123, 620-660)  norm_type = 1
 *   SearchNoNorm (include/nonorm.cuh:12-49, src/nonorm.cu:50-122, 559-589) norm_type = 0
 * constructed per micrograph by main (src/main.cu:42-48) from a Config, the EulerData and a
 * device id, and driven by work(Templates, Image|TileImages, ostream).  This header exposes the
 * same capabilities as plain C entry points: an engine bound to one device and one host
 * thread, template loading, a per-micrograph search, and the records the reference prints.
 *
 * Conventions: every pointer argument is caller-owned host memory and is never modified;
 * arrays returned by the library are freed with gs_free.  Functions return a gs_status and
 * never exit or throw; gs_last_error() describes the last failure on the calling thread.
 * Separate engines may be used concurrently from separate threads.
 */
#ifndef GISSPA_H
#define GISSPA_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  GS_OK = 0,
  GS_ERR_INVALID = 1,     /* bad argument or configuration */
  GS_ERR_HIP = 2,         /* a HIP runtime call failed */
  GS_ERR_UNSUPPORTED = 3, /* valid request the library does not handle (e.g. an unsupported size) */
  GS_ERR_NOMEM = 4,       /* device or host allocation failed */
  GS_ERR_STATE = 5        /* call out of order (e.g. search before templates) */
} gs_status;

/* Search parameters.  Each field is the reference config key of the same meaning
 * (include/EMReader/DataReader2.h:17-44 keys, :73-83 Parameters):
 *   angpix, phistep, kk, energy, cs, Highres, Lowres, diameter, threshold,
 *   window_size, overlap, norm_type, invert, phase_flip.
 * nms_radius is new: radius (px) of the greedy non-maximum suppression applied to the
 * records before they are returned (0 returns every record above threshold).
 * window_size: any even size, as in the reference (norm_type=0: at most 720).  320 and 720 run
 * the tiled correlation kernels; other sizes run a generic correlation on runtime-planned
 * transforms (same results to float rounding, slower), for which the kernel variants,
 * gs_autotune and the two-stream pipeline have no effect.
 * reference_records is new (0 default): with norm_type=1, 1 makes the records of a run the
 * reference's own output (src/norm.cu:560-600): one per (angle, block, template, 256-element
 * chunk of the block's correlation index y*l + x) whose largest z reaches the threshold, the
 * chunk maximum chosen by get_peak_pos's tree (src/kernels.cu:766-793), centre in
 * [d_m, W - d_m) x [d_m, H - d_m), in the reference's order and without NMS.  It costs a second
 * correlation pass (on the generic correlation) and needs window_size^2 divisible by 256. */
typedef struct {
  float angpix, phistep, kk, energy, cs, highres, lowres, diameter, threshold;
  int32_t window_size, overlap, norm_type, invert, phase_flip;
  float nms_radius;
  int32_t reference_records;
} gs_params;

/* Per-micrograph CTF of an LST line (src/EMReader/DataReader2.cpp:80-101):
 * defocus and dfdiff in micrometres, dfang in degrees. */
typedef struct {
  double defocus, dfdiff, dfang;
} gs_ctf;

/* One output record: the fields the reference prints per line (src/norm.cu:610-615,
 * src/nonorm.cu:548-553): centre, template index J (Euler angles alt[J], az[J]), in-plane angle
 * psi (norm: 360-theta; nonorm: theta) and score; plus bookkeeping (angle loop index, block or
 * tile, in-block pixel) and, for norm_type=0, the tile-local centre the reference prints. */
typedef struct {
  float x, y;
  int32_t template_id;
  float psi;
  float score;
  int32_t angle_index;
  int32_t block;
  int32_t pixel;
  float local_x, local_y;
} gs_record;

typedef struct gs_engine gs_engine;

/* Replaces SearchNorm/SearchNoNorm construction (src/norm.cu:48-123, src/nonorm.cu:50-122):
 * binds an engine to `device` and validates the parameters. */
gs_status gs_create(const gs_params* params, int device, gs_engine** out);

/* Replaces Templates + LoadTemplate (src/utils/templates.cpp:11-31, src/norm.cu:132-154):
 * n templates of w x h float32, row-major, stacked; alt_az holds n (alt, az) pairs from the
 * Euler file (src/EMReader/DataReader2.cpp:55-78) and may be NULL. */
gs_status gs_load_templates(gs_engine* e, const float* data, int n, int w, int h,
                            const float* alt_az);

/* Replaces Image/LoadImage + SetParams (src/utils/image.cpp:5-18, src/norm.cu:156-192): takes
 * one W x H float32 micrograph and its CTF, applies the reference's host-side clipping and
 * inversion, and uploads it.  The caller's pixels are not modified. */
gs_status gs_set_image(gs_engine* e, const float* image, int width, int height,
                       const gs_ctf* ctf);

/* Replaces work() (src/norm.cu:620-660, src/nonorm.cu:559-589) minus the text output: runs the
 * whole device pipeline on the image set by gs_set_image.  Asynchronous work completes before
 * return. */
gs_status gs_run(gs_engine* e);

/* Records of the last gs_run, sorted by descending score, after NMS when nms_radius > 0.
 * *out is allocated by the library (free with gs_free); *n_out receives the count. */
gs_status gs_get_records(gs_engine* e, gs_record** out, size_t* n_out);

/* gs_set_image + gs_run + gs_get_records. */
gs_status gs_search(gs_engine* e, const float* image, int width, int height, const gs_ctf* ctf,
                    gs_record** out, size_t* n_out);

/* Formats one record as the reference's output line (src/norm.cu:610-615):
 * "%d\t%s\tdefocus=%f\tdfdiff=%f\tdfang=%f\teuler=%f,%f,%f\tcenter=%f,%f\tscore=%f\n".
 * Writes at most cap bytes including the terminator; returns the untruncated length. */
int gs_format_record(const gs_engine* e, const gs_record* r, int image_index, const char* path,
                     const gs_ctf* ctf, int tile_local, char* buf, size_t cap);

void gs_free(void* p);
void gs_destroy(gs_engine* e);
const char* gs_last_error(void);

/* ---- inspection entry points (no counterpart in the reference; used by the parity tests) ---- */
typedef enum {
  GS_STAGE_TEMPLATES = 0,  /* norm: processed real-space templates, n*l*l float32;
                              nonorm: Fourier-space templates, n*(l/2+1)*l complex64 [t][kx][ky] */
  GS_STAGE_IMAGE = 1,      /* preprocessed micrograph (norm) W*H float32 */
  GS_STAGE_BLOCKS = 2,     /* norm: block spectra nb*(l/2+1)*l complex64 [b][kx][ky], scaled by 1/l^2 */
  GS_STAGE_ROTATED = 3,    /* norm: template spectra of the last angle run, n*(l/2+1)*l complex64 */
  GS_STAGE_MU = 4,         /* norm: per block pixel mean, nb*l*l float64 */
  GS_STAGE_SIGMA = 5,      /* norm: per block pixel sigma, nb*l*l float64 */
  GS_STAGE_BEST = 6,       /* norm: per block pixel best correlation, nb*l*l float32 */
  GS_STAGE_ARG = 7,        /* norm: per block pixel argmax angle_index*n+template, nb*l*l int32 */
  GS_STAGE_PLANES = 8      /* nonorm: peak, index, sum, sum2 float64 per [template][angle][tile][block] */
} gs_stage;

/* Size in bytes of a stage buffer after the last run (0 if not available). */
size_t gs_stage_bytes(const gs_engine* e, int stage);
/* Copies a stage buffer to host memory (bytes must equal gs_stage_bytes). */
gs_status gs_fetch_stage(gs_engine* e, int stage, void* out, size_t bytes);
/* Restricts the next gs_run to the first n_angles angles and n_templates templates (<= 0 means
 * all); mu and sigma then use the restricted counts.  For tests and timing samples. */
gs_status gs_set_limits(gs_engine* e, int n_angles, int n_templates);
/* Standalone transform of `batch` contiguous complex rows of length n on `device` (host
 * buffers, interleaved re/im float32, in place, unnormalised; sign -1 forward, +1 inverse). */
gs_status gs_fft_c2c(int device, float* data, int n, int batch, int sign);
/* Standalone 2-D transform of `batch` row-major ny x nx complex planes (in place). */
gs_status gs_fft2_c2c(int device, float* data, int ny, int nx, int batch, int sign);
/* gs_fft2_c2c without any allocation, for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  No counterpart in the
 * reference.  d_data: `batch` row-major ny x nx complex planes (interleaved re/im float32) in
 * device memory of `device`, transformed in place; the work is queued on `stream` (a hipStream_t,
 * NULL = the default stream) and the call returns without waiting for it.  The transform's work
 * buffer (batch * ny * nx + nx + ny complex elements) is a static 64 MiB device array of the
 * library, one per device, held until the process ends: larger requests fail with
 * GS_ERR_INVALID before any work is queued.  Nothing in it needs clearing between calls (every
 * element read is written first by the same call); calls that overlap on one device share that
 * array and race. */
gs_status gs_fft2_c2c_noalloc(int device, float* d_data, int ny, int nx, int batch, int sign,
                              void* stream);
/* Measured device-memory copy rate (read + write bytes per second, GB/s) of a plain streaming
 * copy of `bytes` repeated `reps` times: the practical HBM ceiling reported next to the peak. */
gs_status gs_copy_bandwidth(int device, size_t bytes, int reps, double* gbps);
/* Wall time of the device pipeline of the last gs_run, and of its angle loop, in ms. */
gs_status gs_last_timing(const gs_engine* e, double* total_ms, double* corr_ms);
/* When on, every correlation launch is bracketed by HIP events on the engine's stream. */
gs_status gs_set_profiling(gs_engine* e, int on);
/* Average duration (ms) and count of the timed launches: which = 0 column pass, 1 row pass. */
gs_status gs_kernel_timing(const gs_engine* e, int which, double* avg_ms, long long* count);
/* Templates per correlation launch (0 = automatic, negative: GS_ERR_INVALID); a result-invariant
 * tuning knob. */
gs_status gs_set_chunk(gs_engine* e, int templates_per_launch);
/* Selects a result-invariant code variant of a correlation kernel.  kernel 0 = column pass:
 * 0 one workgroup per plane (radix 4*4*4*5), 1 persistent over templates, 2 persistent +
 * prefetch, 3 as 0 with radix 4*4*20 (window 720: as 0), 4/5 as 3/0 with the
 * bank-conflict-free LDS layout (swizzled slots; per-lane twiddles in registers for 4), 6 as 3
 * with an unpadded column buffer and twiddles from global memory, 7 as 6 with twiddles formed
 * per lane by recurrence, 8 as 7 writing the row-packed inter-pass layout, 9 as 8
 * looping each workgroup over 8 templates with the lane twiddles in registers, 10 as 8 with the
 * lane twiddles loaded into registers, 11 as 9 with the block's columns held in registers and
 * the next template's columns loaded during the current transform, 12 (default) and 13 as 8 with
 * the column group the fastest index of the workgroup order, tiles of 4 x 4 and 2 x 8 (block,
 * template) planes (window 720: 7 and above as 3);
 * kernel 1 = row pass: 0 plain, 1 software-prefetched loads, 2/3 as 1/0 with one row and eight
 * templates per wave iteration, 4/5/6 as 1/0/2 with the bank-conflict-free LDS layout and
 * streaming loads, 7 as 6 reading the row-packed layout, 8 (default) as 7 with the running
 * maximum started from the earlier launches' record and the per-template argmax update taken
 * only where a value beats it, 9 as 6 with that update (window 720: one variant).
 * Column variants 8 to 13 (row-packed layout) and row variants 7, 8 go together (window 320,
 * norm_type=1): a search with only one side of a pair fails with GS_ERR_INVALID;
 * kernel 2 = pipeline: 0 both passes on one stream (default), 1 row pass on a second stream
 * overlapping the next column pass through a double-buffered inter-pass buffer. */
gs_status gs_set_variant(gs_engine* e, int kernel, int variant);

/* ---- batching ---- */

/* One micrograph of a batch.  image/width/height/ctf as in gs_search.  templates (n_templates x
 * template_height x template_width float32) and alt_az (may be NULL) replace the engine's
 * template set before this micrograph, as gs_load_templates; templates == NULL keeps the set of
 * the previous item (or of the last gs_load_templates for the first item). */
typedef struct {
  const float* image;
  int32_t width, height;
  gs_ctf ctf;
  const float* templates;
  int32_t n_templates, template_width, template_height;
  const float* alt_az;
} gs_batch_item;

/* Replaces main's per-micrograph loop (src/main.cu:37-49, one engine built and run per
 * micrograph): searches n_items micrographs in order on one engine whose buffers persist, with
 * the host-side clipping of micrograph i+1 staged on a helper thread while micrograph i runs on
 * the device.  counts[i] receives the number of records of item i; *out receives all records
 * concatenated in item order (free with gs_free).  The records of each item equal those of
 * gs_search on that item with the same template set.  All items are validated before any
 * work starts. */
gs_status gs_search_batch(gs_engine* e, const gs_batch_item* items, int n_items, gs_record** out,
                          size_t* counts);

/* When on (default off), the norm_type=1 angle loop is captured as a HIP graph the first time a
 * problem shape and buffer set occurs and replayed for later runs of the same shape; turning it
 * off releases the captured graphs.  Profiling runs are never captured. */
gs_status gs_set_graph(gs_engine* e, int on);
/* Number of graph captures and graph launches so far. */
gs_status gs_graph_stats(const gs_engine* e, long long* captures, long long* launches);

/* Times the result-invariant options of the norm_type=1 correlation kernels (column-pass and
 * row-pass variants, templates per launch) on a one-angle sample of the current image and
 * templates, keeps the fastest and reports them (any output pointer may be NULL).  Overwrites
 * the results of the last run. */
gs_status gs_autotune(gs_engine* e, int* col_variant, int* row_variant, int* templates_per_launch);

/* Records of the last norm_type=1 run selected again at another threshold and NMS radius from
 * the per-pixel best records kept on the device (the scan of src/norm.cu:583-597 repeated
 * without a new search); the run's own records are unchanged.  Free with gs_free. */
gs_status gs_select(gs_engine* e, float threshold, float nms_radius, gs_record** out, size_t* n_out);

/* ---- adjoint refinement (no reference counterpart: the reference's picks stay on the pixel
 * grid, the phistep angle grid and the micrograph defocus) ----
 * A pick on the micrograph of the last norm_type=1 run: its block, in-block pixel (y*l + x) and
 * template (all from its gs_record), plus the continuous parameters refined: sub-pixel offset
 * (dx, dy) px from the pixel, template rotation angle theta in degrees (the search's angle-list
 * value; psi = 360 - theta) and a defocus offset ddf in Angstrom added to the micrograph
 * defocus.  The score is z = (c - mu) / sigma, with c the search's correlation made continuous
 * in all four parameters (band-limited shift, bilinear rotation, template preprocessing at the
 * offset defocus; the micrograph side stays as searched) and mu, sigma frozen at the pick's
 * block pixel.  At zero offsets and theta on the angle grid, z is the search's score. */
typedef struct {
  int32_t block, pixel, template_id;
  double dx, dy, theta, ddf;
} gs_peak;

/* Ascent options: `iterations` steps p += step * scale^2 * dz/dp for each parameter p in
 * (dx, dy, theta, ddf), i.e. gradient ascent with step `step` in the coordinates p / scale. */
typedef struct {
  int32_t iterations;
  double step;
  double scale[4];
} gs_refine_opts;

/* Defaults: 20 iterations, step 0.1, scale (2 px, 2 px, 3 deg, 500 A). */
gs_refine_opts gs_refine_defaults(void);

typedef struct {
  gs_peak peak;    /* refined parameters */
  double score;    /* z at the refined parameters */
  double score0;   /* z at the starting parameters */
  double grad[4];  /* dz/d(dx, dy, theta, ddf) at the refined parameters (per px, px, deg, A) */
} gs_refined;

/* The pick of a norm_type=1 record: block, pixel, template, theta of its angle index, zero
 * offsets. */
gs_status gs_peak_from_record(const gs_engine* e, const gs_record* r, gs_peak* out);
/* z and dz/d(dx, dy, theta, ddf) of k picks (score: k values, grad: k x 4). */
gs_status gs_score_and_grad(gs_engine* e, const gs_peak* peaks, size_t k, double* score, double* grad);
/* Gradient ascent of k picks (opts NULL = defaults); out receives k results. */
gs_status gs_refine(gs_engine* e, const gs_peak* peaks, size_t k, const gs_refine_opts* opts,
                    gs_refined* out);
/* Device time (ms) of the last gs_refine or gs_score_and_grad call. */
gs_status gs_refine_timing(const gs_engine* e, double* ms);

/* gs_search_batch (norm_type=1) with refinement: after each micrograph is searched, its first
 * top_k records (its best picks) are refined on it with opts (NULL = defaults) before the next
 * micrograph.  *refined receives the refinement results in item order (free with gs_free) and
 * refined_counts[i] their number for item i. */
gs_status gs_search_batch_refine(gs_engine* e, const gs_batch_item* items, int n_items, int top_k,
                                 const gs_refine_opts* opts, gs_record** out, size_t* counts,
                                 gs_refined** refined, size_t* refined_counts);

#ifdef __cplusplus
}
#endif

#endif /* GISSPA_H */
