/*
 * FP64 CPU oracle for the GisSPA template-matching search.
 *
 * TEST INFRASTRUCTURE ONLY.  This library restates, in double precision, the
 * scoring of the reference (src/kernels.cu, src/norm.cu, src/nonorm.cu) so that
 * the HIP path can be checked against it.  Only tests/, the smoke check and the
 * CPU-baseline leg of bench.py may load it; the product path never does.
 *
 * Conventions shared with the HIP library (include/gisspa.h):
 *  - images are row-major float32, width W (x, fastest) by height H (y);
 *  - templates are a stack of n row-major w x h float32 images;
 *  - CTF parameters are (defocus um, dfdiff um, dfang deg) as in the LST file
 *    (src/EMReader/DataReader2.cpp:80-101).
 */
#pragma once
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Search parameters: the reference's config keys (include/EMReader/DataReader2.h:17-44,73-83). */
typedef struct {
  float angpix, phistep, kk, energy, cs, highres, lowres, diameter, threshold;
  int window_size, overlap, norm_type, invert, phase_flip;
  float nms_radius; /* radius of the greedy non-maximum suppression, px (0 = off) */
} orc_config;

/* One pick: centre in micrograph pixels, template index, in-plane angle as printed by the
 * reference (norm: 360-theta, src/norm.cu:654-655; nonorm: theta, src/nonorm.cu:584),
 * score, plus bookkeeping used by the tests. */
typedef struct {
  float x, y;
  int32_t template_id;
  float psi;
  float score;
  int32_t angle_index;  /* position of theta in the angle loop */
  int32_t block;        /* norm: block index; nonorm: tile index */
  int32_t pixel;        /* norm: in-block pixel index; nonorm: in-block peak index */
  float local_x, local_y; /* nonorm: tile-local centre as the reference prints it */
} orc_pick;

/* ---- FFT (mixed radix, unnormalised, sign -1 forward / +1 inverse) ---- */
int orc_fft1d(double* data, int n, int sign);          /* interleaved complex, in place */
int orc_fft2d(double* data, int ny, int nx, int sign); /* row-major ny x nx, in place */

/* ---- scalar pieces ---- */
double orc_ctf(int x, int y, int nx, int ny, double apix, double defocus, double dfdiff,
               double dfang, double energy, double cs, double ampconst, int mode);
int orc_radial_bin(int dx, int dy); /* floor(hypot(dx,dy)+0.5)-1, exact */
int orc_block_offsets(int length, int l, int overlap, int* out, int max_out);
int orc_tile_geometry(int W, int H, int* ul_x, int* ul_y, int max_out);
int orc_angles(float phistep, float* out, int max_out); /* float loop of src/norm.cu:642 */

/* ---- stages ---- */
/* Template preprocessing.  norm (norm_type=1): real-space processed templates (n*l*l, re/im
 * interleaved).  nonorm: Fourier-space processed templates (n*l*l complex). */
int orc_prep_templates(const orc_config* c, const float* tmpl, int n, int w, int h,
                       const double* ctf3, double* out);
/* Host clipping of src/norm.cu:156-181 (norm only; modifies img in place). */
int orc_clip_image(float* img, int W, int H, int invert);
/* Whole-frame preprocessing of src/norm.cu:356-402 (norm) or src/nonorm.cu:346-389 (nonorm,
 * on one tile).  img is the already clipped/inverted image; out is W*H real. */
int orc_prep_image(const orc_config* c, const float* img, int W, int H, const double* ctf3,
                   double* out);
/* Rotated, masked, standardised template spectra for angle theta (norm): n*l*l complex. */
int orc_rotate_templates(const orc_config* c, const double* prepped, int n, float theta,
                         double* out);

/* ---- full searches ---- */
/* norm_type=1 search with per-(block,pixel) best records (single pass; equivalent to the
 * reference's two passes because z is monotone in the correlation at fixed pixel).
 * Optional outputs (may be NULL): mu/sigma/best/arg per block pixel (nb*l*l each) where
 * arg = angle_index*n + template.  n_angles_limit/n_templates_limit (<=0 = all) restrict the
 * work for timing samples; mu/sigma then use the restricted counts.  Returns #picks. */
int orc_search_norm(const orc_config* c, const float* image, int W, int H, const float* tmpl,
                    int n, int w, int h, const double* ctf3, orc_pick* picks, int max_picks,
                    double* mu, double* sigma, double* best, int32_t* arg, int n_angles_limit,
                    int n_templates_limit);
/* norm_type=0 search: every (tile, theta, template) record above threshold (the reference's
 * output lines) into recs, then NMS picks.  Returns #picks; *n_recs gets #records. */
int orc_search_nonorm(const orc_config* c, const float* image, int W, int H, const float* tmpl,
                      int n, int w, int h, const double* ctf3, orc_pick* recs, int max_recs,
                      int* n_recs, orc_pick* picks, int max_picks);
/* Reference records of the norm_type=1 search (the second pass of src/norm.cu:560-600): one
 * per (angle, block, template, 256-element chunk) whose chunk maximum of z reaches the
 * threshold with its centre inside [d_m, W - d_m) x [d_m, H - d_m).  margin: the chunk maximum
 * minus the largest other z of the chunk (how far the chosen pixel is from a tie). */
typedef struct {
  float x, y;
  int32_t template_id, angle_index, block, chunk, pixel;
  float score;
  double margin;
} orc_refrec;

/* norm_type=1 statistics of the listed blocks only (each block's mu/sigma/best depend on that
 * block alone): per listed block pixel ([list index][pixel]) mu, sigma, the best and second-best
 * correlation c/l^2 and their hypotheses (angle_index * n + template).  Limits as in
 * orc_search_norm. */
int orc_search_norm_blocks(const orc_config* c, const float* image, int W, int H, const float* tmpl,
                           int n, int w, int h, const double* ctf3, const int* blocks, int nlist,
                           int n_angles_limit, int n_templates_limit, double* mu, double* sigma,
                           double* best, double* second, int32_t* arg, int32_t* arg2);
/* Reference records over the listed blocks (blocks NULL or nlist 0: all), in the reference's
 * order: angle, block, template, chunk.  Returns the number of records (l*l must be divisible
 * by 256). */
int orc_search_norm_records(const orc_config* c, const float* image, int W, int H, const float* tmpl,
                            int n, int w, int h, const double* ctf3, const int* blocks, int nlist,
                            int n_angles_limit, int n_templates_limit, orc_refrec* recs, int max_recs);

/* ---- stage pieces (pinned one by one in the tests) ---- */
/* SQRSum_by_circle + whiten_Tmp (k:94-149) on one l x l complex spectrum, in place. */
int orc_whiten_template(double* F, int l);
/* apply_weighting_function (k:243-288) on one l x l complex spectrum, in place. */
int orc_weight_template(const orc_config* c, const double* ctf3, double* F, int l);
/* do_phase_flip + SQRSum_by_circle (mode 1) + whiten_filter_weight_Img (k:94-202, 828-839) on a
 * scaled whole-frame spectrum (ny x nx complex), in place. */
int orc_filter_image_spectrum(const orc_config* c, const double* ctf3, double* F, int nx, int ny);
/* rotate_subIMG (k:534-603) of one l x l complex plane. */
int orc_rotate_plane(const double* in, double* out, int l, float theta);
/* norm_type=0 plane reduction (get_peak_and_SUM, k:725-763): stats4 = peak, peak index, sum,
 * sum of squares over an l x l complex plane's real parts (border d_m/4). */
int orc_nonorm_plane(const double* C, int l, float d_m, double* stats4);
/* norm_type=0 score peak / sd of the other l*l - 1 values (src/nonorm.cu:511-519). */
double orc_nonorm_score(const double* stats4, int l);
/* norm_type=0 centre of a peak index in block (bi, bj) at angle theta (src/nonorm.cu:521-524). */
void orc_nonorm_center(long long pos, int bi, int bj, int l, int overlap, float theta, float* cx, float* cy);

/* Greedy NMS: sort by score (desc; ties by y, x, template, angle), keep a candidate if no kept
 * pick lies strictly closer than radius.  In place; returns the number kept. */
int orc_nms(orc_pick* cand, int n, float radius);
/* compute_corner_CCG + inverse transform (k:645-678): out = IDFT_unnorm(I conj(T) checker),
 * complex l x l planes. */
int orc_correlate(const double* I, const double* T, int l, double* out);

/* ---- adjoint refinement of norm_type=1 picks (no reference counterpart) ----
 * A pick is its block, in-block pixel (y*l + x) and template, plus the continuous parameters
 * refined: sub-pixel offset (dx, dy) px, template rotation angle theta (deg; the search's
 * angle-list value, psi = 360 - theta) and defocus offset ddf (Angstrom, added to the
 * micrograph defocus).  z = (c - mu) / sigma with mu, sigma frozen at the pick's block pixel. */
typedef struct {
  int32_t block, pixel, template_id;
  double dx, dy, theta, ddf;
} orc_peak;

/* Micrograph context: preprocessed block spectra and whitened template spectra.  NULL on
 * failure; free with orc_refine_free. */
void* orc_refine_create(const orc_config* c, const float* image, int W, int H, const float* tmpl,
                        int n, int w, int h, const double* ctf3);
void orc_refine_free(void* ctx);
/* 1: rotation sample positions in float as the search computes them; 0 (default): double. */
void orc_refine_geometry(void* ctx, int float_geometry);
/* z, gradient dz/d(dx, dy, theta, ddf) (k x 4, per px, deg, Angstrom) and optionally c. */
int orc_refine_eval(void* ctx, const orc_peak* peaks, int k, const double* mu, const double* sigma,
                    double* z, double* grad, double* c);
/* iters ascent steps p += step * scale^2 * dz/dp per parameter, then the final evaluation;
 * peaks are updated in place; z0 (may be NULL) receives the starting z. */
int orc_refine_run(void* ctx, orc_peak* peaks, int k, const double* mu, const double* sigma,
                   int iters, double step, const double* scale4, double* z, double* grad, double* z0);
/* Correlation plane Re(IDFT(I_b conj(T) checker))/l^2 of template t rotated by theta with its
 * preprocessing at defocus offset ddf, through the search's own stage functions. */
int orc_refine_plane(void* ctx, int block, int t, double theta, double ddf, double* out);

int orc_num_threads(void);
void orc_set_num_threads(int n);

#ifdef __cplusplus
}
#endif
