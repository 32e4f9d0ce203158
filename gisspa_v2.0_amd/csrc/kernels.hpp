// Device parameter block and kernel launchers of the GisSPA pipeline.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace gs {

class FftCache;

// Per-micrograph parameters (reference Parameters, include/EMReader/DataReader2.h:73-83, with
// SetParams, src/norm.cu:183-192).  CTF quantities are kept in double: the CTF phase reaches
// tens of radians on whole frames, and these kernels run once per micrograph.
struct DevParams {
  float apix, kk, highres, lowres, d_m, thresh;
  double lambda, dfu, dfv, dfdiff, dfang, cs, ampconst;
  float ehw, a, b, b2, bf, bf2, bf3;
  // Band-pass radii and thresholds: float expressions of config values in the reference
  // (k:179,183-189,265-273), evaluated once on the host.
  float hr_t, lr_t;          // template band at l
  float hr_i, lr_i, hi_thr;  // whole-frame band at max(nx, ny)
  float ds;                  // 1/(apix*l), src/norm.cu:191
};

// Candidate pick written by the finalise kernel (norm_type=1).
struct Cand {
  float x, y, score;
  int32_t arg;    // angle_index * n_templates + template
  int32_t block;  // block index
  int32_t pixel;  // in-block pixel index
};

// ---- template preparation (src/norm.cu:194-354, src/nonorm.cu:178-344) ----
void launch_pad_templates(const float* raw, int n, int w, int h, int l, float2* out, hipStream_t s);
// part: scratch of 64 doubles per template.
void launch_template_moments(const float2* T, int n, int l, double* part, double* mom, hipStream_t s);
void launch_template_standardise(float2* T, int n, int l, const double* mom, hipStream_t s);
// part: scratch of 16 * (l/2) doubles per plane.
// Workgroups per plane of the chunked per-plane reductions (template moments, radial power, area
// power): 16 for many planes, more for few so that a small batch still fills the GPU.  Their
// partial buffers hold n * plane_reduction_chunks(n) entries per reduced quantity.
int plane_reduction_chunks(int n);
void launch_radial_power_planes(const float2* F, int n, int l, double* part, double* ra, hipStream_t s);
void launch_whiten_planes(float2* F, int n, int l, const double* ra, const double* rb, hipStream_t s);
void launch_weight_planes(float2* F, int n, int l, const DevParams& p, hipStream_t s);
void launch_mask_planes(float2* T, int n, int l, float d_m, float ehw, hipStream_t s);
// Spectral-normalisation factor m per plane (compute_area_sum_ofSQR k:319-382 + the host mean of
// src/norm.cu:334-342); part is scratch of 16 double2 per plane.
void launch_area_power_full(const float2* F, int n, int l, double2* part, double* m, hipStream_t s);
void launch_area_power_half(const float2* F, int n, int l, double2* part, double* m, hipStream_t s);
void launch_scale_planes(float2* F, int n, long long plane, const double* m, float extra,
                         hipStream_t s);
void launch_extract_real(const float2* F, float* out, long long count, hipStream_t s);

// ---- whole-frame preparation (src/norm.cu:356-402, src/nonorm.cu:346-389) ----
void launch_real_to_complex(const float* in, float2* out, long long count, float sign, hipStream_t s);
// The frame functions take nf frames of nx x ny, contiguous; ra holds max(nx, ny) / 2 bins per
// frame and part is scratch of nf * radial_image_groups(nf) * (max(nx, ny) / 2) doubles.
void launch_phase_flip(float2* F, int nx, int ny, int nf, const DevParams& p, hipStream_t s);
int radial_image_groups(int nf);
void launch_radial_power_image(const float2* F, int nx, int ny, int nf, double* part, double* ra, hipStream_t s);
void launch_whiten_filter_image(float2* F, int nx, int ny, int nf, const double* ra, const double* rb,
                                const DevParams& p, hipStream_t s);
void launch_zero_dc(float2* F, hipStream_t s);
// ntiles tiles of tile x tile (upper-left corners ul[2k], ul[2k + 1], device array) as the real
// parts of complex frames, times sign.
void launch_extract_tiles(const float* img, int W, const int* ul, int tile, int ntiles, float sign, float2* out,
                          hipStream_t s);

// ---- blocks (src/norm.cu:404-477, src/nonorm.cu:391-450) ----
void launch_split_blocks(const float* img, int nx, int ny, int l, const int* offx, const int* offy,
                         int bx, int nb, float2* out, hipStream_t s);
// Blocks of every tile from the real parts of complex tiles: out[k * nbk^2 + b], block b at
// (off[b % nbk], off[b / nbk]) of tile k.
void launch_split_tile_blocks(const float2* F, int tile, int l, const int* off, int nbk, int ntiles, float2* out,
                              hipStream_t s);

// ---- per-angle rotation (src/norm.cu:479-521, src/nonorm.cu:457) ----
// Template pairs (2p, 2p+1) rotated and masked into the real and imaginary parts of out[p]
// ((n + 1) / 2 complex l x l planes); partial holds n x rotation_groups(l) x 2 double sums.
int rotation_groups(int l);
void launch_rotate_mask_pairs(const float* tp, int n, int l, float cose, float sine, float d_m, float ehw,
                              float2* out, double* partial, hipStream_t s);
void launch_rotation_stats(const double* partial, int n, int l, float* stats, hipStream_t s);
// Half spectra out[i * stride + g][kx 0..l/2][ky] (times scale; standardised when stats is not
// null) of items i < n from the pair transforms Z[j * stride + g] holding items 2j and 2j + 1.
void launch_unpack_pair_spectra(const float2* Z, int n, int stride, int l, const float* stats, float scale,
                                float2* out, hipStream_t s);
// Rotated copies of nb real blocks in[g] for na angles cs[a] = (cos, sin) (device array), two
// angles per complex plane: out[j * nb + g] = rot(in[g], cs[2j]) + i rot(in[g], cs[2j+1]).
void launch_rotate_real_pairs(const float* in, int nb, int l, const float2* cs, int na, float2* out,
                              hipStream_t s);

// ---- correlation: product, inverse column and row transforms, per-pixel or per-plane reductions ----
struct CorrArgs {
  const float2* that;  // [>=t0+tc][H+1][L] template spectra
  const float2* ihat;  // [nb][H+1][L] image spectra
  float2* Z;           // [tc][nb][L][H] inter-pass buffer
  int t0, tc, nb;
};
// Column pass: packed conj-multiply + inverse column FFTs into Z (returns false if L
// unsupported).  variant (result-invariant): 0 one workgroup per plane and column group,
// 1 persistent over templates, 2 persistent with prefetch.
bool launch_colpass(int L, const CorrArgs& a, int variant, hipStream_t s);
// Row pass, norm_type=1: C2R rows + per-(block,pixel) running moments and best record.
// variant (result-invariant): 0 plain, 1 software-prefetched Z loads.
bool launch_rowpass_pixel(int L, const float2* Z, int tc, int nb, double* S, double* SS,
                          float* best, int32_t* arg, int arg_base, int variant, hipStream_t s);
// Column pass of the norm_type=0 correlation: as launch_colpass, except at L = 720, where one
// column per wavefront writes the column-grouped inter-pass layout that the L = 720 row pass of
// launch_rowpass_plane reads (variant ignored there).
bool launch_colpass_plane(int L, const CorrArgs& a, int variant, hipStream_t s);
// Row pass, norm_type=0: C2R rows + per-(template,block,row group) peak/sum partials.
struct RefRec;
// The reference's records of one launch from the row-packed inter-pass buffer (l = 320; false
// elsewhere): per template and 256-element chunk part of each row, the best z at or above
// thresh, appended to rec (RefRec::chunk, pixel, score); the parts of a chunk are merged, and
// the centre range applied, by the caller.
bool launch_rowpass_chunk_records(int L, const float2* Z, int tc, int nb, const double* mu, const double* sigma,
                                  float thresh, int angle, int t0, RefRec* rec, int* count, int max_out,
                                  hipStream_t s);
bool launch_rowpass_plane(int L, const float2* Z, int tc, int nb, float4* partial, float border,
                          hipStream_t s);
int rowpass_groups(int L);
// True where the tiled correlation kernels exist (L = 320 and 720); any other even L runs the
// generic path below.
bool corr_supported(int L);
// Generic correlation for any even L (corr_generic.hip): P and W are two caller buffers of
// corr_generic_buffer(L, tc, nb) elements each; the pixel form updates the same per-pixel
// records as launch_rowpass_pixel, the plane form writes one partial per (template, block)
// (finalise with groups = 1).  corr_generic_prepare builds its plans and tables (nothing is
// then allocated on later calls).
size_t corr_generic_buffer(int L, int tc, int nb);
void corr_generic_prepare(FftCache& cache, int L);
void launch_corr_generic_pixel(FftCache& cache, const CorrArgs& a, int L, float2* P, float2* W, double* S,
                               double* SS, float* best, int32_t* arg, int arg_base, hipStream_t s);
void launch_corr_generic_plane(FftCache& cache, const CorrArgs& a, int L, float2* P, float2* W, float4* partial,
                               float border, hipStream_t s);
// Reference records (the second pass of src/norm.cu:560-600): one per (template, block,
// 256-element chunk of the correlation index y*L + x) whose maximum of z = (c - mu)/sigma, taken
// by get_peak_pos's tree (k:766-793), reaches thresh with its centre in [d_m, W - d_m) x
// [d_m, H - d_m); appended in any order (count is an atomic counter).  L*L % 256 == 0.
struct RefRec {
  int32_t angle, block, templ, chunk, pixel;
  float score;
};
void launch_corr_generic_records(FftCache& cache, const CorrArgs& a, int L, float2* P, float2* W,
                                 const double* mu, const double* sigma, int angle, float thresh,
                                 const int* offx, const int* offy, int bx, int Wimg, int Himg, float d_m,
                                 RefRec* out, int* count, int max_out, hipStream_t s);
// Builds the correlation kernels' twiddle tables on the current device (they are otherwise
// built on first use, which must not happen inside a stream capture).
void corr_prepare();
// In-place batched 2-D transform of `batch` full L x L complex planes (row-major), unnormalised
// then multiplied by scale; LDS transforms for L = 320 and 720 (returns false otherwise).
bool launch_plane_fft2(float2* data, int batch, int L, int sign, float scale, hipStream_t s);
// The same transform of `in` written to `out` (in is left unchanged).  Input rows y with
// min(y, L - y) > band are taken as zero without being read (default: every row is read).  The
// result in `out` is complete in the columns [clo, chi] (default: all; the second pass skips
// the groups of columns outside).
bool launch_plane_fft2_from(const float2* in, float2* out, int batch, int L, int sign, float scale, hipStream_t s,
                            int band = -1, int clo = 0, int chi = -1);
// Forward transforms of npairs real-pair planes `in` (items 2j and 2j + 1 of group g in the real
// and imaginary parts of plane j * stride + g; npairs = ceil(n / 2) * stride) split into the half
// spectra out[i * stride + g][kx 0..L/2][ky] of the n items per group, times scale and, with
// stats (mean, variance per item), standardised as by launch_unpack_pair_spectra.  work holds
// npairs planes (may be `in`).  L = 320, 720 (returns false otherwise).
bool launch_plane_fft2_pairs(const float2* in, float2* work, int npairs, int L, int n, int stride,
                             const float* stats, float scale, float2* out, hipStream_t s);
// The flat twiddle table exp(-2 pi i e / L) of the plane transforms on the current device
// (L = 320, 720; nullptr otherwise).
const float2* plane_twiddles(int L);

// ---- adjoint refinement of norm_type=1 picks ----
// Device record of one pick: block, in-block pixel, template, and the refined parameters
// (sub-pixel shift px, rotation angle deg, defocus offset Angstrom).
struct RefPeak {
  int32_t block, pixel, t, pad;
  double dx, dy, theta, ddf;
};
// Defocus-independent factors of the template weight (apply_weighting_function + CTF_AST,
// k:243-317) at one spectrum entry where the band-pass is not zero.
struct RefWeight {
  int32_t idx;    // y * l + x
  int32_t flags;  // bit 0: in the normalisation region (x <= l/2, radial bin < l/2); bit 1: CTF applies
  double band;    // band-pass factor
  double a0;      // cs * 2.5e6 * lambda^3 * s^4 (spherical-aberration term of gamma)
  double s2;      // s^2 of CTF_AST (1/A^2)
  double cs_ast;  // cos(2 (dfang - angle))
  double N;       // noise / signal curve N/S' of the Wiener-like weight
};
// Number of half-plane spectrum entries with radial bin < l/2 (the normalisation count).
double ref_region_count(int l);
// Spectrum indices y*l + x whose template band-pass factor can be non-zero (radial bin below
// hr + 8 or inside [lr - 8, lr)), in increasing order.
std::vector<int32_t> ref_band_indices(int l, float hr, float lr);
// Weight table at the given indices for the micrograph parameters p.
void launch_ref_table(const int32_t* idx, int n, int l, const DevParams& p, RefWeight* tab, hipStream_t s);
// Weighted, normalised spectrum of each pick's template at its defocus, packed with its
// defocus derivative: H = G + i dG (K planes of l*l).  Only the table entries are written (the
// weight is zero elsewhere): the rest of H must hold zeros (launch_ref_clear).
void launch_ref_template(const float2* white, const RefWeight* tab, int ntab, const RefPeak* pk, int K,
                         int l, const DevParams& p, double2* partial, double count, float2* H, hipStream_t s);
// Zeros n complex values (a fill kernel, so that the call may be captured in a graph).
void launch_ref_clear(float2* H, long long n, hipStream_t s);
// From the inverse transform t + i dt in H: rotation and mask with their theta and defocus
// derivatives, packed into AB (A = r0 + i dr0/dtheta of pick j in plane j, B = dr0/d ddf of
// picks 2m and 2m+1 in the real and imaginary parts of plane K + m), and per pick the box
// statistics (stats: 8 doubles) of the standardisation.
// The rotation's soft mask (diameter d_m, edge half-width ehw) as an l x l float table, and the
// rows [*ylo, *yhi] outside which it is zero.
void launch_ref_mask(int l, float d_m, float ehw, float* mask, int* ylo, int* yhi, hipStream_t s);

// launch_ref_rotate fused with the row pass of the forward transform (l = 320, 720; returns
// false otherwise), with the mask table of launch_ref_mask: AB receives the A and B planes with
// their rows already transformed, except rows outside [ylo, yhi], which the mask makes zero and
// which are left unwritten: finish with launch_ref_cols_dot over that row support.  partial6
// needs K * l / 8 * 6 doubles, stats as launch_ref_rotate's.
bool launch_ref_rotate_rows(const float2* H, const RefPeak* pk, int K, int l, const float* mask, int ylo,
                            int yhi, double* partial6, double* stats, float2* AB, hipStream_t s);
void launch_ref_rotate(const float2* H, const RefPeak* pk, int K, int l, float d_m, float ehw,
                       double* partial6, double* stats, float2* AB, hipStream_t s);
// c and dc/d(dx, dy, theta, ddf) per pick from the transformed AB and the statistics (the
// standardisation is applied to the spectra); ihat is the block spectra transposed to
// [b][ky][kx 0..l/2].  out: K x kRefDotSplit x 5 partial sums over consecutive shares of the
// half spectrum, summed by launch_ref_update.
constexpr int kRefDotSplit = 4;
bool launch_ref_dot(const float2* ihat, const float2* AB, const RefPeak* pk, const double* stats, int K, int l,
                    double* out, hipStream_t s);
// The column pass of the AB planes (left by launch_ref_rotate_rows, rows outside [ylo, yhi]
// unwritten) fused with launch_ref_dot, the spectra scaled by `scale` (l = 320, 720; returns false
// otherwise); AB is only read.  ihat: the block half spectra [b][kx 0..l/2][ky].  ph: scratch for
// the picks' shift phase tables, K x (2 l + 2 (l/2 + 1)) entries.  out: K x ref_cols_dot_parts(l)
// x 5 partial sums.
int ref_cols_dot_parts(int l);
bool launch_ref_cols_dot(const float2* ihat, const float2* AB, const RefPeak* pk, const double* stats, int K, int l,
                         int ylo, int yhi, float scale, float2* ph, double* out, hipStream_t s);
// z and gradient per pick from cg's `parts` partial sums per pick; when update, one ascent step
// p += step * scale^2 * dz/dp.
void launch_ref_update(RefPeak* pk, int K, const double* cg, int parts, const double* mu, const double* sigma, int l,
                       double step, const double* scale4, bool first, bool update, double* z, double* z0,
                       double* grad, hipStream_t s);
// Reduce plane partials to (peak, index, sum, sum2) per (template, block): the record of
// (template t, block b) goes to out[(t * tstride + b) * 4].
void launch_finalize_plane(const float4* partial, int tc, int nb, int groups, long long tstride, double* out,
                           hipStream_t s);
void launch_init_accumulators(double* S, double* SS, float* best, int32_t* arg, long long count,
                              hipStream_t s);
// Per block pixel: mu, sigma, z of the best record; candidates above threshold.
void launch_finalize_norm(const double* S, const double* SS, const float* best, const int32_t* arg,
                          int nb, int l, double total_n, float thresh, const int* offx,
                          const int* offy, int bx, int W, int H, float d_m, double* mu,
                          double* sigma, Cand* cand, int* count, int max_cand, hipStream_t s);

}  // namespace gs
