// Preprocessing kernels: templates, whole frame, blocks and per-angle rotation.
//
// These stages run once per micrograph (or once per angle on the template side) and are small
// next to the correlation.  Every reduction is deterministic: double accumulation in an order
// fixed by the data layout and the launch shape, never by the scheduling of the hardware.
#include <cmath>

#include "common.hpp"
#include "device_common.hpp"
#include "kernels.hpp"

namespace gs {

namespace {

constexpr double kRefPi = 3.1415926535;  // include/utils/constants.h:3
constexpr int BS = 256;

inline unsigned nblk(long long n, int bs = BS) { return (unsigned)((n + bs - 1) / bs); }

using dev::band_image;
using dev::band_template;
using dev::kRotTX;
using dev::kRotTY;
using dev::block_sum;
using dev::ctf_ast;
using dev::rbin;

// ---------------------------------------------------------------- templates
__global__ void k_pad_templates(const float* __restrict__ raw, int n, int w, int h, int l,
                                float2* __restrict__ out) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long L2 = (long long)l * l;
  if (i >= L2 * n) return;
  const int t = (int)(i / L2);
  const int p = (int)(i - (long long)t * L2);
  const int x = p % l, y = p / l;
  const int sx = (l - w) / 2, sy = (l - h) / 2;  // src/norm.cu:137-138
  float v = 0.f;
  if (x >= sx && x < sx + w && y >= sy && y < sy + h)
    v = raw[(long long)t * w * h + (long long)(y - sy) * w + (x - sx)];
  out[i] = make_float2(v, 0.f);
}

// Per-plane reductions are split over plane_reduction_chunks(n) workgroups per plane (slices of
// the plane, gridDim.x of the reduction kernels), whose partials a finishing kernel adds in
// chunk order.
// Workgroups (row slices) of the whole-frame radial power.
constexpr int kImageGroups = 256;

// Ring sum/count (generate_mask + multiCount_dot, k:28-80) and box sums (UpdateSigma, k:4-26);
// grid (chunks, n), partials [t][chunk][4].
__global__ void k_template_moments(const float2* __restrict__ T, int l, double* __restrict__ part) {
  const int t = blockIdx.y, ch = blockIdx.x;
  const float2* P = T + (long long)t * l * l;
  const float r = l / 2.f - 2.f;
  float up = 0.f, low = 0.f;
  if (r > 1) {
    up = (r + 1) * (r + 1);
    low = (r - 1) * (r - 1);
  }
  const int chunks = gridDim.x;
  const int per = (l * l + chunks - 1) / chunks;
  const int p0 = ch * per, p1 = min(l * l, p0 + per);
  double ring = 0, cnt = 0, s = 0, s2 = 0;
  for (int p = p0 + threadIdx.x; p < p1; p += blockDim.x) {
    const int x = p % l, y = p / l;
    const float rr = (float)((x - l / 2) * (x - l / 2) + (y - l / 2) * (y - l / 2));
    const double v = P[p].x;
    if (rr >= low && rr <= up) {
      ring += v;
      cnt += 1;
    }
    s += v;
    s2 += v * v;
  }
  ring = block_sum(ring);
  cnt = block_sum(cnt);
  s = block_sum(s);
  s2 = block_sum(s2);
  if (threadIdx.x == 0) {
    double* o = part + ((long long)t * chunks + ch) * 4;
    o[0] = ring;
    o[1] = cnt;
    o[2] = s;
    o[3] = s2;
  }
}

__global__ void k_template_moments_finish(const double* __restrict__ part, int n, int chunks,
                                          double* __restrict__ mom) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  double a[4] = {0, 0, 0, 0};
  for (int ch = 0; ch < chunks; ++ch)
    for (int k = 0; k < 4; ++k) a[k] += part[((long long)t * chunks + ch) * 4 + k];
  for (int k = 0; k < 4; ++k) mom[4 * t + k] = a[k];
}

// scale_each (k:82-90) with the host sigma of src/norm.cu:268-273.
__global__ void k_template_standardise(float2* __restrict__ T, int n, int l, const double* __restrict__ mom) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long L2 = (long long)l * l;
  if (i >= L2 * n) return;
  const int t = (int)(i / L2);
  const double ring = mom[4 * t], cnt = mom[4 * t + 1], s = mom[4 * t + 2], s2 = mom[4 * t + 3];
  const double em = cnt != 0 ? ring / cnt : ring;
  const double mean = s / (double)L2;
  double sigma = sqrt(s2 / (double)L2 - mean * mean);
  if (sigma < 0 || !isfinite(sigma)) sigma = 0;
  if (sigma < 1e-8 && sigma > -1e-8) return;
  T[i].x = (float)(((double)T[i].x - em) / sigma);
}

// Radial power of one row of a half plane (x = 0 .. hx - 1, folded row index fy) added to a
// wave-private histogram.  Along such a row the bin floor(hypot(x, fy) + 1/2) - 1 is
// non-decreasing in x, so the pixels of one bin among 64 consecutive x are one run of lanes: a
// segmented inclusive scan over the wave (keyed by bin) leaves each run's sum in its last lane,
// which adds it to the histogram.  Runs have distinct bins, so no two lanes touch one entry, and
// every addition happens in an order fixed by the data layout (the reference's float atomics,
// k:123-124, sum in whatever order the hardware schedules them).
__device__ __forceinline__ void radial_row(const float2* __restrict__ row, int hx, int fy, int nbins,
                                           double* __restrict__ hist, int lane) {
  for (int x0 = 0; x0 < hx; x0 += 64) {
    const int x = x0 + lane;
    int key = -1 - lane;  // a lane outside the row or the bins: a run of its own, never added
    double v = 0.0;
    if (x < hx) {
      const int r = rbin(x, fy);
      if (r >= 0 && r < nbins) {
        const float2 f = row[x];
        v = (double)f.x * f.x + (double)f.y * f.y;
        key = r;
      }
    }
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const double pv = __shfl_up(v, d, 64);
      const int pk = __shfl_up(key, d, 64);
      if (lane >= d && pk == key) v += pv;
    }
    const int nk = __shfl_down(key, 1, 64);
    if (key >= 0 && (lane == 63 || nk != key)) hist[key] += v;
  }
}

// Sums the wave-private histograms in wave order into out[0 .. nbins).
__device__ __forceinline__ void radial_merge(const double* __restrict__ hist, int waves, int nbins,
                                             double* __restrict__ out) {
  for (int i = threadIdx.x; i < nbins; i += blockDim.x) {
    double s = 0.0;
    for (int w = 0; w < waves; ++w) s += hist[w * nbins + i];
    out[i] = s;
  }
}

// Radial power over the half plane x <= l/2 of each template spectrum (SQRSum_by_circle,
// k:94-126); grid (chunks, n), each workgroup a slice of rows (wave w takes rows w, w + 4, ...
// of it), partials [t][chunk][bin].
__global__ void k_radial_power_planes(const float2* __restrict__ F, int l, double* __restrict__ part) {
  extern __shared__ double hist[];
  const int t = blockIdx.y, ch = blockIdx.x, nbins = l / 2;
  const int waves = blockDim.x / 64, wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  for (int i = threadIdx.x; i < waves * nbins; i += blockDim.x) hist[i] = 0.0;
  __syncthreads();
  const float2* P = F + (long long)t * l * l;
  const int per = (l + (int)gridDim.x - 1) / (int)gridDim.x;
  const int y0 = ch * per, y1 = min(l, y0 + per);
  for (int y = y0 + wave; y < y1; y += waves)
    radial_row(P + (long long)y * l, l / 2 + 1, min(y, l - y), nbins, hist + wave * nbins, lane);
  __syncthreads();
  radial_merge(hist, waves, nbins, part + ((long long)t * gridDim.x + ch) * nbins);
}

// ra[t][bin] = sum over chunks of part[t][chunk][bin], one wavefront per (t, bin): lane j adds
// chunks j, j + 64, ... in order, then a fixed butterfly over the lanes (deterministic).
__global__ void k_radial_power_finish(const double* __restrict__ part, int n, int nbins, int chunks,
                                      double* __restrict__ ra) {
  const long long w = (long long)blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
  const int lane = threadIdx.x % 64;
  if (w >= (long long)n * nbins) return;  // uniform per wavefront
  const long long t = w / nbins, b = w % nbins;
  double s = 0;
  for (int ch = lane; ch < chunks; ch += 64) s += part[(t * chunks + ch) * nbins + b];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if (lane == 0) ra[w] = s;
}

// whiten_Tmp (k:128-149): |F| /= sqrt(ra[r]/rb[r]) for r in [0, l/2).
__global__ void k_whiten_planes(float2* __restrict__ F, int n, int l, const double* __restrict__ ra,
                                const double* __restrict__ rb) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long L2 = (long long)l * l;
  if (i >= L2 * n) return;
  const int t = (int)(i / L2);
  const int p = (int)(i - (long long)t * L2);
  const int x = p % l, y = p / l;
  const int r = rbin(min(x, l - x), min(y, l - y));
  if (r >= 0 && r < l / 2) {
    const double f = 1.0 / sqrt(ra[(long long)t * (l / 2) + r] / rb[r]);
    F[i] = make_float2((float)(F[i].x * f), (float)(F[i].y * f));
  }
}

// apply_weighting_function (k:243-288): band-pass and |CTF| Wiener-like weighting.  The factor
// depends on the pixel only, so each thread evaluates it once for its pixel and applies it to
// kWeightT planes (blockIdx.y selects the group of planes).
constexpr int kWeightT = 8;
__global__ void k_weight_planes(float2* __restrict__ F, int n, int l, DevParams p) {
  const long long L2 = (long long)l * l;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= L2) return;
  const int t0 = blockIdx.y * kWeightT;
  float2* const Fq = F + (long long)t0 * L2 + q;
  const int nt = min(kWeightT, n - t0);
  float2 v[kWeightT];
#pragma unroll
  for (int k = 0; k < kWeightT; ++k)
    if (k < nt) v[k] = Fq[k * L2];
  const int x = q % l, y = q / l;
  const int fx = min(x, l - x), fy = min(y, l - y);
  const int rr = rbin(fx, fy);
  double f = band_template(rr, p.hr_t, p.lr_t);
  if (rr < l / 2 && rr >= 0) {
    const double r = hypot((double)fy, (double)fx);
    const double ss = r * (double)p.ds;
    const double c = ctf_ast(x, (y + l / 2) % l, l, l, p, 2);
    const double signal = exp(p.bf * ss * ss + p.bf2 * ss + p.bf3) / ((double)p.kk + 1);
    const double Ncurve = exp(p.a * ss * ss + p.b * ss + p.b2) / signal;
    f *= c * sqrt(1 / (Ncurve + (double)p.kk * c * c));
  }
#pragma unroll
  for (int k = 0; k < kWeightT; ++k)
    if (k < nt) Fq[k * L2] = make_float2((float)(v[k].x * f), (float)(v[k].y * f));
}

// apply_mask (k:225-241) on the real part.
__global__ void k_mask_planes(float2* __restrict__ T, int n, int l, float rin, float rout, float ehw) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long L2 = (long long)l * l;
  if (i >= L2 * n) return;
  const int q = (int)(i % L2);
  const int x = q % l, y = q / l;
  const double r = hypot((double)(x - l / 2), (double)(y - l / 2));
  if (r > rout)
    T[i].x = 0.f;
  else if (r >= rin)
    T[i].x = (float)(T[i].x * (0.5 * cos(kRefPi * (r - rin) / (2.0 * ehw)) + 0.5));
}

// compute_area_sum_ofSQR (k:319-382) + host mean (src/norm.cu:334-342); full [y][x] planes.
// Partial sums over a slice of the region; grid (chunks, n), then k_area_power_finish.

__global__ void k_area_power_full(const float2* __restrict__ F, int l, double2* __restrict__ part) {
  const int t = blockIdx.y, ch = blockIdx.x;
  const float2* P = F + (long long)t * l * l;
  const int hx = l / 2 + 1;
  const int per = (l * hx + (int)gridDim.x - 1) / (int)gridDim.x;
  const int p0 = ch * per, p1 = min(l * hx, p0 + per);
  double s = 0, c = 0;
  for (int p = p0 + threadIdx.x; p < p1; p += blockDim.x) {
    const int y = p / hx, x = p % hx;
    const int r = rbin(min(x, l - x), min(y, l - y));
    if (r < l / 2 && r >= 0) {
      const float2 v = P[(long long)y * l + x];
      s += (double)v.x * v.x + (double)v.y * v.y;
      c += 1;
    }
  }
  s = block_sum(s);
  c = block_sum(c);
  if (threadIdx.x == 0) part[(long long)t * gridDim.x + ch] = make_double2(s, c);
}

// Same over half spectra stored column-major [kx 0..l/2][ky].
__global__ void k_area_power_half(const float2* __restrict__ F, int l, double2* __restrict__ part) {
  const int t = blockIdx.y, ch = blockIdx.x;
  const int hx = l / 2 + 1;
  const float2* P = F + (long long)t * hx * l;
  const int per = (l * hx + (int)gridDim.x - 1) / (int)gridDim.x;
  const int p0 = ch * per, p1 = min(l * hx, p0 + per);
  double s = 0, c = 0;
  for (int p = p0 + threadIdx.x; p < p1; p += blockDim.x) {
    const int x = p / l, y = p % l;
    const int r = rbin(min(x, l - x), min(y, l - y));
    if (r < l / 2 && r >= 0) {
      const float2 v = P[p];
      s += (double)v.x * v.x + (double)v.y * v.y;
      c += 1;
    }
  }
  s = block_sum(s);
  c = block_sum(c);
  if (threadIdx.x == 0) part[(long long)t * gridDim.x + ch] = make_double2(s, c);
}

// m = sqrt(sum / count^2) per plane from the chunk partials (host mean of src/norm.cu:334-342).
__global__ void k_area_power_finish(const double2* __restrict__ part, int n, int chunks, double* __restrict__ m) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  double s = 0, c = 0;
  for (int ch = 0; ch < chunks; ++ch) {
    s += part[(long long)t * chunks + ch].x;
    c += part[(long long)t * chunks + ch].y;
  }
  m[t] = sqrt(s / (c * c));
}

// normalize (k:445-458): F /= m (when m != 0), times an extra factor.
__global__ void k_scale_planes(float2* __restrict__ F, int n, long long plane, const double* __restrict__ m,
                               float extra) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= plane * n) return;
  const int t = (int)(i / plane);
  const double mt = m[t];
  const double f = (mt != 0 ? 1.0 / mt : 1.0) * (double)extra;
  F[i] = make_float2((float)(F[i].x * f), (float)(F[i].y * f));
}

__global__ void k_extract_real(const float2* __restrict__ F, float* __restrict__ out, long long count) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) out[i] = F[i].x;
}

// ---------------------------------------------------------------- whole frame
__global__ void k_real_to_complex(const float* __restrict__ in, float2* __restrict__ out, long long count,
                                  float sign) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) out[i] = make_float2(sign * in[i], 0.f);
}

// do_phase_flip (k:828-839) on nf frames of nx x ny.
__global__ void k_phase_flip(float2* __restrict__ F, int nx, int ny, int nf, DevParams p) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long N = (long long)nx * ny;
  if (i >= N * nf) return;
  const int q = (int)(i % N);
  const int x = q % nx, y = q / nx;
  const float v = (float)ctf_ast(x, (y + ny / 2) % ny, nx, ny, p, 0);
  F[i] = make_float2(F[i].x * v, F[i].y * v);
}

// SQRSum_by_circle mode 1 (k:94-126) over whole frames: grid (groups, frames), workgroup g of
// frame f takes a slice of its rows (its waves interleaved), partials [f][g][bin];
// k_radial_power_finish sums them in order.
__global__ void k_radial_power_image(const float2* __restrict__ F, int nx, int ny, double* __restrict__ part) {
  extern __shared__ double hist[];
  const int nbins = max(nx, ny) / 2;
  const int waves = blockDim.x / 64, wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  for (int i = threadIdx.x; i < waves * nbins; i += blockDim.x) hist[i] = 0.0;
  __syncthreads();
  const float2* P = F + (long long)blockIdx.y * nx * ny;
  const int per = (ny + gridDim.x - 1) / gridDim.x;
  const int y0 = blockIdx.x * per, y1 = min(ny, y0 + per);
  for (int y = y0 + wave; y < y1; y += waves)
    radial_row(P + (long long)y * nx, nx / 2 + 1, min(y, ny - y), nbins, hist + wave * nbins, lane);
  __syncthreads();
  radial_merge(hist, waves, nbins, part + ((long long)blockIdx.y * gridDim.x + blockIdx.x) * nbins);
}

// whiten_filter_weight_Img (k:151-202) on nf frames; ra holds max(nx, ny) / 2 bins per frame.
__global__ void k_whiten_filter_image(float2* __restrict__ F, int nx, int ny, int nf, const double* __restrict__ ra,
                                      const double* __restrict__ rb, DevParams p) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long N = (long long)nx * ny;
  if (i >= N * nf) return;
  const int fr = (int)(i / N), q = (int)(i - (long long)fr * N);
  const int x = q % nx, y = q / nx;
  const int dxi = min(x, nx - x), dyi = min(y, ny - y);
  const double dx = dxi, dy = dyi;
  const int r = rbin(dxi, dyi);
  const double ss = sqrt((dx * dx / (double)(nx * nx) + dy * dy / (double)(ny * ny)) /
                         ((double)p.apix * (double)p.apix));
  const int l = max(nx, ny);
  double f = 1.0, v = 0, signal = 0, Ncurve = 0;
  if (r < l / 2 && r >= 0) {
    v = ctf_ast(x, (y + ny / 2) % ny, nx, ny, p, 2);
    signal = exp(p.bf * ss * ss + p.bf2 * ss + p.bf3);
    Ncurve = exp(p.a * ss * ss + p.b * ss + p.b2);
    f *= sqrt((signal * v * v + Ncurve) / signal) / sqrt(ra[(long long)fr * (l / 2) + r] / rb[r]);
    if (r > p.hi_thr) f *= exp(-100 * ss * ss);
  }
  f *= band_image(r, p.hr_i, p.lr_i);
  if (r < l / 2 && r >= 0) {
    signal /= ((double)p.kk + 1);
    Ncurve /= signal;
    f *= sqrt(1 / (Ncurve + (double)p.kk * v * v));
  }
  F[i] = make_float2((float)(F[i].x * f), (float)(F[i].y * f));
}

__global__ void k_zero_dc(float2* F) {
  if (threadIdx.x == 0 && blockIdx.x == 0) F[0] = make_float2(0.f, 0.f);
}

// Tile copies of TileImages (src/utils/tileimages.cpp:66-71) with optional inversion, as the
// real parts of complex tiles out[k] (k < ntiles) whose upper-left corners are ul[2k], ul[2k+1].
__global__ void k_extract_tiles(const float* __restrict__ img, int W, const int* __restrict__ ul, int tile,
                                int ntiles, float sign, float2* __restrict__ out) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long T2 = (long long)tile * tile;
  if (i >= T2 * ntiles) return;
  const int k = (int)(i / T2), q = (int)(i - (long long)k * T2);
  const int r = q / tile, c = q % tile;
  out[i] = make_float2(sign * img[(long long)(ul[2 * k + 1] + r) * W + ul[2 * k] + c], 0.f);
}

// ---------------------------------------------------------------- blocks
// split_IMG (k:605-643): block b = (i, j), i = b % bx, at (offx[i], offy[j]); zero outside.
__global__ void k_split_blocks(const float* __restrict__ img, int nx, int ny, int l,
                               const int* __restrict__ offx, const int* __restrict__ offy, int bx,
                               int nb, float2* __restrict__ out) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long L2 = (long long)l * l;
  if (i >= L2 * nb) return;
  const int b = (int)(i / L2);
  const int p = (int)(i - (long long)b * L2);
  const int x = p % l, y = p / l;
  const int X = offx[b % bx] + x, Y = offy[b / bx] + y;
  const float v = (X < nx && Y < ny) ? img[(long long)Y * nx + X] : 0.f;
  out[i] = make_float2(v, 0.f);
}

// split_IMG of every tile: the real parts of ntiles complex tile x tile frames into
// out[k * nbk^2 + b] (block b = (i, j), i = b % nbk, at (off[i], off[j]); zero outside).
__global__ void k_split_tile_blocks(const float2* __restrict__ F, int tile, int l, const int* __restrict__ off,
                                    int nbk, int ntiles, float2* __restrict__ out) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long L2 = (long long)l * l;
  const int nbb = nbk * nbk;
  if (i >= L2 * nbb * ntiles) return;
  const int g = (int)(i / L2);
  const int p = (int)(i - (long long)g * L2);
  const int k = g / nbb, b = g - k * nbb;
  const int X = off[b % nbk] + p % l, Y = off[b / nbk] + p / l;
  const float v = (X < tile && Y < tile) ? F[((long long)k * tile + Y) * tile + X].x : 0.f;
  out[i] = make_float2(v, 0.f);
}

// ---------------------------------------------------------------- rotation
using dev::Bil;
using dev::bil_coords;

// Rotate + mask one template row per block; per-row double partial sums for mean/var.
// RotateTemplate (src/norm.cu:479-521) for the template pair (2p, 2p+1): rotate_subIMG +
// apply_mask (k:534-603, 225-241) of both templates into one complex plane (template 2p in the
// real part, 2p+1 in the imaginary part; their joint forward transform gives both spectra),
// plus compute_sum_sqr's double sums (k:384-443) per template and output tile.  The sample
// geometry and the mask factor are computed once for both templates.
__global__ void __launch_bounds__(256)
    k_rotate_mask_pair(const float* __restrict__ tp, int n, int l, float cose, float sine, float rin,
                       float rout, float ehw, float2* __restrict__ out, double* __restrict__ partial) {
  const int g = blockIdx.x, p = blockIdx.y, groups = gridDim.x;
  const int gx = (l + kRotTX - 1) / kRotTX;
  const int ta = 2 * p, tb = 2 * p + 1;
  const bool hb = tb < n;
  const float* A = tp + (long long)ta * l * l;
  const float* B = tp + (long long)(hb ? tb : ta) * l * l;
  // each wavefront covers a 16 x 8 part of the tile as two 8 x 8 squares, so that the bilinear
  // taps of one load instruction fall on a compact source patch at every angle
  static_assert(kRotTX == 32 && kRotTY == 16, "wavefront squares");
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int j = (g / gx) * kRotTY + (wave >> 1) * 8 + (lane >> 3);
  double sa = 0, sa2 = 0, sb = 0, sb2 = 0;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int i = (g % gx) * kRotTX + (wave & 1) * 16 + h * 8 + (lane & 7);
    if (i >= l || j >= l) continue;
    float ra = 0.f, rb = 0.f;
    const float r = hypotf((float)(i - l / 2), (float)(j - l / 2));
    const Bil c = bil_coords(i, j, l, cose, sine);
    if (c.ok && !(r > rout)) {  // outside the mask the sample is not needed: it is zeroed
      const float tt = 1.0f - c.t, uu = 1.0f - c.u;
      ra = A[c.k0] * tt * uu + A[c.k1] * c.t * uu + A[c.k2] * c.t * c.u + A[c.k3] * tt * c.u;
      if (hb) rb = B[c.k0] * tt * uu + B[c.k1] * c.t * uu + B[c.k2] * c.t * c.u + B[c.k3] * tt * c.u;
    }
    if (r > rout) {
      ra = 0.f;
      rb = 0.f;
    } else if (r >= rin) {
      const float f = 0.5f * cosf((float)kRefPi * (r - rin) / (2.0f * ehw)) + 0.5f;
      ra *= f;
      rb *= f;
    }
    out[((long long)p * l + j) * l + i] = make_float2(ra, rb);
    sa += ra;
    sa2 += (double)ra * ra;
    sb += rb;
    sb2 += (double)rb * rb;
  }
  sa = block_sum(sa);
  sa2 = block_sum(sa2);
  sb = block_sum(sb);
  sb2 = block_sum(sb2);
  if (threadIdx.x == 0) {
    partial[((long long)ta * groups + g) * 2 + 0] = sa;
    partial[((long long)ta * groups + g) * 2 + 1] = sa2;
    if (hb) {
      partial[((long long)tb * groups + g) * 2 + 0] = sb;
      partial[((long long)tb * groups + g) * 2 + 1] = sb2;
    }
  }
}

// Host reduction of RotateTemplate (src/norm.cu:498-507): mean and E[x^2]-mean^2 per template
// from its row-group partials.
__global__ void k_rotation_stats(const double* __restrict__ partial, int n, int groups, int l,
                                 float* __restrict__ stats) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  double s = 0, s2 = 0;
  for (int g = 0; g < groups; ++g) {
    s += partial[((long long)t * groups + g) * 2];
    s2 += partial[((long long)t * groups + g) * 2 + 1];
  }
  const double L2 = (double)l * l;
  const double mean = s / L2;
  stats[2 * t] = (float)mean;
  stats[2 * t + 1] = (float)(s2 / L2 - mean * mean);
}

// Splits the joint forward transform Z of a pair of real planes (item 2j in the real part, 2j+1
// in the imaginary part) into the two spectra, A = (Z(k) + conj Z(-k)) / 2 and B = (Z(k) -
// conj Z(-k)) / 2i, and writes the column-major half spectra out[o][kx 0..l/2][ky] times scale.
// Pairs are laid out [j][g] for g < stride, and item i of group g goes to out plane o = i *
// stride + g (stride 1: template pairs, o = t).  With stats, the template standardisation
// substract_by_mean + divided_by_var (k:460-477) is applied in the frequency domain (it is
// affine: the mean changes only the DC term by mean * l^2, the division scales every term), with
// the scale of the reference's forward transform (src/norm.cu:516-519).  One workgroup per 32 x 32
// tile of (ky, kx) and pair; both tiles (k and -k) are staged in LDS, read along kx and written
// along ky.
__global__ void __launch_bounds__(256)
    k_unpack_pair_spectra(const float2* __restrict__ Z, int n, int stride, int l, const float* __restrict__ stats,
                          float scale, float2* __restrict__ out) {
  constexpr int TS = 32;
  __shared__ float2 zk[TS][TS + 1], zm[TS][TS + 1];
  const int hx = l / 2 + 1;
  const int kx0 = blockIdx.x * TS, ky0 = blockIdx.y * TS, p = blockIdx.z;
  const int j = p / stride, g = p - j * stride;
  const float2* P = Z + (long long)p * l * l;
  const int c = threadIdx.x % TS;
  // all loads of a thread issued before the first is used
  constexpr int NR = TS / (256 / TS);
  float2 vk[NR], vm[NR];
#pragma unroll
  for (int q = 0; q < NR; ++q) {
    const int r = threadIdx.x / TS + q * (256 / TS), ky = ky0 + r, kx = kx0 + c;
    if (ky < l && kx < hx) {
      vk[q] = P[(long long)ky * l + kx];
      vm[q] = P[(long long)(ky ? l - ky : 0) * l + (kx ? l - kx : 0)];
    }
  }
#pragma unroll
  for (int q = 0; q < NR; ++q) {
    const int r = threadIdx.x / TS + q * (256 / TS), ky = ky0 + r, kx = kx0 + c;
    if (ky < l && kx < hx) {
      zk[r][c] = vk[q];
      zm[r][c] = vm[q];
    }
  }
  __syncthreads();
  const int b = threadIdx.x % TS, ky = ky0 + b;
  float mean[2] = {0.f, 0.f}, sd[2] = {1.f, 1.f};
  if (stats) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int o = min(2 * j + h, n - 1) * stride + g;
      mean[h] = stats[2 * o];
      const float var = stats[2 * o + 1];
      sd[h] = var != 0 ? sqrtf(var) : 1.f;
    }
  }
  for (int a = threadIdx.x / TS; a < TS; a += 256 / TS) {
    const int kx = kx0 + a;
    if (kx >= hx || ky >= l) continue;
    const float2 u = zk[b][a], w = zm[b][a];
    float2 v[2] = {make_float2(0.5f * (u.x + w.x), 0.5f * (u.y - w.y)),
                   make_float2(0.5f * (u.y + w.y), 0.5f * (w.x - u.x))};
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int i = 2 * j + h;
      if (i >= n) break;
      float2 x = v[h];
      if (kx == 0 && ky == 0 && mean[h] != 0) x.x -= mean[h] * (float)((long long)l * l);
      x.x = x.x / sd[h] * scale;
      x.y = x.y / sd[h] * scale;
      out[((long long)(i * stride + g) * hx + kx) * l + ky] = x;
    }
  }
}

// rotate_subIMG (k:534-603) of real planes, two angles per complex output plane: for the
// pair plane p = j * nb + g, out[p] = rotate(in[g], cs[2j]) + i rotate(in[g], cs[2j + 1]) (the
// imaginary part zero when 2j + 1 = na).  The blocks are real images (split_IMG of a real tile;
// the phase-flip normalisation keeps them real up to rounding), so one forward transform of the
// pair plane gives the spectra of both rotated copies (k_unpack_pair_spectra).  Workgroup: a
// 16 x 16 output tile, each wavefront an 8 x 8 square of it, so that the bilinear taps of a
// wavefront fall on a compact patch of the source at every angle (a 64-pixel output row maps to a
// source line that crosses 64 rows near 90 degrees).
__global__ void __launch_bounds__(256)
    k_rotate_real_pairs(const float* __restrict__ in, int nb, int l, const float2* __restrict__ cs, int na,
                        float2* __restrict__ out) {
  const long long L2 = (long long)l * l;
  const int gx = (l + 15) / 16;
  const int p = blockIdx.y, j = p / nb, g = p - j * nb;
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int x = ((int)blockIdx.x % gx) * 16 + (wave & 1) * 8 + (lane & 7);
  const int y = ((int)blockIdx.x / gx) * 16 + (wave >> 1) * 8 + (lane >> 3);
  if (x >= l || y >= l) return;  // no barrier follows
  const float* P = in + (long long)g * L2;
  Bil b[2];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int a = min(2 * j + h, na - 1);
    const float2 c = cs[a];
    b[h] = bil_coords(x, y, l, c.x, c.y);
    if (2 * j + h >= na) b[h].ok = false;
  }
  float v[2][4];
#pragma unroll
  for (int h = 0; h < 2; ++h)
    if (b[h].ok) {
      v[h][0] = P[b[h].k0];
      v[h][1] = P[b[h].k1];
      v[h][2] = P[b[h].k2];
      v[h][3] = P[b[h].k3];
    }
  float res[2] = {0.f, 0.f};
#pragma unroll
  for (int h = 0; h < 2; ++h)
    if (b[h].ok) {
      const float tt = 1.0f - b[h].t, uu = 1.0f - b[h].u;
      const float w0 = tt * uu, w1 = b[h].t * uu, w2 = b[h].t * b[h].u, w3 = tt * b[h].u;
      res[h] = v[h][0] * w0 + v[h][1] * w1 + v[h][2] * w2 + v[h][3] * w3;
    }
  st_stream(out + p * L2 + (long long)y * l + x, make_float2(res[0], res[1]));  // read once, by the transform
}

}  // namespace

// ---------------------------------------------------------------- launchers
int plane_reduction_chunks(int n) { return n >= 64 ? 16 : std::min(128, std::max(16, 1024 / std::max(n, 1))); }

void launch_pad_templates(const float* raw, int n, int w, int h, int l, float2* out, hipStream_t s) {
  k_pad_templates<<<nblk((long long)n * l * l), BS, 0, s>>>(raw, n, w, h, l, out);
  GS_CHECK_LAUNCH();
}
void launch_template_moments(const float2* T, int n, int l, double* part, double* mom, hipStream_t s) {
  const int ch = plane_reduction_chunks(n);
  k_template_moments<<<dim3(ch, n), BS, 0, s>>>(T, l, part);
  GS_CHECK_LAUNCH();
  k_template_moments_finish<<<nblk(n), BS, 0, s>>>(part, n, ch, mom);
  GS_CHECK_LAUNCH();
}
void launch_template_standardise(float2* T, int n, int l, const double* mom, hipStream_t s) {
  k_template_standardise<<<nblk((long long)n * l * l), BS, 0, s>>>(T, n, l, mom);
  GS_CHECK_LAUNCH();
}
void launch_radial_power_planes(const float2* F, int n, int l, double* part, double* ra, hipStream_t s) {
  const int nbins = l / 2;
  GS_REQUIRE(4 * nbins * sizeof(double) <= 64 * 1024, kUnsupported, "window too large for radial histogram");
  const int ch = plane_reduction_chunks(n);
  k_radial_power_planes<<<dim3(ch, n), BS, 4 * sizeof(double) * nbins, s>>>(F, l, part);
  GS_CHECK_LAUNCH();
  k_radial_power_finish<<<nblk((long long)n * nbins * 64), BS, 0, s>>>(part, n, nbins, ch, ra);
  GS_CHECK_LAUNCH();
}
void launch_whiten_planes(float2* F, int n, int l, const double* ra, const double* rb, hipStream_t s) {
  k_whiten_planes<<<nblk((long long)n * l * l), BS, 0, s>>>(F, n, l, ra, rb);
  GS_CHECK_LAUNCH();
}
void launch_weight_planes(float2* F, int n, int l, const DevParams& p, hipStream_t s) {
  GS_REQUIRE((n + kWeightT - 1) / kWeightT < 65536, kInvalid, "template count out of range");
  k_weight_planes<<<dim3(nblk((long long)l * l), (unsigned)((n + kWeightT - 1) / kWeightT)), BS, 0, s>>>(F, n, l, p);
  GS_CHECK_LAUNCH();
}
void launch_mask_planes(float2* T, int n, int l, float d_m, float ehw, hipStream_t s) {
  const float dm = 1.5f * d_m;  // apply_mask: d_m = 1.5 * d_m (k:233)
  const float rin = dm / 2, rout = dm / 2 + 2 * ehw;
  k_mask_planes<<<nblk((long long)n * l * l), BS, 0, s>>>(T, n, l, rin, rout, ehw);
  GS_CHECK_LAUNCH();
}
void launch_area_power_full(const float2* F, int n, int l, double2* part, double* m, hipStream_t s) {
  const int ch = plane_reduction_chunks(n);
  k_area_power_full<<<dim3(ch, n), BS, 0, s>>>(F, l, part);
  GS_CHECK_LAUNCH();
  k_area_power_finish<<<nblk(n), BS, 0, s>>>(part, n, ch, m);
  GS_CHECK_LAUNCH();
}
void launch_area_power_half(const float2* F, int n, int l, double2* part, double* m, hipStream_t s) {
  const int ch = plane_reduction_chunks(n);
  k_area_power_half<<<dim3(ch, n), BS, 0, s>>>(F, l, part);
  GS_CHECK_LAUNCH();
  k_area_power_finish<<<nblk(n), BS, 0, s>>>(part, n, ch, m);
  GS_CHECK_LAUNCH();
}
void launch_scale_planes(float2* F, int n, long long plane, const double* m, float extra, hipStream_t s) {
  k_scale_planes<<<nblk(plane * n), BS, 0, s>>>(F, n, plane, m, extra);
  GS_CHECK_LAUNCH();
}
void launch_extract_real(const float2* F, float* out, long long count, hipStream_t s) {
  k_extract_real<<<nblk(count), BS, 0, s>>>(F, out, count);
  GS_CHECK_LAUNCH();
}
void launch_real_to_complex(const float* in, float2* out, long long count, float sign, hipStream_t s) {
  k_real_to_complex<<<nblk(count), BS, 0, s>>>(in, out, count, sign);
  GS_CHECK_LAUNCH();
}
void launch_phase_flip(float2* F, int nx, int ny, int nf, const DevParams& p, hipStream_t s) {
  k_phase_flip<<<nblk((long long)nx * ny * nf), BS, 0, s>>>(F, nx, ny, nf, p);
  GS_CHECK_LAUNCH();
}
int radial_image_groups(int nf) { return nf <= 1 ? kImageGroups : std::max(16, kImageGroups / nf); }
void launch_radial_power_image(const float2* F, int nx, int ny, int nf, double* part, double* ra, hipStream_t s) {
  const int nbins = std::max(nx, ny) / 2, groups = radial_image_groups(nf);
  // wave-private histograms within 64 KB of LDS per workgroup
  const int waves = (int)std::min<size_t>(4, (64 * 1024) / (sizeof(double) * std::max(1, nbins)));
  GS_REQUIRE(waves >= 1, kUnsupported, "frame too large for radial histogram");
  GS_REQUIRE(nf >= 1 && nf < 65536, kInvalid, "frame count out of range");
  k_radial_power_image<<<dim3(groups, nf), 64 * waves, sizeof(double) * nbins * waves, s>>>(F, nx, ny, part);
  GS_CHECK_LAUNCH();
  k_radial_power_finish<<<nblk((long long)nf * nbins * 64), BS, 0, s>>>(part, nf, nbins, groups, ra);
  GS_CHECK_LAUNCH();
}
void launch_whiten_filter_image(float2* F, int nx, int ny, int nf, const double* ra, const double* rb,
                                const DevParams& p, hipStream_t s) {
  k_whiten_filter_image<<<nblk((long long)nx * ny * nf), BS, 0, s>>>(F, nx, ny, nf, ra, rb, p);
  GS_CHECK_LAUNCH();
}
void launch_zero_dc(float2* F, hipStream_t s) {
  k_zero_dc<<<1, 64, 0, s>>>(F);
  GS_CHECK_LAUNCH();
}
void launch_extract_tiles(const float* img, int W, const int* ul, int tile, int ntiles, float sign, float2* out,
                          hipStream_t s) {
  k_extract_tiles<<<nblk((long long)tile * tile * ntiles), BS, 0, s>>>(img, W, ul, tile, ntiles, sign, out);
  GS_CHECK_LAUNCH();
}
void launch_split_blocks(const float* img, int nx, int ny, int l, const int* offx, const int* offy,
                         int bx, int nb, float2* out, hipStream_t s) {
  k_split_blocks<<<nblk((long long)nb * l * l), BS, 0, s>>>(img, nx, ny, l, offx, offy, bx, nb, out);
  GS_CHECK_LAUNCH();
}
void launch_split_tile_blocks(const float2* F, int tile, int l, const int* off, int nbk, int ntiles, float2* out,
                              hipStream_t s) {
  k_split_tile_blocks<<<nblk((long long)nbk * nbk * ntiles * l * l), BS, 0, s>>>(F, tile, l, off, nbk, ntiles, out);
  GS_CHECK_LAUNCH();
}
int rotation_groups(int l) { return ((l + kRotTX - 1) / kRotTX) * ((l + kRotTY - 1) / kRotTY); }
void launch_rotate_mask_pairs(const float* tp, int n, int l, float cose, float sine, float d_m, float ehw,
                              float2* out, double* partial, hipStream_t s) {
  const float dm = 1.5f * d_m;
  const float rin = dm / 2, rout = dm / 2 + 2 * ehw;
  GS_REQUIRE((n + 1) / 2 < 65536, kInvalid, "template count out of range");
  k_rotate_mask_pair<<<dim3(rotation_groups(l), (n + 1) / 2), 256, 0, s>>>(tp, n, l, cose, sine, rin, rout,
                                                                           ehw, out, partial);
  GS_CHECK_LAUNCH();
}
void launch_rotation_stats(const double* partial, int n, int l, float* stats, hipStream_t s) {
  k_rotation_stats<<<nblk(n, 64), 64, 0, s>>>(partial, n, rotation_groups(l), l, stats);
  GS_CHECK_LAUNCH();
}
void launch_unpack_pair_spectra(const float2* Z, int n, int stride, int l, const float* stats, float scale,
                                float2* out, hipStream_t s) {
  const long long pairs = (long long)((n + 1) / 2) * stride;
  GS_REQUIRE(stride >= 1 && pairs < 65536, kInvalid, "plane count out of range");
  const dim3 grid((unsigned)((l / 2 + 1 + 31) / 32), (unsigned)((l + 31) / 32), (unsigned)pairs);
  k_unpack_pair_spectra<<<grid, 256, 0, s>>>(Z, n, stride, l, stats, scale, out);
  GS_CHECK_LAUNCH();
}
void launch_rotate_real_pairs(const float* in, int nb, int l, const float2* cs, int na, float2* out,
                              hipStream_t s) {
  const long long pairs = (long long)nb * ((na + 1) / 2);
  GS_REQUIRE(pairs >= 1 && pairs < 65536, kInvalid, "plane count out of range");
  const int gx = (l + 15) / 16;
  k_rotate_real_pairs<<<dim3((unsigned)(gx * gx), (unsigned)pairs), 256, 0, s>>>(in, nb, l, cs, na, out);
  GS_CHECK_LAUNCH();
}

}  // namespace gs
