// FP64 CPU oracle of the GisSPA search (TEST INFRASTRUCTURE ONLY; see gisspa_oracle.h).
//
// Every function restates one reference routine in double precision.  Parameter-level scalars
// that drive integer decisions (band-pass radii, rotation sample coordinates) are evaluated in
// float exactly as the HIP path evaluates them, so both sides take the same branches; all
// data arithmetic is FP64.  Radial bins are computed exactly from integer squared radii.
#include "gisspa_oracle.h"

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <memory>
#include <numeric>
#include <stdexcept>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "fft64.hpp"
#include "oracle_impl.hpp"

namespace orc_impl {

P make_params(const orc_config* c, const double* ctf3) {
  P p{};
  p.apix = c->angpix; p.kk = c->kk; p.energy = c->energy; p.cs = c->cs;
  p.highres = c->highres; p.lowres = c->lowres; p.d_m = c->diameter; p.thresh = c->threshold;
  const double def = ctf3[0], dfdiff = ctf3[1], dfang = ctf3[2];
  p.defocus = -def;
  p.dfdiff = dfdiff;
  p.dfang = dfang;
  p.dfu = -def + dfdiff;  // src/norm.cu:188
  p.dfv = -def - dfdiff;  // src/norm.cu:189
  const double E = p.energy;
  p.lambda = 12.2639 / std::sqrt(E * 1000.0 + 0.97845 * E * E);  // src/norm.cu:190
  return p;
}

// CTF_AST (src/kernels.cu:290-317) in FP64: the signed CTF value and, when dv is given, its
// derivative with respect to a defocus offset in Angstrom added to the micrograph defocus
// (dfu and dfv both move by -offset*1e-4 um, so d gamma / d offset = PI * lambda * s^2).
double ctf_signed(int x1, int y1, int nx, int ny, const P& p, double* dv) {
  y1 = x1 > nx / 2 ? ny - y1 : y1;
  x1 = std::min(x1, nx - x1);
  const double dx = std::min(x1, nx - x1);
  const double dy = y1 - ny / 2;
  const double ss = (dx * dx / (double)(nx * nx) + dy * dy / (double)(ny * ny)) /
                    ((double)p.apix * (double)p.apix);
  const double ag = std::atan2((double)(y1 - ny / 2), (double)x1);
  const double df_ast =
      0.5 * (p.dfu + p.dfv + 2 * p.dfdiff * std::cos(2 * (p.dfang * PI / 180 - ag)));
  const double lam = p.lambda;
  const double gamma =
      -2 * PI * ((double)p.cs * 2.5e6 * lam * lam * lam * ss * ss + df_ast * 5000.0 * lam * ss);
  const double A = p.ph.ampconst, B = std::sqrt(1.0 - A * A);
  if (dv) *dv = (B * std::cos(gamma) - A * std::sin(gamma)) * (PI * lam * ss);
  return B * std::sin(gamma) + A * std::cos(gamma);
}

double ctf_ast(int x1, int y1, int nx, int ny, const P& p, int mode) {
  const double v = ctf_signed(x1, y1, nx, ny, p, nullptr);
  if (mode == 0) return v > 0 ? 1.0 : -1.0;
  if (mode == 2) return std::fabs(v);
  return v;
}

Band band(int l, const P& p) {
  Band b;
  b.hr = (float)l * p.apix / p.highres;
  b.lr = (float)l * p.apix / p.lowres;
  return b;
}

// Template band-pass factor of apply_weighting_function (k:263-275).
double bandpass_template(int r, const Band& b) {
  if (r < b.hr && r >= 0) return 1.0;
  if (r >= b.hr && r < b.hr + 8) return 0.5 * std::cos(PI * (r - (double)b.hr) / (2 * 8)) + 0.5;
  if (r >= (b.lr - 8) && r < b.lr && r >= 0)
    return 0.5 * std::cos(PI * ((double)b.lr - r) / (2 * 8)) + 0.5;
  return 0.0;
}

// Image band-pass factor of whiten_filter_weight_Img (k:183-189).
double bandpass_image(int r, const Band& b) {
  if (r < b.hr && r >= b.lr) return 1.0;
  if (r >= b.hr && r < b.hr + 8) return 0.5 * std::cos(PI * (r - (double)b.hr) / (2 * 8)) + 0.5;
  if (r >= (b.lr - 8) && r < b.lr && r >= 0)
    return 0.5 * std::cos(PI * ((double)b.lr - r) / (2 * 8)) + 0.5;
  return 0.0;
}

// Zero-padded centring of LoadTemplate (src/norm.cu:132-154).
void pad_template(const float* t, int w, int h, int l, cd* out) {
  std::fill(out, out + (size_t)l * l, cd(0));
  const int sx = (l - w) / 2, sy = (l - h) / 2;
  for (int j = 0; j < h; ++j)
    for (int i = 0; i < w; ++i) out[(size_t)(sy + j) * l + (sx + i)] = cd(t[(size_t)j * w + i], 0);
}

// Edge-mean / box-sigma standardisation: generate_mask, multiCount_dot, UpdateSigma,
// scale_each (k:4-90) with the host reductions of src/norm.cu:194-280.
void standardise_template(cd* t, int l) {
  const float r = l / 2.f - 2.f;
  float up = 0, low = 0;
  if (r > 1) {
    up = (r + 1) * (r + 1);
    low = (r - 1) * (r - 1);
  }
  double ring = 0, cnt = 0, s = 0, s2 = 0;
  for (int y = 0; y < l; ++y)
    for (int x = 0; x < l; ++x) {
      const float rr = (float)((x - l / 2) * (x - l / 2) + (y - l / 2) * (y - l / 2));
      const double v = t[(size_t)y * l + x].real();
      if (rr >= low && rr <= up) {
        ring += v;
        cnt += 1;
      }
      s += v;
      s2 += v * v;
    }
  const double em = cnt != 0 ? ring / cnt : ring;
  const double mean = s / ((double)l * l);
  double sigma = std::sqrt(s2 / ((double)l * l) - mean * mean);
  if (sigma < 0 || !std::isfinite(sigma)) sigma = 0;
  if (sigma < EPS && sigma > -EPS) return;
  for (size_t i = 0; i < (size_t)l * l; ++i) t[i] = cd((t[i].real() - em) / sigma, t[i].imag());
}

// SQRSum_by_circle (k:94-126) + whiten_Tmp (k:128-149) on one l x l spectrum.
void whiten_template(cd* F, int l) {
  std::vector<double> ra(l / 2, 0.0), rb(l / 2, 0.0);
  for (int y = 0; y < l; ++y)
    for (int x = 0; x <= l / 2; ++x) {
      const int r = rbin(std::min(x, l - x), std::min(y, l - y));
      if (r >= 0 && r < l / 2) {
        ra[r] += std::norm(F[(size_t)y * l + x]);
        rb[r] += 1;
      }
    }
  for (int y = 0; y < l; ++y)
    for (int x = 0; x < l; ++x) {
      const int r = rbin(std::min(x, l - x), std::min(y, l - y));
      if (r >= 0 && r < l / 2) F[(size_t)y * l + x] /= std::sqrt(ra[r] / rb[r]);
    }
}

// apply_weighting_function (k:243-288): band-pass then |CTF| Wiener-like weight.
void weight_template(cd* F, int l, const P& p) {
  const Band b = band(l, p);
  const double ds = (double)(1.0f / (p.apix * (float)l));  // para.ds, src/norm.cu:191
  for (int y = 0; y < l; ++y)
    for (int x = 0; x < l; ++x) {
      const int fx = std::min(x, l - x), fy = std::min(y, l - y);
      const double r = std::hypot((double)fy, (double)fx);
      const int rr = rbin(fx, fy);
      double f = bandpass_template(rr, b);
      if (rr < l / 2 && rr >= 0) {
        const double ss = r * ds;
        const double v = ctf_ast(x, (y + l / 2) % l, l, l, p, 2);
        const double signal =
            std::exp(p.ph.bf * ss * ss + p.ph.bf2 * ss + p.ph.bf3) / ((double)p.kk + 1);
        const double Ncurve = std::exp(p.ph.a * ss * ss + p.ph.b * ss + p.ph.b2) / signal;
        f *= v * std::sqrt(1 / (Ncurve + (double)p.kk * v * v));
      }
      F[(size_t)y * l + x] *= f;
    }
}

// compute_area_sum_ofSQR (k:319-382) + host mean (src/norm.cu:334-342) + normalize (k:445-458).
void spectral_normalise(cd* F, int l) {
  double s = 0, cnt = 0;
  for (int y = 0; y < l; ++y)
    for (int x = 0; x <= l / 2; ++x) {
      const int r = rbin(std::min(x, l - x), std::min(y, l - y));
      if (r < l / 2 && r >= 0) {
        s += std::norm(F[(size_t)y * l + x]);
        cnt += 1;
      }
    }
  const double m = std::sqrt(s / (cnt * cnt));
  if (m != 0)
    for (size_t i = 0; i < (size_t)l * l; ++i) F[i] /= m;
}

// apply_mask (k:225-241): soft disc on the real part; mask_factor is its value at (x, y).
double mask_factor(int x, int y, int l, double d_m_cfg, double ehw) {
  const float d_m = 1.5f * (float)d_m_cfg;
  const double rout = (double)(d_m / 2 + 2 * (float)ehw), rin = (double)(d_m / 2);
  const double r = std::hypot((double)(x - l / 2), (double)(y - l / 2));
  if (r > rout) return 0.0;
  if (r >= rin) return 0.5 * std::cos(PI * (r - rin) / (2 * ehw)) + 0.5;
  return 1.0;
}

void apply_mask(cd* t, int l, double d_m_cfg, double ehw) {
  for (int y = 0; y < l; ++y)
    for (int x = 0; x < l; ++x) {
      cd& v = t[(size_t)y * l + x];
      v = cd(v.real() * mask_factor(x, y, l, d_m_cfg, ehw), v.imag());
    }
}

void scale_all(cd* x, size_t n, double s) {
  for (size_t i = 0; i < n; ++i) x[i] *= s;
}

void rotate_plane(const cd* in, cd* out, int l, float theta) {
  float cose, sine;
  rot_cs(theta, cose, sine);
  for (int j = 0; j < l; ++j)
    for (int i = 0; i < l; ++i) {
      Bil b;
      cd r = 0;
      if (bilinear_coords(i, j, l, cose, sine, b)) {
        const double t = b.t, u = b.u, tt = 1.0 - t, uu = 1.0 - u;
        r = in[b.k0] * (tt * uu) + in[b.k1] * (t * uu) + in[b.k3] * (tt * u) + in[b.k2] * (t * u);
      }
      out[(size_t)j * l + i] = r;
    }
}

std::vector<int> block_offsets(int length, int l, int overlap) {  // src/norm.cu:68-81
  std::vector<int> o{0};
  int off = 0;
  while (off + l < length) {
    off += l - overlap;
    if (off + l >= length) off = length - l;
    o.push_back(off);
  }
  return o;
}

std::vector<float> angle_list(float phistep) {  // src/norm.cu:642
  std::vector<float> a;
  for (float e = 0.0f; e < 360.0f; e += phistep) a.push_back(e);
  return a;
}

// Template preprocessing, norm (src/norm.cu:194-354) or nonorm (src/nonorm.cu:178-344).
void prep_templates(const orc_config* c, const P& p, const float* tmpl, int n, int w, int h,
                    cd* out) {
  const int l = c->window_size;
  const size_t L2 = (size_t)l * l;
#pragma omp parallel for schedule(dynamic)
  for (int t = 0; t < n; ++t) {
    cd* T = out + (size_t)t * L2;
    pad_template(tmpl + (size_t)t * w * h, w, h, l, T);
    standardise_template(T, l);
    orc::fft2(T, l, l, -1);
    scale_all(T, L2, 1.0 / (double)L2);
    whiten_template(T, l);
    if (!c->norm_type) {  // src/nonorm.cu:291-299
      orc::fft2(T, l, l, +1);
      apply_mask(T, l, p.d_m, p.ph.edge_half_width);
      orc::fft2(T, l, l, -1);
      scale_all(T, L2, 1.0 / (double)L2);
    }
    weight_template(T, l, p);
    spectral_normalise(T, l);
    if (c->norm_type) orc::fft2(T, l, l, +1);  // src/norm.cu:350
  }
}

// do_phase_flip (k:828-839) + SQRSum_by_circle mode 1 (k:94-126) + whiten_filter_weight_Img
// (k:151-202) on the scaled whole-frame spectrum F (ny x nx, row-major), in place.
void filter_image_spectrum(const P& p, cd* F, int nx, int ny) {
  const int l = std::max(nx, ny);
  // do_phase_flip (k:828-839)
#pragma omp parallel for schedule(static)
  for (int y = 0; y < ny; ++y)
    for (int x = 0; x < nx; ++x)
      F[(size_t)y * nx + x] *= ctf_ast(x, (y + ny / 2) % ny, nx, ny, p, 0);
  // SQRSum_by_circle mode 1 (k:94-126): radial power over the half plane x <= nx/2
  std::vector<double> ra(l / 2, 0.0), rb(l / 2, 0.0);
  for (int y = 0; y < ny; ++y)
    for (int x = 0; x <= nx / 2; ++x) {
      const int r = rbin(std::min(x, nx - x), std::min(y, ny - y));
      if (r < l / 2 && r >= 0) {
        ra[r] += std::norm(F[(size_t)y * nx + x]);
        rb[r] += 1;
      }
    }
  // whiten_filter_weight_Img (k:151-202)
  const Band b = band(l, p);
  const float hi_thr = (float)l * p.apix / 6;
#pragma omp parallel for schedule(static)
  for (int y = 0; y < ny; ++y)
    for (int x = 0; x < nx; ++x) {
      const double dx = std::min(x, nx - x), dy = std::min(y, ny - y);
      const int r = rbin((long long)dx, (long long)dy);
      const double ss =
          std::sqrt((dx * dx / (double)(nx * nx) + dy * dy / (double)(ny * ny)) /
                    ((double)p.apix * p.apix));
      double f = 1.0, v = 0, signal = 0, Ncurve = 0;
      if (r < l / 2 && r >= 0) {
        v = ctf_ast(x, (y + ny / 2) % ny, nx, ny, p, 2);
        signal = std::exp(p.ph.bf * ss * ss + p.ph.bf2 * ss + p.ph.bf3);
        Ncurve = std::exp(p.ph.a * ss * ss + p.ph.b * ss + p.ph.b2);
        f *= std::sqrt((signal * v * v + Ncurve) / signal) / std::sqrt(ra[r] / rb[r]);
        if (r > hi_thr) f *= std::exp(-100 * ss * ss);
      }
      f *= bandpass_image(r, b);
      if (r < l / 2 && r >= 0) {
        signal /= ((double)p.kk + 1);
        Ncurve /= signal;
        f *= std::sqrt(1 / (Ncurve + (double)p.kk * v * v));
      }
      F[(size_t)y * nx + x] *= f;
    }
}

// Whole-frame preprocessing: src/norm.cu:356-402 (dc_zero=1) / src/nonorm.cu:346-389.
void prep_image(const P& p, const float* img, int nx, int ny, bool dc_zero, double* out) {
  const size_t N = (size_t)nx * ny;
  std::vector<cd> F(N);
  for (size_t i = 0; i < N; ++i) F[i] = cd(img[i], 0);
  orc::fft2(F.data(), ny, nx, -1);
  scale_all(F.data(), N, 1.0 / (double)N);
  filter_image_spectrum(p, F.data(), nx, ny);
  if (dc_zero) F[0] = 0;  // set_0Hz_to_0_at_RI (k:204-223)
  orc::fft2(F.data(), ny, nx, +1);
  for (size_t i = 0; i < N; ++i) out[i] = F[i].real();
}

// Host clipping of LoadImage (src/norm.cu:156-181): float sequential sums, int bounds.
void clip_image(float* img, size_t N, int nx, int ny, int invert) {
  float sum = 0, sum_s2 = 0;
  for (size_t i = 0; i < N; ++i) {
    const float cur = img[i];
    sum += cur / nx / ny;
    sum_s2 += (cur * cur / nx / ny);
  }
  const float avg = sum;
  const float var = std::sqrt(sum_s2 - avg * avg);
  // (int) as x86 converts: a value outside int's range (or NaN) gives INT_MIN, stated here
  // instead of left undefined
  auto to_int = [](float v) { return (v >= -2147483648.f && v < 2147483648.f) ? (int)v : INT_MIN; };
  const int up = to_int(avg + 6 * var);
  const int low = to_int(avg - 6 * var);
  if (sum_s2 > avg * avg)
    for (size_t i = 0; i < N; ++i)
      if (img[i] > up || img[i] < low) img[i] = avg;
  if (invert)
    for (size_t i = 0; i < N; ++i) img[i] = -img[i];
}

// Rotate + mask + standardise + forward FFT of one processed template (RotateTemplate,
// src/norm.cu:479-521).
void rotate_template(const cd* Tp, cd* out, int l, float theta, const P& p) {
  const size_t L2 = (size_t)l * l;
  rotate_plane(Tp, out, l, theta);
  apply_mask(out, l, p.d_m, p.ph.edge_half_width);
  double s = 0, s2 = 0;
  for (size_t i = 0; i < L2; ++i) {
    const double v = out[i].real();
    s += v;
    s2 += v * v;
  }
  const double mean = s / (double)L2;
  const double var = s2 / (double)L2 - mean * mean;
  for (size_t i = 0; i < L2; ++i) {
    double v = out[i].real();
    if (mean != 0) v -= mean;
    if (var != 0) v /= std::sqrt(var);
    out[i] = cd(v, out[i].imag());
  }
  orc::fft2(out, l, l, -1);
  scale_all(out, L2, 1.0 / (double)L2);
}

// rotate_template for two templates with one forward transform: the rotated, masked and
// standardised planes are real (the imaginary rounding residue of the processed templates is
// dropped), so the transform of a + i b gives both spectra through Hermitian symmetry:
// A(k) = (Z(k) + conj Z(-k)) / 2, B(k) = (Z(k) - conj Z(-k)) / 2i.
void rotate_template_pair(const cd* Ta, const cd* Tb, cd* oa, cd* ob, int l, float theta, const P& p) {
  const size_t L2 = (size_t)l * l;
  std::vector<cd> Z(L2);
  double* planes[2];
  std::vector<double> ra(L2), rb(L2);
  planes[0] = ra.data();
  planes[1] = rb.data();
  const cd* src[2] = {Ta, Tb};
  for (int u = 0; u < 2; ++u) {
    rotate_plane(src[u], oa, l, theta);  // oa as scratch
    apply_mask(oa, l, p.d_m, p.ph.edge_half_width);
    double sm = 0, s2 = 0;
    for (size_t i = 0; i < L2; ++i) {
      const double v = oa[i].real();
      sm += v;
      s2 += v * v;
    }
    const double mean = sm / (double)L2;
    const double var = s2 / (double)L2 - mean * mean;
    for (size_t i = 0; i < L2; ++i) {
      double v = oa[i].real();
      if (mean != 0) v -= mean;
      if (var != 0) v /= std::sqrt(var);
      planes[u][i] = v;
    }
  }
  for (size_t i = 0; i < L2; ++i) Z[i] = cd(ra[i], rb[i]);
  orc::fft2(Z.data(), l, l, -1);
  const double sc = 0.5 / (double)L2;
  for (int y = 0; y < l; ++y)
    for (int x = 0; x < l; ++x) {
      const size_t k = (size_t)y * l + x, km = (size_t)((l - y) % l) * l + (l - x) % l;
      const cd z = Z[k], zc = std::conj(Z[km]);
      oa[k] = (z + zc) * sc;
      ob[k] = (z - zc) * cd(0, -1) * sc;
    }
}

// Block spectra (SplitImage, src/norm.cu:404-477): Fourier-space, spectrally normalised.
void split_blocks(const double* img, int nx, int ny, int l, const std::vector<int>& ox,
                  const std::vector<int>& oy, cd* out) {
  const size_t L2 = (size_t)l * l;
  const int bx = (int)ox.size(), by = (int)oy.size();
#pragma omp parallel for schedule(dynamic)
  for (int b = 0; b < bx * by; ++b) {
    cd* B = out + (size_t)b * L2;
    const int i = b % bx, j = b / bx;
    for (int y = 0; y < l; ++y)
      for (int x = 0; x < l; ++x) {
        const int X = ox[i] + x, Y = oy[j] + y;
        B[(size_t)y * l + x] = (X < nx && Y < ny) ? cd(img[(size_t)Y * nx + X], 0) : cd(0);
      }
    orc::fft2(B, l, l, -1);
    scale_all(B, L2, 1.0 / (double)L2);
    spectral_normalise(B, l);
  }
}

void correlate(const cd* I, const cd* T, cd* C, int l) {
  for (int y = 0; y < l; ++y)
    for (int x = 0; x < l; ++x) {
      const size_t k = (size_t)y * l + x;
      C[k] = I[k] * std::conj(T[k]) * checker(x, y, l);
    }
  orc::fft2(C, l, l, +1);
}

// get_peak_and_SUM (k:725-763) over one correlation plane plus the host merge of its chunk
// results (src/nonorm.cu:497-509): the peak is the largest value off the border of width d_m/4,
// which the reference sets to 0 before the max (so the peak is at least 0; first maximum in
// raster order here), and the sums run over the whole plane, border included.
PlaneStats nonorm_plane(const cd* C, int l, float dm4) {
  PlaneStats o{0.0, 0.0, 0.0, 0};
  for (int y = 0; y < l; ++y)
    for (int x = 0; x < l; ++x) {
      const double v = C[(size_t)y * l + x].real();
      o.sum += v;
      o.sum2 += v * v;
      const bool border = x < dm4 || x > l - dm4 || y < dm4 || y > l - dm4;
      if (!border && v > o.peak) {  // strictly greater: first maximum wins
        o.peak = v;
        o.pos = (long long)y * l + x;
      }
    }
  return o;
}

// Score of src/nonorm.cu:511-519: the peak over the standard deviation of the other l^2 - 1
// values (ra, rb: their sum and sum of squares; rc: their count); 0 when that deviation is 0.
double nonorm_score(const PlaneStats& s, int l) {
  const double ra = s.sum - s.peak, rb = s.sum2 - s.peak * s.peak;
  const double rc = (double)l * l - 1;
  const double sd = std::sqrt(rb / rc - (ra / rc) * (ra / rc));
  return sd == 0 ? 0.0 : s.peak / sd;
}

// Centre of src/nonorm.cu:521-524: the peak position of the correlation of the block rotated by
// e, rotated back about (l/2, l/2), plus the block's offset in the tile.
void nonorm_center(long long pos, int bi, int bj, int l, int overlap, float e, float& cx, float& cy) {
  const double ce = std::cos((double)e * PI / 180), se = std::sin((double)e * PI / 180);
  const int px = (int)(pos % l), py = (int)(pos / l);
  cx = (float)(bi * (l - overlap) + (px - l / 2) * ce + (py - l / 2) * se + l / 2);
  cy = (float)(bj * (l - overlap) + (py - l / 2) * ce - (px - l / 2) * se + l / 2);
}

bool pick_less(const orc_pick& a, const orc_pick& b) {  // NMS order: a before b
  if (a.score != b.score) return a.score > b.score;
  if (a.y != b.y) return a.y < b.y;
  if (a.x != b.x) return a.x < b.x;
  if (a.template_id != b.template_id) return a.template_id < b.template_id;
  if (a.angle_index != b.angle_index) return a.angle_index < b.angle_index;
  if (a.block != b.block) return a.block < b.block;
  return a.pixel < b.pixel;
}

int nms(orc_pick* c, int n, float radius) {
  std::stable_sort(c, c + n, pick_less);
  if (radius <= 0) return n;
  const double R2 = (double)radius * radius;
  int kept = 0;
  for (int i = 0; i < n; ++i) {
    bool ok = true;
    for (int k = 0; k < kept && ok; ++k) {
      const double dx = (double)c[i].x - c[k].x, dy = (double)c[i].y - c[k].y;
      if (dx * dx + dy * dy < R2) ok = false;
    }
    if (ok) c[kept++] = c[i];
  }
  return kept;
}

void check_config(const orc_config* c) {
  if (!c) throw std::invalid_argument("null config");
  if (c->window_size <= 0 || c->window_size % 2) throw std::invalid_argument("window_size must be even and > 0");
  if (c->phistep <= 0) throw std::invalid_argument("phistep must be > 0");
}


// ---- norm_type=1 search (src/norm.cu:479-660) over a subset of the blocks ----
// mu_p and sigma_p are per block pixel over all (template, angle) hypotheses (k:681-721), so a
// block's statistics, best records and reference records depend on that block alone and any
// subset of blocks gives exactly the values the full search gives for them.
struct NormCtx {
  const orc_config* c = nullptr;
  P p{};
  int W = 0, H = 0, n = 0, l = 0, nt = 0, na = 0;
  std::vector<int> ox, oy, blist;
  std::vector<cd> blocks;  // spectra of the listed blocks, in list order
  std::vector<cd> T;       // processed templates (real space)
  std::vector<float> angles;
  double total_n = 0;
};

void norm_prepare(NormCtx& X, const orc_config* c, const float* image, int W, int H, const float* tmpl,
                  int n, int w, int h, const double* ctf3, const int* blocks, int nlist, int na_lim,
                  int nt_lim) {
  check_config(c);
  X.c = c;
  X.p = make_params(c, ctf3);
  X.W = W;
  X.H = H;
  X.n = n;
  X.l = c->window_size;
  const int l = X.l;
  const size_t L2 = (size_t)l * l;
  if (W < 1 || H < 1 || n < 1) throw std::invalid_argument("empty input");
  // A7: clip + invert on a private copy (the reference mutates the caller's pixels)
  std::vector<float> img(image, image + (size_t)W * H);
  clip_image(img.data(), img.size(), W, H, c->invert);
  std::vector<double> pre((size_t)W * H);
  prep_image(X.p, img.data(), W, H, true, pre.data());
  X.ox = block_offsets(W, l, c->overlap);
  X.oy = block_offsets(H, l, c->overlap);
  const int nb = (int)(X.ox.size() * X.oy.size());
  if (blocks && nlist > 0) {
    X.blist.assign(blocks, blocks + nlist);
    for (int b : X.blist)
      if (b < 0 || b >= nb) throw std::invalid_argument("block index out of range");
  } else {
    X.blist.resize(nb);
    std::iota(X.blist.begin(), X.blist.end(), 0);
  }
  std::vector<cd> all((size_t)nb * L2);
  split_blocks(pre.data(), W, H, l, X.ox, X.oy, all.data());
  X.blocks.resize(X.blist.size() * L2);
  for (size_t i = 0; i < X.blist.size(); ++i)
    std::copy(all.begin() + (size_t)X.blist[i] * L2, all.begin() + (size_t)(X.blist[i] + 1) * L2,
              X.blocks.begin() + i * L2);
  X.T.resize((size_t)n * L2);
  prep_templates(c, X.p, tmpl, n, w, h, X.T.data());
  X.angles = angle_list(c->phistep);
  X.na = (int)X.angles.size();
  if (na_lim > 0) X.na = std::min(X.na, na_lim);
  X.nt = nt_lim > 0 ? std::min(n, nt_lim) : n;
  // set_CCG_mean (k:697-706): n = N_t * int(360/phistep) (src/norm.cu:549-550)
  X.total_n = (double)n * (int)(360 / c->phistep);
  if (na_lim > 0 || nt_lim > 0) X.total_n = (double)X.nt * X.na;
}

// Runs fn(a, first listed block, number of blocks, planes) for every angle and group of listed
// blocks, with planes[(i * nt + t) * L2 + k] = c/l^2 (the real correlation of add_CCG_to_sum,
// k:691) of template t against listed block first + i.  Groups bound the plane memory.
template <typename Fn>
void norm_correlations(const NormCtx& X, Fn&& fn) {
  const int l = X.l, nt = X.nt;
  const size_t L2 = (size_t)l * l;
  const int nlist = (int)X.blist.size();
  const size_t budget = (size_t)2 << 30;
  const int group = (int)std::max<size_t>(1, std::min<size_t>(nlist, budget / (sizeof(double) * nt * L2)));
  std::vector<cd> R((size_t)nt * L2);
  std::vector<double> planes((size_t)group * nt * L2);
  for (int a = 0; a < X.na; ++a) {
#pragma omp parallel for schedule(dynamic)
    for (int t = 0; t < nt; t += 2) {
      if (t + 1 < nt)
        rotate_template_pair(X.T.data() + t * L2, X.T.data() + (t + 1) * L2, R.data() + t * L2,
                             R.data() + (t + 1) * L2, l, X.angles[a], X.p);
      else
        rotate_template(X.T.data() + t * L2, R.data() + t * L2, l, X.angles[a], X.p);
    }
    for (int g0 = 0; g0 < nlist; g0 += group) {
      const int ng = std::min(group, nlist - g0);
      // Templates t and t + 1 share one inverse transform: both products are Hermitian (real
      // block and template planes; the checkerboard is even), so the transform of
      // P_t + i P_(t+1) is c_t + i c_(t+1), equal to two separate transforms up to FP64 rounding.
      const int tp = (nt + 1) / 2;
#pragma omp parallel
      {
        std::vector<cd> C(L2);
#pragma omp for schedule(dynamic)
        for (int q = 0; q < ng * tp; ++q) {
          const int i = q / tp, t = 2 * (q % tp);
          const cd* I = X.blocks.data() + (size_t)(g0 + i) * L2;
          double* o0 = planes.data() + ((size_t)i * nt + t) * L2;
          if (t + 1 < nt) {
            const cd* T0 = R.data() + t * L2;
            const cd* T1 = R.data() + (t + 1) * L2;
            for (int y = 0; y < l; ++y)
              for (int x = 0; x < l; ++x) {
                const size_t k = (size_t)y * l + x;
                const double ch = checker(x, y, l);
                C[k] = (I[k] * std::conj(T0[k]) + cd(0, 1) * (I[k] * std::conj(T1[k]))) * ch;
              }
            orc::fft2(C.data(), l, l, +1);
            double* o1 = o0 + L2;
            for (size_t k = 0; k < L2; ++k) {
              o0[k] = C[k].real() / (double)L2;
              o1[k] = C[k].imag() / (double)L2;
            }
          } else {
            correlate(I, R.data() + t * L2, C.data(), l);
            for (size_t k = 0; k < L2; ++k) o0[k] = C[k].real() / (double)L2;
          }
        }
      }
      fn(a, g0, ng, planes.data());
    }
  }
}

// Pass 1 (ComputeCCGSum + ComputeCCGMean, src/norm.cu:523-552): per listed block pixel the sum,
// sum of squares, the best and second-best correlation and their hypotheses (angle * n +
// template, first strict maximum in (angle, template) order).
struct NormStats {
  std::vector<double> S, SS, best, second;
  std::vector<int32_t> arg, arg2;
};

NormStats norm_pass1(const NormCtx& X) {
  const size_t L2 = (size_t)X.l * X.l, nl = X.blist.size();
  NormStats st;
  st.S.assign(nl * L2, 0.0);
  st.SS.assign(nl * L2, 0.0);
  st.best.assign(nl * L2, -INFINITY);
  st.second.assign(nl * L2, -INFINITY);
  st.arg.assign(nl * L2, -1);
  st.arg2.assign(nl * L2, -1);
  const int nt = X.nt, n = X.n;
  norm_correlations(X, [&](int a, int g0, int ng, const double* planes) {
#pragma omp parallel for schedule(static)
    for (long long e = 0; e < (long long)ng * (long long)L2; ++e) {
      const int i = (int)(e / (long long)L2);
      const size_t k = (size_t)(e % (long long)L2);
      const size_t q = (size_t)(g0 + i) * L2 + k;
      double s = st.S[q], s2 = st.SS[q], b1 = st.best[q], b2 = st.second[q];
      int32_t a1 = st.arg[q], a2 = st.arg2[q];
      for (int t = 0; t < nt; ++t) {
        const double v = planes[((size_t)i * nt + t) * L2 + k];
        s += v;
        s2 += v * v;
        const int32_t id = a * n + t;
        if (v > b1) {
          b2 = b1;
          a2 = a1;
          b1 = v;
          a1 = id;
        } else if (v > b2) {
          b2 = v;
          a2 = id;
        }
      }
      st.S[q] = s;
      st.SS[q] = s2;
      st.best[q] = b1;
      st.second[q] = b2;
      st.arg[q] = a1;
      st.arg2[q] = a2;
    }
  });
  return st;
}

double norm_mu(const NormCtx& X, const NormStats& st, size_t q) { return st.S[q] / X.total_n; }
double norm_sigma(const NormCtx& X, const NormStats& st, size_t q) {
  const double m = st.S[q] / X.total_n;
  return std::sqrt(st.SS[q] / X.total_n - m * m);
}

// Pass 2 (PickParticles, src/norm.cu:560-600, with update_CCG and get_peak_pos, k:709-793): the
// reference's own records.  Per (angle, block, template) plane, z = (c/l^2 - mu)/sigma (0 when
// sigma is not positive) and, per 256-element chunk of the plane in index order, the maximum of
// get_peak_pos's tree (at stride s = 128, 64, ..., 1 element i keeps its value unless element
// i + s is strictly greater); a record when that maximum reaches the threshold and its centre
// (block offset + index % l, block offset + index / l) lies in [d_m, W - d_m) x [d_m, H - d_m).
// margin: the chunk maximum minus the largest other value of the chunk.
std::vector<orc_refrec> norm_records(const NormCtx& X, const NormStats& st) {
  const int l = X.l, nt = X.nt;
  const size_t L2 = (size_t)l * l;
  if (L2 % 256) throw std::invalid_argument("reference records need l*l divisible by 256");
  const int chunks = (int)(L2 / 256);
  const float dm = X.p.d_m;
  const int bx = (int)X.ox.size();
  std::vector<orc_refrec> out;
  norm_correlations(X, [&](int a, int g0, int ng, const double* planes) {
    std::vector<std::vector<orc_refrec>> per((size_t)ng * nt);
#pragma omp parallel for schedule(dynamic)
    for (int q = 0; q < ng * nt; ++q) {
      const int i = q / nt, t = q % nt;
      const int b = X.blist[g0 + i];
      const size_t sb = (size_t)(g0 + i) * L2;
      const double* C = planes + (size_t)q * L2;
      double z[256];
      int pos[256];
      for (int ch = 0; ch < chunks; ++ch) {
        for (int j = 0; j < 256; ++j) {
          const size_t k = (size_t)ch * 256 + j;
          const double m = norm_mu(X, st, sb + k), sd = norm_sigma(X, st, sb + k);
          z[j] = sd > 0 ? (C[k] - m) / sd : 0.0;
          pos[j] = (int)k;
        }
        double zz[256];
        std::copy(z, z + 256, zz);
        for (int s = 128; s > 0; s >>= 1)
          for (int j = 0; j < s; ++j)
            if (zz[j + s] > zz[j]) {
              zz[j] = zz[j + s];
              pos[j] = pos[j + s];
            }
        const double zmax = zz[0];
        if (!(zmax >= X.p.thresh)) continue;
        const int k = pos[0];
        const int cx = X.ox[b % bx] + k % l, cy = X.oy[b / bx] + k / l;
        if (!(cx >= dm && cx < X.W - dm && cy >= dm && cy < X.H - dm)) continue;
        double other = -INFINITY;
        for (int j = 0; j < 256; ++j)
          if (ch * 256 + j != k) other = std::max(other, z[j]);
        orc_refrec r{};
        r.score = (float)zmax;
        r.x = (float)cx;
        r.y = (float)cy;
        r.template_id = t;
        r.angle_index = a;
        r.block = b;
        r.chunk = ch;
        r.pixel = k;
        r.margin = zmax - other;
        per[q].push_back(r);
      }
    }
    for (auto& v : per) out.insert(out.end(), v.begin(), v.end());
  });
  return out;
}

}  // namespace orc_impl

using namespace orc_impl;

extern "C" {

int orc_fft1d(double* data, int n, int sign) {
  try {
    std::vector<cd> w(n);
    orc::fft(reinterpret_cast<cd*>(data), w.data(), n, sign);
    return 0;
  } catch (...) { return -1; }
}

int orc_fft2d(double* data, int ny, int nx, int sign) {
  try {
    orc::fft2(reinterpret_cast<cd*>(data), ny, nx, sign);
    return 0;
  } catch (...) { return -1; }
}

double orc_ctf(int x, int y, int nx, int ny, double apix, double defocus, double dfdiff,
               double dfang, double energy, double cs, double ampconst, int mode) {
  orc_config c{};
  c.angpix = (float)apix;
  c.energy = (float)energy;
  c.cs = (float)cs;
  const double ctf3[3] = {defocus, dfdiff, dfang};
  P p = make_params(&c, ctf3);
  p.ph.ampconst = ampconst;
  return ctf_ast(x, y, nx, ny, p, mode);
}

int orc_radial_bin(int dx, int dy) { return rbin(dx, dy); }

int orc_block_offsets(int length, int l, int overlap, int* out, int max_out) {
  auto o = block_offsets(length, l, overlap);
  for (int i = 0; i < (int)o.size() && i < max_out; ++i) out[i] = o[i];
  return (int)o.size();
}

// TileImages geometry (src/utils/tileimages.cpp:18-70): upper-left corners, x-major order.
int orc_tile_geometry(int W, int H, int* ul_x, int* ul_y, int max_out) {
  const int cols = (int)std::ceil((double)W / TILE), rows = (int)std::ceil((double)H / TILE);
  const int x_step = W / cols, y_step = H / rows;
  int k = 0;
  for (int j = 0; j < cols; ++j) {
    int cx = j == 0 ? j * TILE + TILE / 2 : (j == cols - 1 ? W - TILE / 2 : x_step * (2 * j + 1) / 2);
    for (int i = 0; i < rows; ++i) {
      int cy = i == 0 ? i * TILE + TILE / 2 : (i == rows - 1 ? H - TILE / 2 : y_step * (2 * i + 1) / 2);
      if (k < max_out) {
        ul_x[k] = std::max(cx - TILE / 2, 0);
        ul_y[k] = std::max(cy - TILE / 2, 0);
      }
      ++k;
    }
  }
  return k;
}

int orc_angles(float phistep, float* out, int max_out) {
  auto a = angle_list(phistep);
  for (int i = 0; i < (int)a.size() && i < max_out; ++i) out[i] = a[i];
  return (int)a.size();
}

int orc_prep_templates(const orc_config* c, const float* tmpl, int n, int w, int h,
                       const double* ctf3, double* out) {
  try {
    check_config(c);
    P p = make_params(c, ctf3);
    prep_templates(c, p, tmpl, n, w, h, reinterpret_cast<cd*>(out));
    return 0;
  } catch (...) { return -1; }
}

int orc_clip_image(float* img, int W, int H, int invert) {
  clip_image(img, (size_t)W * H, W, H, invert);
  return 0;
}

int orc_prep_image(const orc_config* c, const float* img, int W, int H, const double* ctf3,
                   double* out) {
  try {
    check_config(c);
    P p = make_params(c, ctf3);
    prep_image(p, img, W, H, c->norm_type != 0, out);
    return 0;
  } catch (...) { return -1; }
}

int orc_rotate_templates(const orc_config* c, const double* prepped, int n, float theta,
                         double* out) {
  try {
    check_config(c);
    const double ctf3[3] = {1, 0, 0};
    P p = make_params(c, ctf3);
    const int l = c->window_size;
    const size_t L2 = (size_t)l * l;
    const cd* in = reinterpret_cast<const cd*>(prepped);
    cd* o = reinterpret_cast<cd*>(out);
#pragma omp parallel for schedule(dynamic)
    for (int t = 0; t < n; ++t) rotate_template(in + t * L2, o + t * L2, l, theta, p);
    return 0;
  } catch (...) { return -1; }
}

int orc_search_norm(const orc_config* c, const float* image, int W, int H, const float* tmpl,
                    int n, int w, int h, const double* ctf3, orc_pick* picks, int max_picks,
                    double* mu_out, double* sigma_out, double* best_out, int32_t* arg_out,
                    int n_angles_limit, int n_templates_limit) {
  try {
    NormCtx X;
    norm_prepare(X, c, image, W, H, tmpl, n, w, h, ctf3, nullptr, 0, n_angles_limit, n_templates_limit);
    const NormStats st = norm_pass1(X);
    const int l = X.l;
    const size_t L2 = (size_t)l * l;
    const int nb = (int)X.blist.size();
    std::vector<orc_pick> cand;
    const float dm = X.p.d_m;
    for (int b = 0; b < nb; ++b) {
      const int bi = b % (int)X.ox.size(), bj = b / (int)X.ox.size();
      for (size_t k = 0; k < L2; ++k) {
        const size_t q = b * L2 + k;
        const double m = norm_mu(X, st, q);
        const double sd = norm_sigma(X, st, q);
        const double z = sd > 0 ? (st.best[q] - m) / sd : 0.0;  // update_CCG (k:709-721)
        if (mu_out) mu_out[q] = m;
        if (sigma_out) sigma_out[q] = sd;
        if (best_out) best_out[q] = st.best[q];
        if (arg_out) arg_out[q] = st.arg[q];
        if (z >= X.p.thresh) {  // src/norm.cu:585-590
          const int cx = X.ox[bi] + (int)(k % l), cy = X.oy[bj] + (int)(k / l);
          if (cx >= dm && cx < W - dm && cy >= dm && cy < H - dm) {
            orc_pick pk{};
            pk.x = (float)cx;
            pk.y = (float)cy;
            pk.template_id = st.arg[q] % n;
            pk.angle_index = st.arg[q] / n;
            float e = 360.0f - X.angles[pk.angle_index];  // src/norm.cu:654-655
            if (e >= 360.0f) e -= 360.0f;
            pk.psi = e;
            pk.score = (float)z;
            pk.block = b;
            pk.pixel = (int)k;
            pk.local_x = (float)(k % l);
            pk.local_y = (float)(k / l);
            cand.push_back(pk);
          }
        }
      }
    }
    const int kept = nms(cand.data(), (int)cand.size(), c->nms_radius);
    for (int i = 0; i < kept && i < max_picks; ++i) picks[i] = cand[i];
    return kept;
  } catch (...) { return -1; }
}

int orc_search_norm_blocks(const orc_config* c, const float* image, int W, int H, const float* tmpl,
                           int n, int w, int h, const double* ctf3, const int* blocks, int nlist,
                           int n_angles_limit, int n_templates_limit, double* mu, double* sigma,
                           double* best, double* second, int32_t* arg, int32_t* arg2) {
  try {
    if (!blocks || nlist < 1) throw std::invalid_argument("empty block list");
    NormCtx X;
    norm_prepare(X, c, image, W, H, tmpl, n, w, h, ctf3, blocks, nlist, n_angles_limit, n_templates_limit);
    const NormStats st = norm_pass1(X);
    const size_t N = X.blist.size() * (size_t)X.l * X.l;
    for (size_t q = 0; q < N; ++q) {
      mu[q] = norm_mu(X, st, q);
      sigma[q] = norm_sigma(X, st, q);
      best[q] = st.best[q];
      second[q] = st.second[q];
      arg[q] = st.arg[q];
      arg2[q] = st.arg2[q];
    }
    return 0;
  } catch (...) { return -1; }
}

int orc_search_norm_records(const orc_config* c, const float* image, int W, int H, const float* tmpl,
                            int n, int w, int h, const double* ctf3, const int* blocks, int nlist,
                            int n_angles_limit, int n_templates_limit, orc_refrec* recs, int max_recs) {
  try {
    NormCtx X;
    norm_prepare(X, c, image, W, H, tmpl, n, w, h, ctf3, blocks, nlist, n_angles_limit, n_templates_limit);
    const NormStats st = norm_pass1(X);
    const auto r = norm_records(X, st);
    for (int i = 0; i < (int)r.size() && i < max_recs; ++i) recs[i] = r[i];
    return (int)r.size();
  } catch (...) { return -1; }
}

int orc_search_nonorm(const orc_config* c, const float* image, int W, int H, const float* tmpl,
                      int n, int w, int h, const double* ctf3, orc_pick* recs, int max_recs,
                      int* n_recs, orc_pick* picks, int max_picks) {
  try {
    check_config(c);
    P p = make_params(c, ctf3);
    const int l = c->window_size;
    const size_t L2 = (size_t)l * l;
    if (W < TILE || H < TILE) throw std::invalid_argument("norm_type=0 needs W,H >= 720");
    if (l > TILE) throw std::invalid_argument("window_size must be <= 720");
    // blocks inside a tile (src/nonorm.cu:71-85), uniform placement (k:605-624)
    const int nblk = (int)block_offsets(TILE, l, c->overlap).size();
    std::vector<cd> T((size_t)n * L2);
    prep_templates(c, p, tmpl, n, w, h, T.data());
    std::vector<int> ulx(64), uly(64);
    int ntiles = orc_tile_geometry(W, H, ulx.data(), uly.data(), 0);
    ulx.resize(ntiles);
    uly.resize(ntiles);
    orc_tile_geometry(W, H, ulx.data(), uly.data(), ntiles);
    const auto angles = angle_list(c->phistep);
    const float dm4 = p.d_m / 4;
    std::vector<orc_pick> all;
    for (int tile = 0; tile < ntiles; ++tile) {
      std::vector<float> tdata((size_t)TILE * TILE);
      for (int r = 0; r < TILE; ++r)
        std::memcpy(&tdata[(size_t)r * TILE], image + (size_t)(uly[tile] + r) * W + ulx[tile],
                    sizeof(float) * TILE);
      if (c->invert)
        for (auto& v : tdata) v = -v;  // src/nonorm.cu:157-165
      std::vector<double> pre((size_t)TILE * TILE);
      if (c->phase_flip == 1)
        prep_image(p, tdata.data(), TILE, TILE, false, pre.data());
      else
        for (size_t i = 0; i < pre.size(); ++i) pre[i] = tdata[i];
      // SplitImage (src/nonorm.cu:391-450)
      std::vector<cd> blk((size_t)nblk * nblk * L2);
      for (int b = 0; b < nblk * nblk; ++b) {
        cd* B = blk.data() + b * L2;
        const int bi = b % nblk, bj = b / nblk;
        for (int y = 0; y < l; ++y)
          for (int x = 0; x < l; ++x) {
            const int X = bi * (l - c->overlap) + x, Y = bj * (l - c->overlap) + y;
            B[(size_t)y * l + x] = (X < TILE && Y < TILE) ? cd(pre[(size_t)Y * TILE + X], 0) : cd(0);
          }
        if (c->phase_flip == 1) {
          orc::fft2(B, l, l, -1);
          scale_all(B, L2, 1.0 / (double)L2);
          spectral_normalise(B, l);
          orc::fft2(B, l, l, +1);
        }
      }
      for (int a = 0; a < (int)angles.size(); ++a) {
        const float e = angles[a];
        // PickParticles (src/nonorm.cu:452-537)
        std::vector<cd> rot(blk.size());
        for (int b = 0; b < nblk * nblk; ++b) {
          rotate_plane(blk.data() + b * L2, rot.data() + b * L2, l, e);
          orc::fft2(rot.data() + b * L2, l, l, -1);
          scale_all(rot.data() + b * L2, L2, 1.0 / (double)L2);
        }
        std::vector<double> sc((size_t)n, 0.0), cxs(n, 0.0), cys(n, 0.0);
        std::vector<int> pk_pix(n, 0), pk_blk(n, 0);
        for (int b = 0; b < nblk * nblk; ++b) {
          const int bi = b % nblk, bj = b / nblk;
          std::vector<PlaneStats> st(n);
#pragma omp parallel
          {
            std::vector<cd> C(L2);
#pragma omp for schedule(dynamic)
            for (int t = 0; t < n; ++t) {
              correlate(rot.data() + b * L2, T.data() + t * L2, C.data(), l);
              st[t] = nonorm_plane(C.data(), l, dm4);
            }
          }
          for (int t = 0; t < n; ++t) {
            const double score = nonorm_score(st[t], l);
            float centerx, centery;
            nonorm_center(st[t].pos, bi, bj, l, c->overlap, e, centerx, centery);
            if (sc[t] < score && centerx >= 0 && centerx < TILE && centery >= 0 && centery < TILE) {
              sc[t] = score;
              cxs[t] = centerx;
              cys[t] = centery;
              pk_pix[t] = (int)st[t].pos;
              pk_blk[t] = b;
            }
          }
        }
        for (int t = 0; t < n; ++t)
          if (sc[t] >= p.thresh) {  // OutputScore (src/nonorm.cu:539-557)
            orc_pick r{};
            r.local_x = (float)cxs[t];
            r.local_y = (float)cys[t];
            r.x = (float)(ulx[tile] + cxs[t]);
            r.y = (float)(uly[tile] + cys[t]);
            r.template_id = t;
            r.angle_index = a;
            r.psi = e;
            r.score = (float)sc[t];
            r.block = tile;
            r.pixel = pk_blk[t] * (int)L2 + pk_pix[t];
            all.push_back(r);
          }
      }
    }
    if (n_recs) *n_recs = (int)all.size();
    for (int i = 0; i < (int)all.size() && i < max_recs; ++i) recs[i] = all[i];
    const int kept = nms(all.data(), (int)all.size(), c->nms_radius);
    for (int i = 0; i < kept && i < max_picks; ++i) picks[i] = all[i];
    return kept;
  } catch (...) { return -1; }
}

// ---- stage pieces, exposed so that each can be pinned on its own ----
int orc_whiten_template(double* F, int l) {
  try {
    if (l <= 0 || l % 2) throw std::invalid_argument("l");
    whiten_template(reinterpret_cast<cd*>(F), l);
    return 0;
  } catch (...) { return -1; }
}

int orc_weight_template(const orc_config* c, const double* ctf3, double* F, int l) {
  try {
    check_config(c);
    P p = make_params(c, ctf3);
    weight_template(reinterpret_cast<cd*>(F), l, p);
    return 0;
  } catch (...) { return -1; }
}

int orc_filter_image_spectrum(const orc_config* c, const double* ctf3, double* F, int nx, int ny) {
  try {
    check_config(c);
    P p = make_params(c, ctf3);
    filter_image_spectrum(p, reinterpret_cast<cd*>(F), nx, ny);
    return 0;
  } catch (...) { return -1; }
}

int orc_rotate_plane(const double* in, double* out, int l, float theta) {
  try {
    rotate_plane(reinterpret_cast<const cd*>(in), reinterpret_cast<cd*>(out), l, theta);
    return 0;
  } catch (...) { return -1; }
}

int orc_nonorm_plane(const double* C, int l, float d_m, double* stats4) {
  try {
    const PlaneStats s = nonorm_plane(reinterpret_cast<const cd*>(C), l, d_m / 4);
    stats4[0] = s.peak;
    stats4[1] = (double)s.pos;
    stats4[2] = s.sum;
    stats4[3] = s.sum2;
    return 0;
  } catch (...) { return -1; }
}

double orc_nonorm_score(const double* stats4, int l) {
  PlaneStats s{stats4[0], stats4[2], stats4[3], (long long)stats4[1]};
  return nonorm_score(s, l);
}

void orc_nonorm_center(long long pos, int bi, int bj, int l, int overlap, float theta, float* cx, float* cy) {
  nonorm_center(pos, bi, bj, l, overlap, theta, *cx, *cy);
}

int orc_nms(orc_pick* cand, int n, float radius) { return nms(cand, n, radius); }

int orc_correlate(const double* I, const double* T, int l, double* out) {
  try {
    std::vector<cd> C((size_t)l * l);
    correlate(reinterpret_cast<const cd*>(I), reinterpret_cast<const cd*>(T), C.data(), l);
    std::memcpy(out, C.data(), sizeof(cd) * C.size());
    return 0;
  } catch (...) { return -1; }
}

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void orc_set_num_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

}  // extern "C"
