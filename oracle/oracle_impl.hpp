// Internal pieces shared by the FP64 oracle's translation units (TEST INFRASTRUCTURE ONLY).
//
// Geometry that drives integer decisions is evaluated in float exactly as the HIP path does;
// data arithmetic is FP64.  Reference citations are on the definitions in gisspa_oracle.cpp.
#pragma once
#include <cmath>
#include <complex>
#include <vector>

#include "fft64.hpp"
#include "gisspa_oracle.h"

namespace orc_impl {
using orc::cd;

constexpr double PI = 3.1415926535;  // include/utils/constants.h:3
constexpr double EPS = 1e-8;         // include/utils/constants.h:4
constexpr int TILE = 720;            // include/utils/constants.h:5

// Physics constants of Parameters (include/EMReader/DataReader2.h:76-77), as the float values
// the reference stores.
struct Phys {
  double edge_half_width = 4.0f, ampconst = 0.07f;
  double a = -10.81f, b = 1.f, b2 = 0.32f, bf = -18.17f, bf2 = -15.22f, bf3 = 1.72f;
};

struct P {
  // config (float in the reference)
  float apix, kk, energy, cs, highres, lowres, d_m, thresh;
  // per micrograph (SetParams, src/norm.cu:183-192)
  double lambda, dfu, dfv, dfdiff, dfang, defocus;
  Phys ph;
};

// Band-pass radii are float expressions of config values in the reference (k:183-189,265-273).
struct Band {
  float hr, lr;
};

// floor(hypot(dx,dy) + 0.5) - 1 computed exactly (reference: k:118,136,165,264,329).
inline int rbin(long long dx, long long dy) {
  long long n = dx * dx + dy * dy;
  long long k = (long long)std::sqrt((double)n);
  while (k * k > n) --k;
  while ((k + 1) * (k + 1) <= n) ++k;
  if (n > k * k + k) ++k;  // sqrt(n) >= k + 1/2  <=>  n > k^2 + k
  return (int)k - 1;
}

// Rotation geometry of rotate_subIMG (k:534-603), evaluated in float exactly as the HIP path
// does; returns false outside [0,l-1]^2.
struct Bil {
  int k0, k1, k2, k3;
  float t, u;
  float x2, y2;  // sample position
};
inline bool bilinear_coords(int i, int j, int l, float cose, float sine, Bil& o) {
  const float x = (float)(i - l / 2), y = (float)(j - l / 2);
  const float half = (float)(l / 2);
  const float x2 = std::fmaf(cose, x, sine * y) + half;
  const float y2 = std::fmaf(-sine, x, cose * y) + half;
  o.x2 = x2;
  o.y2 = y2;
  if (x2 < 0 || x2 > l - 1.0f || y2 < 0 || y2 > l - 1.0f) return false;
  const int ii = (int)std::floor(x2), jj = (int)std::floor(y2);
  o.k0 = ii + jj * l;
  o.k1 = o.k0 + 1;
  o.k2 = o.k0 + l + 1;
  o.k3 = o.k0 + l;
  if (ii == l - 1) { o.k1 -= 1; o.k2 -= 1; }
  if (jj == l - 1) { o.k2 -= l; o.k3 -= l; }
  o.t = x2 - (float)ii;
  o.u = y2 - (float)jj;
  return true;
}

// cos/sin of e*PI/180 rounded to float (k:538-539: double cos of a double argument).
inline void rot_cs(float e, float& c, float& s) {
  c = (float)std::cos((double)e * PI / 180);
  s = (float)std::sin((double)e * PI / 180);
}

// Checkerboard of compute_corner_CCG (k:669-677): -1 where the reference negates.
inline double checker(int x, int y, int l) {
  const int of = (l / 2) % 2;
  const int st = (of == y % 2) ? 1 : 0;
  return ((x - st) % 2 == 0) ? -1.0 : 1.0;
}

P make_params(const orc_config* c, const double* ctf3);
double ctf_signed(int x1, int y1, int nx, int ny, const P& p, double* dv);
double ctf_ast(int x1, int y1, int nx, int ny, const P& p, int mode);
Band band(int l, const P& p);
double bandpass_template(int r, const Band& b);
void pad_template(const float* t, int w, int h, int l, cd* out);
void standardise_template(cd* t, int l);
void whiten_template(cd* F, int l);
void weight_template(cd* F, int l, const P& p);
void spectral_normalise(cd* F, int l);
double mask_factor(int x, int y, int l, double d_m_cfg, double ehw);
void apply_mask(cd* t, int l, double d_m_cfg, double ehw);
void scale_all(cd* x, size_t n, double s);
std::vector<int> block_offsets(int length, int l, int overlap);
std::vector<float> angle_list(float phistep);
void filter_image_spectrum(const P& p, cd* F, int nx, int ny);
void prep_image(const P& p, const float* img, int nx, int ny, bool dc_zero, double* out);
double bandpass_image(int r, const Band& b);
void rotate_plane(const cd* in, cd* out, int l, float theta);
void clip_image(float* img, size_t N, int nx, int ny, int invert);
void rotate_template(const cd* Tp, cd* out, int l, float theta, const P& p);
void split_blocks(const double* img, int nx, int ny, int l, const std::vector<int>& ox,
                  const std::vector<int>& oy, cd* out);
void correlate(const cd* I, const cd* T, cd* C, int l);

// Per-plane reduction of the norm_type=0 scoring.
struct PlaneStats {
  double peak, sum, sum2;
  long long pos;
};
PlaneStats nonorm_plane(const cd* C, int l, float dm4);
double nonorm_score(const PlaneStats& s, int l);
void nonorm_center(long long pos, int bi, int bj, int l, int overlap, float e, float& cx, float& cy);
void check_config(const orc_config* c);

}  // namespace orc_impl
