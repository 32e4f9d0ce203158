// FP64 adjoint refinement of norm_type=1 picks (TEST INFRASTRUCTURE ONLY; see gisspa_oracle.h).
//
// The reference has no refinement: its picks stay on the integer pixel grid, the in-plane
// angle grid (phistep) and the micrograph defocus.  This file states the continuous score the
// refinement ascends and its exact gradient, in double precision:
//
//   z(dx, dy, theta, ddf) = (c - mu_p) / sigma_p,  mu_p and sigma_p frozen at the pick's block
//   pixel p (set_CCG_mean, src/kernels.cu:697-706), and
//   c = (1/l^2) Re sum_k I_b(k) conj(T(k)) checker(k) E(k, p + (dx, dy)),
//
// with I_b the pick's block spectrum (SplitImage, src/norm.cu:404-477), checker the sign
// pattern of compute_corner_CCG (k:669-677) and T the spectrum of the template processed at
// defocus + ddf (whitened spectrum of src/norm.cu:194-318, apply_weighting_function + CTF_AST
// k:243-317, spectral normalisation k:319-382,445-458, inverse transform src/norm.cu:350),
// rotated by theta, masked and standardised (RotateTemplate, src/norm.cu:479-521).
// E(k, q) = ph(kx, qx) ph(ky, qy) is the band-limited shift: ph(k, q) = exp(2 pi i k_s q / l)
// with the signed frequency k_s in (-l/2, l/2), and cos(pi q) at the Nyquist index.  At integer
// (dx, dy), theta on the angle grid and ddf = 0, c is the search's correlation value exactly.
//
// The gradient is the exact derivative of this function: phase ramps for (dx, dy); for theta
// the bilinear interpolation derivative (piecewise; kinks where a sample crosses a grid line)
// through the mask, the box standardisation and the transform; for ddf the derivative of
// |CTF| sqrt(1/(N/S' + kk CTF^2)) through the spectral normalisation, the inverse transform,
// the rotation and the standardisation.  Central differences pin it in the tests.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "oracle_impl.hpp"

namespace orc_impl {
namespace {

struct RefineCtx {
  orc_config cfg{};
  P p{};
  int l = 0, n = 0, nb = 0;
  bool float_geometry = false;
  std::vector<cd> white;   // n x l^2 whitened template spectra [y][x]
  std::vector<cd> blocks;  // nb x l^2 block spectra I_b [y][x]
};

// Template weight of weight_template at (x, y) and its derivative with respect to the defocus
// offset (Angstrom): f = band |v| g with g = (N + kk v^2)^-1/2, df = band sign(v) v' N g^3.
void weight_deriv(int x, int y, int l, const P& p, double* f, double* df) {
  const Band b = band(l, p);
  const double ds = (double)(1.0f / (p.apix * (float)l));
  const int fx = std::min(x, l - x), fy = std::min(y, l - y);
  const double r = std::hypot((double)fy, (double)fx);
  const int rr = rbin(fx, fy);
  const double bp = bandpass_template(rr, b);
  *f = bp;
  *df = 0;
  if (bp == 0.0) {
    *f = 0;
    return;
  }
  if (rr < l / 2 && rr >= 0) {
    const double ss = r * ds;
    double dv = 0;
    const double v = ctf_signed(x, (y + l / 2) % l, l, l, p, &dv);
    const double signal = std::exp(p.ph.bf * ss * ss + p.ph.bf2 * ss + p.ph.bf3) / ((double)p.kk + 1);
    const double N = std::exp(p.ph.a * ss * ss + p.ph.b * ss + p.ph.b2) / signal;
    const double g = std::sqrt(1 / (N + (double)p.kk * v * v));
    *f = bp * std::fabs(v) * g;
    *df = bp * (v > 0 ? 1.0 : (v < 0 ? -1.0 : 0.0)) * dv * N * g * g * g;
  }
}

// Processed template t (real space) and dt/d(ddf) at defocus offset ddf (Angstrom).
void template_at(const RefineCtx& C, int t, double ddf, std::vector<double>& tp, std::vector<double>& dtp) {
  const int l = C.l;
  const size_t L2 = (size_t)l * l;
  P p = C.p;
  p.dfu -= ddf * 1e-4;
  p.dfv -= ddf * 1e-4;
  const cd* F = C.white.data() + (size_t)t * L2;
  std::vector<cd> G(L2), dG(L2);
  for (int y = 0; y < l; ++y)
    for (int x = 0; x < l; ++x) {
      double f, df;
      weight_deriv(x, y, l, p, &f, &df);
      G[(size_t)y * l + x] = F[(size_t)y * l + x] * f;
      dG[(size_t)y * l + x] = F[(size_t)y * l + x] * df;
    }
  // spectral_normalise: m = sqrt(s) / cnt over x <= l/2, radial bin < l/2
  double s = 0, ds = 0, cnt = 0;
  for (int y = 0; y < l; ++y)
    for (int x = 0; x <= l / 2; ++x) {
      const int r = rbin(std::min(x, l - x), std::min(y, l - y));
      if (r < l / 2 && r >= 0) {
        const cd g = G[(size_t)y * l + x];
        s += std::norm(g);
        ds += 2 * (std::conj(g) * dG[(size_t)y * l + x]).real();
        cnt += 1;
      }
    }
  const double m = std::sqrt(s) / cnt, dm = ds / (2 * std::sqrt(s) * cnt);
  for (size_t i = 0; i < L2; ++i) {
    dG[i] = dG[i] / m - G[i] * (dm / (m * m));
    G[i] /= m;
  }
  orc::fft2(G.data(), l, l, +1);
  orc::fft2(dG.data(), l, l, +1);
  tp.resize(L2);
  dtp.resize(L2);
  for (size_t i = 0; i < L2; ++i) {
    tp[i] = G[i].real();
    dtp[i] = dG[i].real();
  }
}

// Rotation sample geometry: float (as the search) or double.
bool sample(const RefineCtx& C, int i, int j, double theta, Bil& b, double& dx2, double& dy2) {
  const int l = C.l;
  const double c = std::cos(theta * PI / 180), s = std::sin(theta * PI / 180);
  const double x = i - l / 2, y = j - l / 2;
  // d(x2, y2)/d theta per degree
  dx2 = (-s * x + c * y) * PI / 180;
  dy2 = (-c * x - s * y) * PI / 180;
  if (C.float_geometry) {
    float cf, sf;
    rot_cs((float)theta, cf, sf);
    return bilinear_coords(i, j, l, cf, sf, b);
  }
  const double x2 = c * x + s * y + l / 2, y2 = -s * x + c * y + l / 2;
  if (x2 < 0 || x2 > l - 1.0 || y2 < 0 || y2 > l - 1.0) return false;
  const int ii = (int)std::floor(x2), jj = (int)std::floor(y2);
  b.k0 = ii + jj * l;
  b.k1 = b.k0 + 1;
  b.k2 = b.k0 + l + 1;
  b.k3 = b.k0 + l;
  if (ii == l - 1) { b.k1 -= 1; b.k2 -= 1; }
  if (jj == l - 1) { b.k2 -= l; b.k3 -= l; }
  b.t = 0;  // unused in double mode
  b.u = 0;
  b.x2 = (float)x2;
  b.y2 = (float)y2;
  return true;
}

struct Fields {
  std::vector<cd> T, Tth, Tdf;  // spectra of r, dr/dtheta, dr/dddf, each DFT/l^2
};

void rotated_fields(const RefineCtx& C, int t, double theta, double ddf, Fields& out) {
  const int l = C.l;
  const size_t L2 = (size_t)l * l;
  std::vector<double> tp, dtp;
  template_at(C, t, ddf, tp, dtp);
  std::vector<double> r0(L2), gth(L2), gdf(L2);
  const double c = std::cos(theta * PI / 180), s = std::sin(theta * PI / 180);
  for (int j = 0; j < l; ++j)
    for (int i = 0; i < l; ++i) {
      const size_t q = (size_t)j * l + i;
      Bil b;
      double dx2, dy2;
      double v = 0, vth = 0, vdf = 0;
      if (sample(C, i, j, theta, b, dx2, dy2)) {
        double tt, uu;
        if (C.float_geometry) {
          tt = b.t;
          uu = b.u;
        } else {
          const double x2 = c * (i - l / 2) + s * (j - l / 2) + l / 2;
          const double y2 = -s * (i - l / 2) + c * (j - l / 2) + l / 2;
          tt = x2 - std::floor(x2);
          uu = y2 - std::floor(y2);
        }
        const double a0 = tp[b.k0], a1 = tp[b.k1], a2 = tp[b.k2], a3 = tp[b.k3];
        v = a0 * (1 - tt) * (1 - uu) + a1 * tt * (1 - uu) + a3 * (1 - tt) * uu + a2 * tt * uu;
        const double dvx = (1 - uu) * (a1 - a0) + uu * (a2 - a3);
        const double dvy = (1 - tt) * (a3 - a0) + tt * (a2 - a1);
        vth = dvx * dx2 + dvy * dy2;
        vdf = dtp[b.k0] * (1 - tt) * (1 - uu) + dtp[b.k1] * tt * (1 - uu) +
              dtp[b.k3] * (1 - tt) * uu + dtp[b.k2] * tt * uu;
      }
      const double m = mask_factor(i, j, l, C.p.d_m, C.p.ph.edge_half_width);
      r0[q] = v * m;
      gth[q] = vth * m;
      gdf[q] = vdf * m;
    }
  double s0 = 0, s1 = 0;
  for (size_t q = 0; q < L2; ++q) {
    s0 += r0[q];
    s1 += r0[q] * r0[q];
  }
  const double mean = s0 / L2, var = s1 / L2 - mean * mean, sd = std::sqrt(var);
  auto standardised = [&](const std::vector<double>& g, std::vector<cd>& out_spec) {
    double gs = 0, rg = 0;
    for (size_t q = 0; q < L2; ++q) {
      gs += g[q];
      rg += r0[q] * g[q];
    }
    const double dmean = gs / L2, dvar = 2 * (rg / L2 - mean * dmean), dsd = dvar / (2 * sd);
    out_spec.resize(L2);
    for (size_t q = 0; q < L2; ++q)
      out_spec[q] = cd((g[q] - dmean) / sd - (r0[q] - mean) * dsd / (sd * sd), 0);
    orc::fft2(out_spec.data(), l, l, -1);
    scale_all(out_spec.data(), L2, 1.0 / (double)L2);
  };
  out.T.resize(L2);
  for (size_t q = 0; q < L2; ++q) out.T[q] = cd((r0[q] - mean) / sd, 0);
  orc::fft2(out.T.data(), l, l, -1);
  scale_all(out.T.data(), L2, 1.0 / (double)L2);
  standardised(gth, out.Tth);
  standardised(gdf, out.Tdf);
}

// Band-limited shift factor ph(k, q) and its derivative d/dq.
void phase(int k, int l, double q, cd& ph, cd& dph) {
  if (2 * k == l) {
    ph = cd(std::cos(M_PI * q), 0);
    dph = cd(-M_PI * std::sin(M_PI * q), 0);
    return;
  }
  const int ks = k < l / 2 ? k : k - l;
  const double a = 2 * M_PI * ks * q / l;
  ph = cd(std::cos(a), std::sin(a));
  dph = cd(0, 2 * M_PI * ks / l) * ph;
}

// z and dz/d(dx, dy, theta, ddf) of one pick; c (the correlation value) optional.
void eval_peak(const RefineCtx& C, const orc_peak& pk, double mu, double sigma, double* z,
               double* grad, double* c_out) {
  const int l = C.l;
  if (pk.block < 0 || pk.block >= C.nb || pk.template_id < 0 || pk.template_id >= C.n ||
      pk.pixel < 0 || pk.pixel >= l * l)
    throw std::invalid_argument("peak out of range");
  Fields f;
  rotated_fields(C, pk.template_id, pk.theta, pk.ddf, f);
  const double qx = pk.pixel % l + pk.dx, qy = pk.pixel / l + pk.dy;
  std::vector<cd> phx(l), dphx(l), phy(l), dphy(l);
  for (int k = 0; k < l; ++k) {
    phase(k, l, qx, phx[k], dphx[k]);
    phase(k, l, qy, phy[k], dphy[k]);
  }
  const cd* I = C.blocks.data() + (size_t)pk.block * l * l;
  double c = 0, gx = 0, gy = 0, gth = 0, gdf = 0;
  for (int y = 0; y < l; ++y)
    for (int x = 0; x < l; ++x) {
      const size_t q = (size_t)y * l + x;
      const cd X = I[q] * checker(x, y, l);
      const cd E = phx[x] * phy[y];
      c += (X * std::conj(f.T[q]) * E).real();
      gx += (X * std::conj(f.T[q]) * dphx[x] * phy[y]).real();
      gy += (X * std::conj(f.T[q]) * phx[x] * dphy[y]).real();
      gth += (X * std::conj(f.Tth[q]) * E).real();
      gdf += (X * std::conj(f.Tdf[q]) * E).real();
    }
  const double inv = 1.0 / ((double)l * l);
  c *= inv;
  if (c_out) *c_out = c;
  *z = (c - mu) / sigma;
  grad[0] = gx * inv / sigma;
  grad[1] = gy * inv / sigma;
  grad[2] = gth * inv / sigma;
  grad[3] = gdf * inv / sigma;
}

}  // namespace
}  // namespace orc_impl

using namespace orc_impl;

extern "C" {

void* orc_refine_create(const orc_config* c, const float* image, int W, int H, const float* tmpl,
                        int n, int w, int h, const double* ctf3) {
  try {
    check_config(c);
    if (!c->norm_type) throw std::invalid_argument("refinement is defined for norm_type=1");
    auto C = std::make_unique<RefineCtx>();
    C->cfg = *c;
    C->p = make_params(c, ctf3);
    const int l = c->window_size;
    const size_t L2 = (size_t)l * l;
    C->l = l;
    C->n = n;
    std::vector<float> img(image, image + (size_t)W * H);
    clip_image(img.data(), img.size(), W, H, c->invert);
    std::vector<double> pre((size_t)W * H);
    prep_image(C->p, img.data(), W, H, true, pre.data());
    const auto ox = block_offsets(W, l, c->overlap), oy = block_offsets(H, l, c->overlap);
    C->nb = (int)(ox.size() * oy.size());
    C->blocks.resize((size_t)C->nb * L2);
    split_blocks(pre.data(), W, H, l, ox, oy, C->blocks.data());
    C->white.resize((size_t)n * L2);
#pragma omp parallel for schedule(dynamic)
    for (int t = 0; t < n; ++t) {  // src/norm.cu:194-318 up to the weighting
      cd* T = C->white.data() + (size_t)t * L2;
      pad_template(tmpl + (size_t)t * w * h, w, h, l, T);
      standardise_template(T, l);
      orc::fft2(T, l, l, -1);
      scale_all(T, L2, 1.0 / (double)L2);
      whiten_template(T, l);
    }
    return C.release();
  } catch (...) { return nullptr; }
}

void orc_refine_free(void* ctx) { delete static_cast<RefineCtx*>(ctx); }

void orc_refine_geometry(void* ctx, int float_geometry) {
  if (ctx) static_cast<RefineCtx*>(ctx)->float_geometry = float_geometry != 0;
}

int orc_refine_eval(void* ctx, const orc_peak* peaks, int k, const double* mu, const double* sigma,
                    double* z, double* grad, double* c) {
  if (!ctx || !peaks || k < 0 || !mu || !sigma || !z || !grad) return -1;
  const RefineCtx& C = *static_cast<RefineCtx*>(ctx);
  int bad = 0;
#pragma omp parallel for schedule(dynamic) reduction(+ : bad)
  for (int i = 0; i < k; ++i) {
    try {
      eval_peak(C, peaks[i], mu[i], sigma[i], z + i, grad + 4 * i, c ? c + i : nullptr);
    } catch (...) { ++bad; }
  }
  return bad ? -1 : 0;
}

int orc_refine_run(void* ctx, orc_peak* peaks, int k, const double* mu, const double* sigma,
                   int iters, double step, const double* scale4, double* z, double* grad, double* z0) {
  if (!ctx || !peaks || k < 0 || !mu || !sigma || !z || !grad || !scale4 || iters < 0) return -1;
  const RefineCtx& C = *static_cast<RefineCtx*>(ctx);
  int bad = 0;
#pragma omp parallel for schedule(dynamic) reduction(+ : bad)
  for (int i = 0; i < k; ++i) {
    try {
      orc_peak pk = peaks[i];
      double zz = 0, g[4] = {0, 0, 0, 0};
      for (int it = 0; it <= iters; ++it) {
        eval_peak(C, pk, mu[i], sigma[i], &zz, g, nullptr);
        if (it == 0 && z0) z0[i] = zz;
        if (it == iters) break;
        // ascent in scaled coordinates u = param / scale: u += step * dz/du
        pk.dx += step * scale4[0] * scale4[0] * g[0];
        pk.dy += step * scale4[1] * scale4[1] * g[1];
        pk.theta += step * scale4[2] * scale4[2] * g[2];
        pk.ddf += step * scale4[3] * scale4[3] * g[3];
      }
      peaks[i] = pk;
      z[i] = zz;
      for (int q = 0; q < 4; ++q) grad[4 * i + q] = g[q];
    } catch (...) { ++bad; }
  }
  return bad ? -1 : 0;
}

// The search path's correlation plane c(p) = Re(C[p]) / l^2 of template t at (theta, ddf) with
// block b: template preprocessing at defocus + ddf (src/norm.cu:319-354 on the whitened
// spectrum), RotateTemplate, compute_corner_CCG and the inverse transform.
int orc_refine_plane(void* ctx, int block, int t, double theta, double ddf, double* out) {
  if (!ctx || !out) return -1;
  try {
    const RefineCtx& C = *static_cast<RefineCtx*>(ctx);
    const int l = C.l;
    const size_t L2 = (size_t)l * l;
    if (block < 0 || block >= C.nb || t < 0 || t >= C.n) return -1;
    P p = C.p;
    p.dfu -= ddf * 1e-4;
    p.dfv -= ddf * 1e-4;
    std::vector<cd> T(C.white.begin() + (size_t)t * L2, C.white.begin() + (size_t)(t + 1) * L2);
    weight_template(T.data(), l, p);
    spectral_normalise(T.data(), l);
    orc::fft2(T.data(), l, l, +1);
    std::vector<cd> R(L2), Cc(L2);
    rotate_template(T.data(), R.data(), l, (float)theta, p);
    correlate(C.blocks.data() + (size_t)block * L2, R.data(), Cc.data(), l);
    for (size_t q = 0; q < L2; ++q) out[q] = Cc[q].real() / (double)L2;
    return 0;
  } catch (...) { return -1; }
}

}  // extern "C"
