// Adjoint refinement of norm_type=1 picks: z(dx, dy, theta, ddf) and its exact gradient.
//
// No reference counterpart (the reference's picks stay on the pixel grid, the phistep angle
// grid and the micrograph defocus).  The score is the search's correlation made continuous:
//
//   c = sum over the half spectrum of w(kx) Re[ Ihat_b(k) (-1)^(kx+ky) conj(T(k)) E(k, p + d) ]
//
// with Ihat_b the block spectrum as the correlation kernels use it (1/l^2 and the odd-l/2 sign
// folded in), T the spectrum / l^2 of the template processed at defocus + ddf (whitened
// spectrum of src/norm.cu:194-318, apply_weighting_function + CTF_AST k:243-317, spectral
// normalisation k:319-382,445-458, inverse transform src/norm.cu:350), rotated by theta,
// masked and standardised (RotateTemplate, src/norm.cu:479-521); E the band-limited shift
// (signed frequency, cos(pi q) at Nyquist) and w = 1 on the kx = 0 and l/2 columns, else 2.
// z = (c - mu_p) / sigma_p with mu_p, sigma_p frozen at the pick's block pixel.  At integer
// shift, theta on the angle grid and ddf = 0 this is the search's z.
//
// One evaluation of K picks (batched, one plane per pick):
//   k_ref_power / k_ref_pack  weighted spectrum G at the pick's defocus and dG/d ddf, spectrally
//                             normalised, packed as G + i dG (both have real inverses)
//   inverse 2-D transform    -> t + i dt (processed template and its defocus derivative)
//   k_ref_rotate / k_ref_stats
//                             r, dr/dtheta (bilinear derivative), dr/d ddf through the mask,
//                             packed as A = r + i dr/dtheta per pick and B = dr/d ddf of two
//                             picks per plane; box sums for the standardisation
//   forward 2-D transform    of the K A planes and (K + 1) / 2 B planes
//   k_ref_dot                 the standardisation applied to the spectra, c and its four
//                             derivatives (phase ramps for the shift)
//   k_ref_update              z, gradient, ascent step
// For l = 320 and 720 the rotation (k_ref_mask_table once, then k_ref_rotate_rows) also does the
// row pass of the forward transform, writing only the rows the mask leaves non-zero, and the
// column pass is fused with the dot (k_ref_cols_dot): the spectra of A and B are never stored.
#include <algorithm>
#include <cmath>
#include <type_traits>

#include "common.hpp"
#include "device_common.hpp"
#include "fft_tile.hpp"
#include "kernels.hpp"

namespace gs {

namespace {

using dev::band_template;
using dev::bil_coords;
using dev::Bil;
using dev::block_sum;
using dev::ctf_signed;
using dev::kRefPi;
using dev::kRotTX;
using dev::kRotTY;
using dev::rbin;

constexpr int BS = 256;
constexpr int kPowerChunks = 16;  // row chunks of the power reduction per pick

__global__ void k_ref_table(const int32_t* __restrict__ idx, int n, int l, DevParams p,
                            RefWeight* __restrict__ tab) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int e = idx[i], y = e / l, x = e % l;
  const int fx = min(x, l - x), fy = min(y, l - y);
  const int rr = rbin(fx, fy);
  RefWeight w{};
  w.idx = e;
  w.band = band_template(rr, p.hr_t, p.lr_t);
  const bool ctf = rr < l / 2 && rr >= 0;
  w.flags = ((x <= l / 2 && ctf) ? 1 : 0) | (ctf ? 2 : 0);
  if (ctf) {
    const double r = hypot((double)fy, (double)fx);
    const double ss = r * (double)p.ds;
    const double signal = exp(p.bf * ss * ss + p.bf2 * ss + p.bf3) / ((double)p.kk + 1);
    w.N = exp(p.a * ss * ss + p.b * ss + p.b2) / signal;
    dev::ctf_geometry(x, (y + l / 2) % l, l, l, p, &w.s2, &w.cs_ast);
    const double lam = p.lambda;
    w.a0 = p.cs * 2.5e6 * lam * lam * lam * w.s2 * w.s2;
  }
  tab[i] = w;
}

// Weight f and df/d ddf of one entry at defocus offset ddf: f = band |v| g with
// g = (N + kk v^2)^-1/2 and df = band sign(v) v' N g^3 (v = CTF, v' = dv/d ddf).
__device__ __forceinline__ void entry_weight(const RefWeight& w, const DevParams& p, double ddf, double* f,
                                             double* df) {
  *f = w.band;
  *df = 0.0;
  if (!(w.flags & 2) || w.band == 0.0) return;
  const double dfu = p.dfu - ddf * 1e-4, dfv = p.dfv - ddf * 1e-4;
  const double df_ast = 0.5 * (dfu + dfv + 2 * p.dfdiff * w.cs_ast);
  const double lam = p.lambda;
  const double gamma = -2 * kRefPi * (w.a0 + df_ast * 5000.0 * lam * w.s2);
  double sg, cg;
  sincos(gamma, &sg, &cg);
  const double B = sqrt(1.0 - p.ampconst * p.ampconst);
  const double v = B * sg + p.ampconst * cg;
  const double dv = (B * cg - p.ampconst * sg) * (kRefPi * lam * w.s2);
  const double g = sqrt(1 / (w.N + (double)p.kk * v * v));
  *f = w.band * fabs(v) * g;
  *df = w.band * (v > 0 ? 1.0 : (v < 0 ? -1.0 : 0.0)) * dv * w.N * g * g * g;
}

// Partial sums of |G|^2 and d|G|^2/d ddf over the normalisation region (x <= l/2, radial bin
// < l/2) of compute_area_sum_ofSQR (k:319-382); grid (kPowerChunks, K).
__global__ void k_ref_power(const float2* __restrict__ white, const RefWeight* __restrict__ tab, int ntab,
                            const RefPeak* __restrict__ pk, int l, DevParams p, double2* __restrict__ partial) {
  const int j = blockIdx.y, ch = blockIdx.x;
  const RefPeak P = pk[j];
  const float2* F = white + (long long)P.t * l * l;
  const int per = (ntab + kPowerChunks - 1) / kPowerChunks;
  const int i0 = ch * per, i1 = min(ntab, i0 + per);
  double s = 0, ds = 0;
  for (int i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
    const RefWeight w = tab[i];
    if (!(w.flags & 1)) continue;
    double f, df;
    entry_weight(w, p, P.ddf, &f, &df);
    const float2 v = F[w.idx];
    const double a2 = (double)v.x * v.x + (double)v.y * v.y;
    s += f * f * a2;
    ds += 2 * f * df * a2;
  }
  s = block_sum(s);
  ds = block_sum(ds);
  if (threadIdx.x == 0) partial[(long long)j * kPowerChunks + ch] = make_double2(s, ds);
}

// H = G/m + i (dG/m - G dm/m^2), m = sqrt(sum |G|^2) / count (normalize, k:445-458); the
// region count is the number of half-plane entries with radial bin < l/2.  Only the table
// entries (the band-pass support) are written; grid (table chunks, K).
__global__ void k_ref_pack(const float2* __restrict__ white, const RefWeight* __restrict__ tab, int ntab,
                           const RefPeak* __restrict__ pk, int l, DevParams p,
                           const double2* __restrict__ partial, double count, float2* __restrict__ H) {
  const int j = blockIdx.y;
  const RefPeak P = pk[j];
  __shared__ double sm[2];
  if (threadIdx.x == 0) {
    double s = 0, ds = 0;
    for (int c = 0; c < kPowerChunks; ++c) {
      s += partial[(long long)j * kPowerChunks + c].x;
      ds += partial[(long long)j * kPowerChunks + c].y;
    }
    const double m = sqrt(s) / count;
    sm[0] = m;
    sm[1] = ds / (2 * sqrt(s) * count);
  }
  __syncthreads();
  const double m = sm[0], dm = sm[1];
  const long long L2 = (long long)l * l;
  const float2* F = white + (long long)P.t * L2;
  float2* out = H + (long long)j * L2;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ntab) return;
  const RefWeight w = tab[i];
  float2 h = make_float2(0.f, 0.f);
  double f, df;
  entry_weight(w, p, P.ddf, &f, &df);
  if (f != 0.0 || df != 0.0) {
    const float2 v = F[w.idx];
    const double a = f / m, b = df / m - f * dm / (m * m);
    // (a + i b) * v
    h = make_float2((float)(a * v.x - b * v.y), (float)(a * v.y + b * v.x));
  }
  out[w.idx] = h;
}

__global__ void k_ref_clear(float2* __restrict__ H, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    H[i] = make_float2(0.f, 0.f);
}

// Rotation (float geometry as the search), bilinear derivative, mask; one kRotTX x kRotTY output
// tile of one pick per block.  Writes A = r0 + i dr0/dtheta and B = dr0/d ddf (before the box
// standardisation, which is affine per pick and is applied to their spectra in k_ref_dot) and
// the tile's six double sums.  The B fields of picks 2m and 2m+1 share plane K + m (real and
// imaginary part: both are real fields, so one forward transform gives both spectra); the last
// pick of an odd K writes a zero imaginary part.
__global__ void __launch_bounds__(256)
    k_ref_rotate(const float2* __restrict__ H, const RefPeak* __restrict__ pk, int l, float rin,
                 float rout, float ehw, float2* __restrict__ AB, double* __restrict__ partial6, int K) {
  const int g = blockIdx.x, j = blockIdx.y, groups = gridDim.x;
  const int gx = (l + kRotTX - 1) / kRotTX;
  const RefPeak P = pk[j];
  const float cose = (float)cos(P.theta * kRefPi / 180);
  const float sine = (float)sin(P.theta * kRefPi / 180);
  const float deg = (float)(kRefPi / 180);
  const long long L2 = (long long)l * l;
  const float2* T = H + (long long)j * L2;
  float2* A = AB + (long long)j * L2;                                  // r0 + i dr0/dtheta
  float* Bf = reinterpret_cast<float*>(AB + ((long long)K + j / 2) * L2);  // dr0/d ddf, part j & 1
  const bool lone = (K & 1) && j == K - 1;
  const int i = (g % gx) * kRotTX + threadIdx.x % kRotTX;
  double a[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
  for (int h = 0; h < kRotTY / 8; ++h) {
    const int jrow = (g / gx) * kRotTY + threadIdx.x / kRotTX + 8 * h;
    if (i >= l || jrow >= l) continue;
    float v = 0.f, vth = 0.f, vdf = 0.f;
    const Bil b = bil_coords(i, jrow, l, cose, sine);
    if (b.ok) {
      const float tt = 1.0f - b.t, uu = 1.0f - b.u;
      const float2 h0 = T[b.k0], h1 = T[b.k1], h2 = T[b.k2], h3 = T[b.k3];
      v = h0.x * tt * uu + h1.x * b.t * uu + h2.x * b.t * b.u + h3.x * tt * b.u;
      vdf = h0.y * tt * uu + h1.y * b.t * uu + h2.y * b.t * b.u + h3.y * tt * b.u;
      const float dvx = uu * (h1.x - h0.x) + b.u * (h2.x - h3.x);
      const float dvy = tt * (h3.x - h0.x) + b.t * (h2.x - h1.x);
      const float x = (float)(i - l / 2), y = (float)(jrow - l / 2);
      vth = (dvx * (-sine * x + cose * y) + dvy * (-cose * x - sine * y)) * deg;
    }
    const float m = dev::mask_factor(i, jrow, l, rin, rout, ehw);
    v *= m;
    vth *= m;
    vdf *= m;
    const long long q = (long long)jrow * l + i;
    A[q] = make_float2(v, vth);
    if (lone)
      reinterpret_cast<float2*>(Bf)[q] = make_float2(vdf, 0.f);
    else
      Bf[2 * q + (j & 1)] = vdf;
    a[0] += v;
    a[1] += (double)v * v;
    a[2] += vth;
    a[3] += (double)v * vth;
    a[4] += vdf;
    a[5] += (double)v * vdf;
  }
  // the six sums reduced together: wavefront shuffles, then one LDS exchange
  __shared__ double red[4][6];
#pragma unroll
  for (int k = 0; k < 6; ++k)
    for (int o = 32; o > 0; o >>= 1) a[k] += __shfl_down(a[k], o, 64);
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < 6; ++k) red[wave][k] = a[k];
  __syncthreads();
  if (threadIdx.x < 6) {
    const int k = threadIdx.x;
    partial6[((long long)j * groups + g) * 6 + k] = red[0][k] + red[1][k] + red[2][k] + red[3][k];
  }
}

// Row transforms of the plane passes (the same lanes per row and radix plan as corr.hip's
// k_plane_rows, so the rows below are bit-identical to its output).
template <int L>
struct RefRowCfg;
template <>
struct RefRowCfg<320> {
  static constexpr int T = 16;
  using Plan = ct::RadixList<4, 4, 4, 5>;
};
template <>
struct RefRowCfg<720> {
  static constexpr int T = 32;
  using Plan = ct::RadixList<4, 4, 5, 9>;
};

// Row layout of the A and B planes between k_ref_rotate_rows and k_ref_cols_dot: position
// q = 2 NU g + s of a row holds column NU g + s (s < NU) or that column's mirror (s >= NU;
// l - u, and l/2 for u = 0), so that the NU column pairs of a k_ref_cols_dot workgroup are one
// contiguous run of 2 NU values per row (128 bytes at l = 320) rather than two runs of NU.
template <int L>
__device__ __forceinline__ int ref_pair_column(int q) {
  constexpr int NU = 256 / RefRowCfg<L>::T / 2;
  const int g = q / (2 * NU), s = q % (2 * NU);
  const int u = g * NU + s % NU;
  return s < NU ? u : (u == 0 ? L / 2 : L - u);
}

// k_ref_rotate and the row pass of the forward transform in one kernel: one workgroup = RPB rows
// of one output plane.  For an A plane (pick p) it samples r and dr/dtheta of its rows, for a B
// plane m = p - K the defocus-derivative fields of picks 2m and 2m + 1 (the same float
// expressions as k_ref_rotate), transforms the rows in LDS and writes them; the planes between
// the rotation and the row pass are never stored.  The six box sums go out per (pick, row group):
// sums 0-3 from the A plane's workgroup, sums 4-5 from the B plane's.
template <int L>
__global__ void __launch_bounds__(256)
    k_ref_rotate_rows(const float2* __restrict__ H, const RefPeak* __restrict__ pk,
                      const float* __restrict__ mask, const float2* __restrict__ twL, float2* __restrict__ AB,
                      double* __restrict__ partial6, int K, int ylo, int yhi) {
  constexpr int T = RefRowCfg<L>::T, RPW = 64 / T, RPB = 4 * RPW, NG = L / RPB;
  static_assert(L % RPB == 0, "row groups");
  // l = 320: the row transform as 4*4*20 with twiddles formed per lane by recurrence from two
  // base twiddles in registers, and the box sums reduced through buf after the rows are written,
  // so that the workgroup's LDS is the 40 KB of buf alone (four workgroups per CU)
  constexpr bool RT = L == 320;
  using RowPlan = std::conditional_t<RT, ct::RadixList<4, 4, 20>, typename RefRowCfg<L>::Plan>;
  __shared__ float2 buf[RPB * L];
  __shared__ float2 tw[RT ? 1 : L];
  __shared__ double red[RT ? 1 : 4][4];
  const int tid = threadIdx.x, g = blockIdx.x, p = blockIdx.y;
  const int y0 = g * RPB;
  const long long L2 = (long long)L * L;
  const float deg = (float)(kRefPi / 180);
  const bool aplane = p < K;
  const int j0 = aplane ? p : 2 * (p - K);
  const bool two = !aplane && j0 + 1 < K;
  if (y0 > yhi || y0 + RPB - 1 < ylo) {  // rows outside the mask's support: zero, left unwritten
    if (tid < 4) {
      if (aplane)
        partial6[((long long)j0 * NG + g) * 6 + tid] = 0.0;
      else if (tid < 2 || two)
        partial6[((long long)(j0 + tid / 2) * NG + g) * 6 + 4 + tid % 2] = 0.0;
    }
    return;  // uniform per workgroup
  }
  // rotation of the plane's pick(s)
  float cs[2], sn[2];
  const float2* Tp[2];
  for (int q = 0; q < 2; ++q) {
    const int j = min(j0 + q, K - 1);
    const RefPeak P = pk[j];
    cs[q] = (float)cos(P.theta * kRefPi / 180);
    sn[q] = (float)sin(P.theta * kRefPi / 180);
    Tp[q] = H + (long long)j * L2;
  }
  dev::RegStage<RT ? 1 : L> twr;
  float2 wl[ct::lane_base_count(RowPlan{})];
  if constexpr (RT)
    ct::load_lane_bases<L, T>(wl, twL, (tid % 64) % T, RowPlan{});
  else
    twr.load(twL, tid);
  double a[4] = {0, 0, 0, 0};
  // work items n = tid + 256 k in batches of BT, all gathers of a batch in flight together; the
  // mask factors of the next batch are loaded while the current one is processed.  Item n is
  // pixel sq(n): each wavefront's 64 items are an 8 x 8 square of the workgroup's rows, so that
  // its bilinear taps fall on a compact patch of the source plane at every angle.
  constexpr int N = RPB * L, NE = (N + 255) / 256, BT = RT ? 3 : 4;
  static_assert(RPB % 8 == 0 && L % 8 == 0, "8 x 8 squares");
  auto sq = [](int n) {
    const int s = n / 64, ln = n % 64, sr = s / (L / 8), sc = s - sr * (L / 8);
    return (8 * sr + ln / 8) * L + 8 * sc + ln % 8;
  };
  const float* Mrows = mask + (long long)y0 * L;
  float mnext[BT];
#pragma unroll
  for (int b = 0; b < BT; ++b) {
    const int n = tid + 256 * b;
    mnext[b] = (b < NE && n < N) ? Mrows[sq(n)] : 0.f;
  }
  for (int k0 = 0; k0 < NE; k0 += BT) {
    float2 h[BT][2][4];
    float bt[BT][2], bu[BT][2], mk[BT];
    bool act[BT][2];
#pragma unroll
    for (int b = 0; b < BT; ++b) mk[b] = mnext[b];
#pragma unroll
    for (int b = 0; b < BT; ++b) {
      const int e = sq(tid + 256 * (k0 + b));
      const int r = e / L, i = e - r * L, jrow = y0 + r;
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        act[b][q] = false;
        // where the mask is zero every field is zero: no samples (the sums gain nothing either)
        if ((q == 1 && !two) || mk[b] == 0.f) continue;
        const Bil c = bil_coords(i, jrow, L, cs[q], sn[q]);
        act[b][q] = c.ok;
        bt[b][q] = c.t;
        bu[b][q] = c.u;
        if (c.ok) {
          const float2* Tq = Tp[q];
          h[b][q][0] = Tq[c.k0];
          h[b][q][1] = Tq[c.k1];
          h[b][q][2] = Tq[c.k2];
          h[b][q][3] = Tq[c.k3];
        }
      }
    }
#pragma unroll
    for (int b = 0; b < BT; ++b) {
      const int n = tid + 256 * (k0 + BT + b);
      mnext[b] = (k0 + BT + b < NE && n < N) ? Mrows[sq(n)] : 0.f;
    }
#pragma unroll
    for (int b = 0; b < BT; ++b) {
      const int n = tid + 256 * (k0 + b);
      if (!(k0 + b < NE && n < N)) continue;
      const int e = sq(n);
      const int r = e / L, i = e - r * L, jrow = y0 + r;
      const float m = mk[b];
      float2 out = make_float2(0.f, 0.f);
      const int nq = m == 0.f ? 0 : (aplane ? 1 : (two ? 2 : 1));
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        if (q >= nq) continue;
        float v = 0.f, vth = 0.f, vdf = 0.f;
        if (act[b][q]) {
          const float t = bt[b][q], u = bu[b][q];
          const float tt = 1.0f - t, uu = 1.0f - u;
          const float2 h0 = h[b][q][0], h1 = h[b][q][1], h2 = h[b][q][2], h3 = h[b][q][3];
          v = h0.x * tt * uu + h1.x * t * uu + h2.x * t * u + h3.x * tt * u;
          vdf = h0.y * tt * uu + h1.y * t * uu + h2.y * t * u + h3.y * tt * u;
          if (aplane) {
            const float dvx = uu * (h1.x - h0.x) + u * (h2.x - h3.x);
            const float dvy = tt * (h3.x - h0.x) + t * (h2.x - h1.x);
            const float x = (float)(i - L / 2), y = (float)(jrow - L / 2);
            vth = (dvx * (-sn[q] * x + cs[q] * y) + dvy * (-cs[q] * x - sn[q] * y)) * deg;
          }
        }
        v *= m;
        vth *= m;
        vdf *= m;
        if (aplane) {
          out = make_float2(v, vth);
          a[0] += v;
          a[1] += (double)v * v;
          a[2] += vth;
          a[3] += (double)v * vth;
        } else {
          if (q == 0)
            out.x = vdf;
          else
            out.y = vdf;
          a[2 * q] += vdf;
          a[2 * q + 1] += (double)v * vdf;
        }
      }
      buf[e] = out;
    }
  }
  const int wave = tid / 64, lane = tid % 64;
#pragma unroll
  for (int k = 0; k < 4; ++k)
    for (int o = 32; o > 0; o >>= 1) a[k] += __shfl_down(a[k], o, 64);
  // the box sums of the workgroup (wave partials in lane 0 of each wave)
  auto store_sums = [&](double (*rd)[4]) {
    if (lane == 0)
#pragma unroll
      for (int k = 0; k < 4; ++k) rd[wave][k] = a[k];
    __syncthreads();
    if (tid < 4) {
      const double v = rd[0][tid] + rd[1][tid] + rd[2][tid] + rd[3][tid];
      if (aplane)
        partial6[((long long)j0 * NG + g) * 6 + tid] = v;
      else if (tid < 2 || two)
        partial6[((long long)(j0 + tid / 2) * NG + g) * 6 + 4 + tid % 2] = v;
    }
  };
  if constexpr (RT) {
    __syncthreads();
    ct::run_stockham<L, T, -1, ct::Sync::WAVE, ct::IdentityMap, ct::TwSrc::RECUR>(
        buf + (wave * RPW + lane / T) * L, wl, lane % T, RowPlan{});
  } else {
    twr.store(tw, tid);
    store_sums(red);
    ct::run_stockham<L, T, -1, ct::Sync::WAVE>(buf + (wave * RPW + lane / T) * L, tw, lane % T, RowPlan{});
  }
  __syncthreads();
  // rows written in the column-pair order of ref_pair_column
  float2* dst = AB + (long long)p * L2 + (long long)y0 * L;
  for (int e = tid; e < RPB * L; e += 256) {
    const int r = e / L, q = e - r * L;
    dst[e] = buf[r * L + ref_pair_column<L>(q)];
  }
  if constexpr (RT) {
    __syncthreads();  // buf's rows are written: its first bytes hold the reduction
    store_sums(reinterpret_cast<double(*)[4]>(buf));
  }
}

// The soft mask of the refinement's rotation (apply_mask, k:225-241, with the search's float
// evaluation) as an l x l table.
__global__ void k_ref_mask_table(int l, float rin, float rout, float ehw, float* __restrict__ M) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= l * l) return;
  M[e] = dev::mask_factor(e % l, e / l, l, rin, rout, ehw);
}

// Per pick: the search's float mean/variance (k_rotation_stats) and the derivatives of the
// mean and standard deviation.  stats[j] = {meanf, varf, mean, sd, dmean_th, dsd_th, dmean_df, dsd_df}
__global__ void k_ref_stats(const double* __restrict__ partial6, int groups, int l, int K,
                            double* __restrict__ stats) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= K) return;
  double s[6] = {0, 0, 0, 0, 0, 0};
  for (int g = 0; g < groups; ++g)
    for (int k = 0; k < 6; ++k) s[k] += partial6[((long long)j * groups + g) * 6 + k];
  const double L2 = (double)l * l;
  const double mean = s[0] / L2, var = s[1] / L2 - mean * mean, sd = sqrt(var);
  const double dmt = s[2] / L2, dvt = 2 * (s[3] / L2 - mean * dmt);
  const double dmd = s[4] / L2, dvd = 2 * (s[5] / L2 - mean * dmd);
  double* o = stats + (long long)j * 8;
  o[0] = (double)(float)mean;
  o[1] = (double)(float)var;
  o[2] = mean;
  o[3] = sd;
  o[4] = dmt;
  o[5] = dvt / (2 * sd);
  o[6] = dmd;
  o[7] = dvd / (2 * sd);
}

// Band-limited shift factor ph(k, q) = exp(2 pi i k_s q / l) (cos(pi q) at Nyquist) and d/dq.
__device__ __forceinline__ void phase(int k, int l, int pint, double frac, float2& ph, float2& dph) {
  if (2 * k == l) {
    const double q = pint + frac;
    double s, c;
    sincospi(q, &s, &c);
    ph = make_float2((float)c, 0.f);
    dph = make_float2((float)(-M_PI * s), 0.f);
    return;
  }
  const int ks = k < l / 2 ? k : k - l;
  // angle / (2 pi) = (ks * pint mod l) / l + ks * frac / l, reduced exactly in integers first
  long long m = ((long long)ks * pint) % l;
  if (m < 0) m += l;
  const double turns = (double)m / l + (double)ks * frac / l;
  double s, c;
  sincospi(2.0 * turns, &s, &c);
  ph = make_float2((float)c, (float)s);
  const double w = 2 * M_PI * ks / l;
  dph = make_float2((float)(-w * s), (float)(w * c));
}

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// c and its derivatives for one pick per block.  The half spectrum (ky rows of HX entries) is
// walked flat so that consecutive lanes read consecutive kx of A, B and of the block spectrum,
// which is given transposed: ihat_t[b][ky][kx].
// LT = 0: any even l = lrt, with the phase tables in dynamic LDS (2 (l/2 + 1) + 2 l entries).
template <int LT>
__global__ __launch_bounds__(256) void k_ref_dot(const float2* __restrict__ ihat_t, const float2* __restrict__ AB,
                                                 const RefPeak* __restrict__ pk, const double* __restrict__ stats,
                                                 int K, double* __restrict__ out, int lrt) {
  const int L = LT ? LT : lrt;
  const int HX = L / 2 + 1;
  constexpr int SX = LT ? LT / 2 + 1 : 1, SY = LT ? LT : 1;
  __shared__ float2 sphx[SX], sdphx[SX], sphy[SY], sdphy[SY];
  extern __shared__ float2 dyn[];
  float2* const phx = LT ? sphx : dyn;
  float2* const dphx = LT ? sdphx : dyn + HX;
  float2* const phy = LT ? sphy : dyn + 2 * HX;
  float2* const dphy = LT ? sdphy : dyn + 2 * HX + L;
  const int j = blockIdx.x, part = blockIdx.y;
  const RefPeak P = pk[j];
  const int px = P.pixel % L, py = P.pixel / L;
  for (int k = threadIdx.x; k < HX; k += blockDim.x) phase(k, L, px, P.dx, phx[k], dphx[k]);
  for (int k = threadIdx.x; k < L; k += blockDim.x) phase(k, L, py, P.dy, phy[k], dphy[k]);
  // box standardisation of the pick (substract_by_mean + divided_by_var, k:460-477, with the
  // float mean and variance the search uses) and its theta and defocus derivatives, applied to
  // the spectra: a constant c only moves the DC term (transforms are scaled by 1/l^2)
  const double* st = stats + (long long)j * 8;
  const float meanf = (float)st[0], varf = (float)st[1];
  const float mean = (float)st[2], sd = (float)st[3];
  const float dmt = (float)st[4], dst = (float)st[5], dmd = (float)st[6], dsd = (float)st[7];
  const float inv_s = varf != 0 ? 1.0f / sqrtf(varf) : 1.0f;
  const float inv = 1.0f / sd, inv2 = 1.0f / (sd * sd);
  __syncthreads();
  const long long L2 = (long long)L * L;
  const float2* I = ihat_t + (long long)P.block * HX * L;
  const float2* A = AB + (long long)j * L2;
  const float2* B = AB + ((long long)K + j / 2) * L2;  // shared with pick j ^ 1
  const bool odd = j & 1;
  double acc[5] = {0, 0, 0, 0, 0};
  // this block's share [elo, ehi) of the half spectrum; elements e = elo + threadIdx.x + 256 k in
  // batches of BT, the loads of a batch in flight together
  constexpr int BT = 4;
  const int NT = (int)(((long long)L * HX * (part + 1)) / kRefDotSplit);
  const int elo = (int)(((long long)L * HX * part) / kRefDotSplit);
  for (int e0 = elo + threadIdx.x; e0 < NT; e0 += BT * 256) {
    float2 va[BT], vam[BT], vb[BT], vbm[BT], vx[BT];
#pragma unroll
    for (int q = 0; q < BT; ++q) {
      const int e = e0 + 256 * q;
      if (e < NT) {
        const int ky = e / HX, kx = e - ky * HX;
        const int my = ky ? L - ky : 0, mx = kx ? L - kx : 0;
        va[q] = A[(long long)ky * L + kx];
        vam[q] = A[(long long)my * L + mx];
        vb[q] = B[(long long)ky * L + kx];
        vbm[q] = B[(long long)my * L + mx];
        vx[q] = I[e];
      }
    }
#pragma unroll
    for (int q = 0; q < BT; ++q) {
      const int e = e0 + 256 * q;
      if (e >= NT) break;
      const int ky = e / HX, kx = e - ky * HX;
      const float2 a = va[q], am = vam[q], b = vb[q], bm = vbm[q];
      // spectra of the three raw fields r0, dr0/dtheta, dr0/d ddf
      float2 R = make_float2(0.5f * (a.x + am.x), 0.5f * (a.y - am.y));
      float2 Gt = make_float2(0.5f * (a.y + am.y), -0.5f * (a.x - am.x));
      float2 Gd = odd ? make_float2(0.5f * (b.y + bm.y), 0.5f * (bm.x - b.x))
                      : make_float2(0.5f * (b.x + bm.x), 0.5f * (b.y - bm.y));
      float2 Rc = R;  // r0 - mean
      if (e == 0) {
        R.x -= meanf;
        Rc.x -= mean;
        Gt.x -= dmt;
        Gd.x -= dmd;
      }
      const float2 T = make_float2(R.x * inv_s, R.y * inv_s);
      const float2 Tt = make_float2(Gt.x * inv - Rc.x * dst * inv2, Gt.y * inv - Rc.y * dst * inv2);
      const float2 Td = make_float2(Gd.x * inv - Rc.x * dsd * inv2, Gd.y * inv - Rc.y * dsd * inv2);
      float2 X = vx[q];
      if ((kx + ky) & 1) X = make_float2(-X.x, -X.y);
      const float w = (kx == 0 || 2 * kx == L) ? 1.f : 2.f;
      const float2 ey = phy[ky], dey = dphy[ky];
      const float2 E = cmul(phx[kx], ey);
      const float2 Ex = cmul(dphx[kx], ey);
      const float2 Ey = cmul(phx[kx], dey);
      // X conj(T) and its partners
      const float2 P0 = make_float2(X.x * T.x + X.y * T.y, X.y * T.x - X.x * T.y);
      const float2 Pt = make_float2(X.x * Tt.x + X.y * Tt.y, X.y * Tt.x - X.x * Tt.y);
      const float2 Pd = make_float2(X.x * Td.x + X.y * Td.y, X.y * Td.x - X.x * Td.y);
      acc[0] += w * (P0.x * E.x - P0.y * E.y);
      acc[1] += w * (P0.x * Ex.x - P0.y * Ex.y);
      acc[2] += w * (P0.x * Ey.x - P0.y * Ey.y);
      acc[3] += w * (Pt.x * E.x - Pt.y * E.y);
      acc[4] += w * (Pd.x * E.x - Pd.y * E.y);
    }
  }
  __shared__ double red[4][5];
#pragma unroll
  for (int k = 0; k < 5; ++k)
    for (int o = 32; o > 0; o >>= 1) acc[k] += __shfl_down(acc[k], o, 64);
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < 5; ++k) red[wave][k] = acc[k];
  __syncthreads();
  const double c = red[0][0] + red[1][0] + red[2][0] + red[3][0];
  const double gx = red[0][1] + red[1][1] + red[2][1] + red[3][1];
  const double gy = red[0][2] + red[1][2] + red[2][2] + red[3][2];
  const double gt = red[0][3] + red[1][3] + red[2][3] + red[3][3];
  const double gd = red[0][4] + red[1][4] + red[2][4] + red[3][4];
  if (threadIdx.x == 0) {
    double* o = out + ((long long)j * kRefDotSplit + part) * 5;
    o[0] = c;
    o[1] = gx;
    o[2] = gy;
    o[3] = gt;
    o[4] = gd;
  }
}

// Per pick the shift phase factors of k_ref_dot: ph[j] = {phase(ky) for ky < l, its derivative
// for ky < l, phase(kx) for kx <= l/2, its derivative}, 2 l + 2 (l/2 + 1) entries; grid (., K).
__global__ void k_ref_phases(const RefPeak* __restrict__ pk, int l, float2* __restrict__ ph) {
  const int j = blockIdx.y, HX = l / 2 + 1, n = 2 * l + 2 * HX;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= l + HX) return;
  const RefPeak P = pk[j];
  float2* o = ph + (long long)j * n;
  if (e < l)
    phase(e, l, P.pixel / l, P.dy, o[e], o[l + e]);
  else
    phase(e - l, l, P.pixel % l, P.dx, o[2 * l + e - l], o[2 * l + HX + e - l]);
}

// The forward column pass of the A and B planes fused with k_ref_dot (l = 320, 720): one
// workgroup = NU mirrored column pairs (u, l - u; pair 0 = columns 0 and l/2) of the three planes
// of a pick pair (B_m, A_2m, A_2m+1), transformed in LDS one plane at a time, each plane's
// spectrum consumed where it is made: the transformed planes never reach memory.  The defocus
// term of the dot is linear in the B and the A spectrum separately,
//   Re[X conj(Gd inv - Rc dsd inv2) E] = inv Re[X conj(Gd) E] - dsd inv2 Re[X conj(Rc) E],
// so the B pass accumulates the first part for both picks and each A pass the second.  Rows
// outside [ylo, yhi] of AB are zero and not read; ihat is the block half spectra [b][kx][ky].
// out[(j * G + g) * 5 + k]: the five partial sums of pick j over workgroup g's columns.
template <int L>
__global__ void __launch_bounds__(256)
    k_ref_cols_dot(const float2* __restrict__ ihat, const float2* __restrict__ AB, const RefPeak* __restrict__ pk,
                   const double* __restrict__ stats, const float2* __restrict__ ph, const float2* __restrict__ twL,
                   int K, int ylo, int yhi, float scale, double* __restrict__ out) {
  constexpr int T = RefRowCfg<L>::T, COLS = 256 / T, NU = COLS / 2, LP = L + 1, HX = L / 2 + 1;
  constexpr int G = (L / 2) / NU;
  static_assert((L / 2) % NU == 0 && COLS % 2 == 0, "column pairs");
  constexpr int NO = (NU * L + L + 255) / 256;  // owned half-spectrum elements per thread, at most
  __shared__ float2 buf[COLS * LP];
  __shared__ float2 tw[L];
  __shared__ float2 phy[2][L], dphy[L];
  __shared__ float2 phx[2][NU + 1], dphx[2][NU + 1];
  __shared__ double red[4][5];
  const int tid = threadIdx.x, g = blockIdx.x, m = blockIdx.y;
  const int j0 = 2 * m;
  const int nq = j0 + 1 < K ? 2 : 1;
  const long long L2 = (long long)L * L;
  // slot cc < NU: column u = g NU + cc; slot NU + v: the mirror column of u = g NU + v
  auto col_of = [&](int cc) {
    const int u = g * NU + cc % NU;
    return cc < NU ? u : (u == 0 ? L / 2 : L - u);
  };
  // owned elements o = tid + 256 k: slot o / L (the extra slot NU, column l/2, in workgroup 0), ky = o % L
  const int nown = NU * L + (g == 0 ? L : 0);
  const float2* const Ib0 = ihat + (long long)pk[j0].block * HX * L;
  const float2* const Ib1 = ihat + (long long)pk[min(j0 + 1, K - 1)].block * HX * L;
  // block spectrum X of pick q at owned element o (clamped into the owned range, so that the
  // loads of a batch need no branch)
  auto xat = [&](int q, int o) __attribute__((always_inline)) {
    o = min(o, nown - 1);
    return (q == 0 ? Ib0 : Ib1)[col_of(o / L) * L + o % L];
  };
  dev::RegStage<L> twr;
  twr.load(twL, tid);
  // phase factors of both picks from the tables of k_ref_phases (shift derivative along y: pick
  // j0 first)
  for (int q = 0; q < nq; ++q) {
    const float2* P = ph + (long long)(j0 + q) * (2 * L + 2 * HX);
    for (int k = tid; k < L; k += 256) {
      phy[q][k] = P[k];
      if (q == 0) dphy[k] = P[L + k];
    }
    for (int sl = tid; sl <= NU; sl += 256) {
      const int kx = sl < NU ? g * NU + sl : L / 2;
      phx[q][sl] = P[2 * L + kx];
      dphx[q][sl] = P[2 * L + HX + kx];
    }
  }
  // Each pass works from its own copy of the thread index, opaque to the compiler: otherwise the
  // per-thread LDS and global addresses of the transform and of the element loops, identical in
  // every pass, are computed once and held in registers for the whole kernel (some seventy more).
  auto pass_tid = [&]() __attribute__((always_inline)) {
    int t = tid;
    asm volatile("" : "+v"(t));
    return t;
  };
  // a thread's elements of a plane lie in one column (256 is a multiple of COLS), rows YS apart;
  // the workgroup's COLS columns are contiguous in each row of AB (ref_pair_column order)
  constexpr int NE = (COLS * L + 255) / 256, YS = 256 / COLS;
  static_assert(COLS == 2 * NU && 2 * NU * G == L, "pair runs tile the row");
  auto load_plane = [&](const float2* P, int t) __attribute__((always_inline)) {
    const int lcol = t % COLS, ybase = t / COLS;
    const float2* src = P + COLS * g + lcol;
    float2 v[NE];
#pragma unroll
    for (int k = 0; k < NE; ++k) {
      const int y = ybase + YS * k;
      v[k] = (y < L && y >= ylo && y <= yhi) ? src[y * L] : make_float2(0.f, 0.f);
    }
#pragma unroll
    for (int k = 0; k < NE; ++k) {
      const int y = ybase + YS * k;
      if (y < L) buf[lcol * LP + y] = v[k];
    }
  };
  auto transform = [&](int t) __attribute__((always_inline)) {
    __syncthreads();
    ct::run_stockham<L, T, -1, ct::Sync::WAVE>(buf + (t / T) * LP, tw, t % T, typename RefRowCfg<L>::Plan{});
    __syncthreads();
  };
  // element o: its value and its mirror's (-kx, -ky), scaled as the plane transform scales
  auto pair_at = [&](int o, float2& v, float2& vm) {
    const int cc = o / L, ky = o % L, kx = col_of(cc);
    const int mc = (cc < NU && kx != 0) ? cc + NU : cc, my = ky ? L - ky : 0;
    v = buf[cc * LP + ky];
    vm = buf[mc * LP + my];
    v = make_float2(v.x * scale, v.y * scale);
    vm = make_float2(vm.x * scale, vm.y * scale);
  };
  auto reduce_store = [&](double (&acc)[5], int j) {
#pragma unroll
    for (int k = 0; k < 5; ++k)
      for (int o = 32; o > 0; o >>= 1) acc[k] += __shfl_down(acc[k], o, 64);
    const int wave = tid / 64, lane = tid % 64;
    if (lane == 0)
#pragma unroll
      for (int k = 0; k < 5; ++k) red[wave][k] = acc[k];
    __syncthreads();
    if (tid < 5) out[((long long)j * G + g) * 5 + tid] = red[0][tid] + red[1][tid] + red[2][tid] + red[3][tid];
  };

  // B pass: Re[X conj(Gd) E] of both picks, owned elements in batches of EB with the
  // block-spectrum loads of a batch in flight together
  const int tb = pass_tid();
  load_plane(AB + ((long long)K + m) * L2, tb);
  twr.store(tw, tid);
  transform(tb);
  double sg[2] = {0, 0};
  constexpr int EB = 4;
#pragma unroll 1
  for (int k0 = 0; k0 < NO; k0 += EB) {
    float2 xb[2][EB];
#pragma unroll
    for (int k = 0; k < EB; ++k) {
      xb[0][k] = xat(0, tb + 256 * (k0 + k));
      xb[1][k] = xat(1, tb + 256 * (k0 + k));
    }
#pragma unroll
    for (int k = 0; k < EB; ++k) {
      const int o = tb + 256 * (k0 + k);
      if (o >= nown) break;
      const int cc = o / L, ky = o % L, kx = col_of(cc);
      float2 b, bm;
      pair_at(o, b, bm);
      const float w = (kx == 0 || 2 * kx == L) ? 1.f : 2.f;
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        if (q >= nq) break;
        float2 Gd = q == 0 ? make_float2(0.5f * (b.x + bm.x), 0.5f * (b.y - bm.y))
                           : make_float2(0.5f * (b.y + bm.y), 0.5f * (bm.x - b.x));
        if (o == 0 && g == 0) Gd.x -= (float)stats[(long long)(j0 + q) * 8 + 6];
        float2 X = xb[q][k];
        if ((kx + ky) & 1) X = make_float2(-X.x, -X.y);
        const float2 E = cmul(phx[q][cc], phy[q][ky]);
        const float2 Pg = make_float2(X.x * Gd.x + X.y * Gd.y, X.y * Gd.x - X.x * Gd.y);
        sg[q] += w * (Pg.x * E.x - Pg.y * E.y);
      }
    }
  }
  // A passes, one per pick
  for (int q = 0; q < nq; ++q) {
    const int j = j0 + q;
    __syncthreads();  // the previous pass has read buf (and dphy)
    if (q == 1) {
      const float2* P = ph + (long long)j * (2 * L + 2 * HX);
      for (int k = tid; k < L; k += 256) dphy[k] = P[L + k];
    }
    const int ta = pass_tid();
    load_plane(AB + (long long)j * L2, ta);
    transform(ta);
    const double* st = stats + (long long)j * 8;
    const float meanf = (float)st[0], varf = (float)st[1];
    const float mean = (float)st[2], sd = (float)st[3];
    const float dmt = (float)st[4], dst = (float)st[5], dsd = (float)st[7];
    const float inv_s = varf != 0 ? 1.0f / sqrtf(varf) : 1.0f;
    const float inv = 1.0f / sd, inv2 = 1.0f / (sd * sd);
    double acc[5] = {0, 0, 0, 0, 0};
    double sr = 0;
#pragma unroll 1
    for (int k0 = 0; k0 < NO; k0 += EB) {
      float2 xa[EB];
#pragma unroll
      for (int k = 0; k < EB; ++k) xa[k] = xat(q, ta + 256 * (k0 + k));
#pragma unroll
      for (int k = 0; k < EB; ++k) {
        const int o = ta + 256 * (k0 + k);
        if (o >= nown) break;
        const int cc = o / L, ky = o % L, kx = col_of(cc);
        float2 a, am;
        pair_at(o, a, am);
        // spectra of r0 and dr0/dtheta
        float2 R = make_float2(0.5f * (a.x + am.x), 0.5f * (a.y - am.y));
        float2 Gt = make_float2(0.5f * (a.y + am.y), -0.5f * (a.x - am.x));
        float2 Rc = R;  // r0 - mean
        if (o == 0 && g == 0) {
          R.x -= meanf;
          Rc.x -= mean;
          Gt.x -= dmt;
        }
        const float2 Tn = make_float2(R.x * inv_s, R.y * inv_s);
        const float2 Tt = make_float2(Gt.x * inv - Rc.x * dst * inv2, Gt.y * inv - Rc.y * dst * inv2);
        float2 X = xa[k];
        if ((kx + ky) & 1) X = make_float2(-X.x, -X.y);
        const float w = (kx == 0 || 2 * kx == L) ? 1.f : 2.f;
        const float2 ey = phy[q][ky], dey = dphy[ky];
        const float2 E = cmul(phx[q][cc], ey);
        const float2 Ex = cmul(dphx[q][cc], ey);
        const float2 Ey = cmul(phx[q][cc], dey);
        const float2 P0 = make_float2(X.x * Tn.x + X.y * Tn.y, X.y * Tn.x - X.x * Tn.y);
        const float2 Pt = make_float2(X.x * Tt.x + X.y * Tt.y, X.y * Tt.x - X.x * Tt.y);
        const float2 Pr = make_float2(X.x * Rc.x + X.y * Rc.y, X.y * Rc.x - X.x * Rc.y);
        acc[0] += w * (P0.x * E.x - P0.y * E.y);
        acc[1] += w * (P0.x * Ex.x - P0.y * Ex.y);
        acc[2] += w * (P0.x * Ey.x - P0.y * Ey.y);
        acc[3] += w * (Pt.x * E.x - Pt.y * E.y);
        sr += w * (Pr.x * E.x - Pr.y * E.y);
      }
    }
    acc[4] = (double)inv * sg[q] - (double)dsd * (double)inv2 * sr;
    reduce_store(acc, j);
  }
}

// z, gradient and (unless last) one ascent step p += step * scale^2 * dz/dp.
__global__ void k_ref_update(RefPeak* __restrict__ pk, int K, const double* __restrict__ cg, int parts,
                             const double* __restrict__ mu, const double* __restrict__ sigma, int l,
                             double step, double4 scale, int first, int update,
                             double* __restrict__ zout, double* __restrict__ z0out,
                             double* __restrict__ gout) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= K) return;
  RefPeak P = pk[j];
  const long long q = (long long)P.block * l * l + P.pixel;
  const double m = mu[q], s = sigma[q];
  double o[5] = {0, 0, 0, 0, 0};
  for (int part = 0; part < parts; ++part)
    for (int k = 0; k < 5; ++k) o[k] += cg[((long long)j * parts + part) * 5 + k];
  const double z = s > 0 ? (o[0] - m) / s : 0.0;
  const double inv = s > 0 ? 1.0 / s : 0.0;
  const double g0 = o[1] * inv, g1 = o[2] * inv, g2 = o[3] * inv, g3 = o[4] * inv;
  zout[j] = z;
  if (first) z0out[j] = z;
  gout[4 * j + 0] = g0;
  gout[4 * j + 1] = g1;
  gout[4 * j + 2] = g2;
  gout[4 * j + 3] = g3;
  if (update) {
    P.dx += step * scale.x * scale.x * g0;
    P.dy += step * scale.y * scale.y * g1;
    P.theta += step * scale.z * scale.z * g2;
    P.ddf += step * scale.w * scale.w * g3;
    pk[j] = P;
  }
}

inline unsigned nblk(long long n, int bs = BS) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace

// floor(hypot(dx, dy) + 0.5) - 1 from the integer squared radius (the device rbin).
static long long rbin_host(long long dx, long long dy) {
  const long long n = dx * dx + dy * dy;
  long long k = (long long)std::sqrt((double)n);
  while (k * k > n) --k;
  while ((k + 1) * (k + 1) <= n) ++k;
  if (n > k * k + k) ++k;
  return k - 1;
}

double ref_region_count(int l) {
  double cnt = 0;
  for (int y = 0; y < l; ++y)
    for (int x = 0; x <= l / 2; ++x) {
      const long long r = rbin_host(std::min(x, l - x), std::min(y, l - y));
      if (r >= 0 && r < l / 2) cnt += 1;
    }
  return cnt;
}

std::vector<int32_t> ref_band_indices(int l, float hr, float lr) {
  // band_template is non-zero only for r < hr + 8 or lr - 8 <= r < lr (r >= 0); one bin of
  // margin keeps every such entry (an extra entry gets a zero factor on the device)
  std::vector<int32_t> idx;
  for (int y = 0; y < l; ++y)
    for (int x = 0; x < l; ++x) {
      const long long r = rbin_host(std::min(x, l - x), std::min(y, l - y));
      if (r < 0) continue;
      if ((float)r < hr + 9.0f || ((float)r >= lr - 9.0f && (float)r < lr + 1.0f)) idx.push_back(y * l + x);
    }
  return idx;
}

void launch_ref_table(const int32_t* idx, int n, int l, const DevParams& p, RefWeight* tab, hipStream_t s) {
  if (n <= 0) return;
  k_ref_table<<<nblk(n, BS), BS, 0, s>>>(idx, n, l, p, tab);
  GS_CHECK_LAUNCH();
}

void launch_ref_template(const float2* white, const RefWeight* tab, int ntab, const RefPeak* pk, int K,
                         int l, const DevParams& p, double2* partial, double count, float2* H, hipStream_t s) {
  k_ref_power<<<dim3(kPowerChunks, K), BS, 0, s>>>(white, tab, ntab, pk, l, p, partial);
  GS_CHECK_LAUNCH();
  if (ntab > 0) {
    k_ref_pack<<<dim3(nblk(ntab, BS), K), BS, 0, s>>>(white, tab, ntab, pk, l, p, partial, count, H);
    GS_CHECK_LAUNCH();
  }
}

void launch_ref_clear(float2* H, long long n, hipStream_t s) {
  k_ref_clear<<<(unsigned)std::min<long long>(4096, (n + BS - 1) / BS), BS, 0, s>>>(H, n);
  GS_CHECK_LAUNCH();
}

void launch_ref_rotate(const float2* H, const RefPeak* pk, int K, int l, float d_m, float ehw,
                       double* partial6, double* stats, float2* AB, hipStream_t s) {
  const float dm = 1.5f * d_m;  // apply_mask: d_m = 1.5 * d_m (k:233)
  const float rin = dm / 2, rout = dm / 2 + 2 * ehw;
  k_ref_rotate<<<dim3(rotation_groups(l), K), 256, 0, s>>>(H, pk, l, rin, rout, ehw, AB, partial6, K);
  GS_CHECK_LAUNCH();
  k_ref_stats<<<nblk(K, 64), 64, 0, s>>>(partial6, rotation_groups(l), l, K, stats);
  GS_CHECK_LAUNCH();
}

void launch_ref_mask(int l, float d_m, float ehw, float* mask, int* ylo, int* yhi, hipStream_t s) {
  const float dm = 1.5f * d_m;  // apply_mask: d_m = 1.5 * d_m (k:233)
  const float rin = dm / 2, rout = dm / 2 + 2 * ehw;
  // rows y with |y - l/2| > rout hold only masked (zero) pixels
  *ylo = std::max(0, (int)std::floor(l / 2 - (double)rout));
  *yhi = std::min(l - 1, (int)std::ceil(l / 2 + (double)rout));
  k_ref_mask_table<<<nblk((long long)l * l, BS), BS, 0, s>>>(l, rin, rout, ehw, mask);
  GS_CHECK_LAUNCH();
}

bool launch_ref_rotate_rows(const float2* H, const RefPeak* pk, int K, int l, const float* mask, int ylo,
                            int yhi, double* partial6, double* stats, float2* AB, hipStream_t s) {
  const float2* tw = plane_twiddles(l);
  if (!tw || K <= 0) return false;
  const int planes = K + (K + 1) / 2;
  GS_REQUIRE(planes < 65536, kInvalid, "refinement batch too large");
  int groups = 0;
  if (l == 320) {
    constexpr int RPB = 4 * (64 / RefRowCfg<320>::T);
    groups = 320 / RPB;
    k_ref_rotate_rows<320><<<dim3(groups, planes), 256, 0, s>>>(H, pk, mask, tw, AB, partial6, K, ylo, yhi);
  } else if (l == 720) {
    constexpr int RPB = 4 * (64 / RefRowCfg<720>::T);
    groups = 720 / RPB;
    k_ref_rotate_rows<720><<<dim3(groups, planes), 256, 0, s>>>(H, pk, mask, tw, AB, partial6, K, ylo, yhi);
  } else {
    return false;
  }
  GS_CHECK_LAUNCH();
  k_ref_stats<<<nblk(K, 64), 64, 0, s>>>(partial6, groups, l, K, stats);
  GS_CHECK_LAUNCH();
  return true;
}

bool launch_ref_dot(const float2* ihat, const float2* AB, const RefPeak* pk, const double* stats, int K, int l,
                    double* out, hipStream_t s) {
  if (l == 320) {
    k_ref_dot<320><<<dim3(K, kRefDotSplit), 256, 0, s>>>(ihat, AB, pk, stats, K, out, l);
  } else if (l == 720) {
    k_ref_dot<720><<<dim3(K, kRefDotSplit), 256, 0, s>>>(ihat, AB, pk, stats, K, out, l);
  } else {
    if (l <= 0 || l % 2) return false;
    const size_t lds = sizeof(float2) * (size_t)(2 * (l / 2 + 1) + 2 * l);
    GS_REQUIRE(lds <= 64 * 1024, kUnsupported, "window_size too large for the refinement");
    k_ref_dot<0><<<dim3(K, kRefDotSplit), 256, lds, s>>>(ihat, AB, pk, stats, K, out, l);
  }
  GS_CHECK_LAUNCH();
  return true;
}

int ref_cols_dot_parts(int l) {
  if (l == 320) return 320 / (256 / RefRowCfg<320>::T);
  if (l == 720) return 720 / (256 / RefRowCfg<720>::T);
  return 0;
}

bool launch_ref_cols_dot(const float2* ihat, const float2* AB, const RefPeak* pk, const double* stats, int K, int l,
                         int ylo, int yhi, float scale, float2* ph, double* out, hipStream_t s) {
  const float2* tw = plane_twiddles(l);
  const int G = ref_cols_dot_parts(l);
  if (!tw || !G || K <= 0) return false;
  k_ref_phases<<<dim3(nblk(l + l / 2 + 1, BS), K), BS, 0, s>>>(pk, l, ph);
  GS_CHECK_LAUNCH();
  const dim3 grid(G, (K + 1) / 2);
  if (l == 320)
    k_ref_cols_dot<320><<<grid, 256, 0, s>>>(ihat, AB, pk, stats, ph, tw, K, ylo, yhi, scale, out);
  else
    k_ref_cols_dot<720><<<grid, 256, 0, s>>>(ihat, AB, pk, stats, ph, tw, K, ylo, yhi, scale, out);
  GS_CHECK_LAUNCH();
  return true;
}

void launch_ref_update(RefPeak* pk, int K, const double* cg, int parts, const double* mu, const double* sigma, int l,
                       double step, const double* scale4, bool first, bool update, double* z, double* z0,
                       double* grad, hipStream_t s) {
  const double4 sc = make_double4(scale4[0], scale4[1], scale4[2], scale4[3]);
  k_ref_update<<<nblk(K, 64), 64, 0, s>>>(pk, K, cg, parts, mu, sigma, l, step, sc, first ? 1 : 0, update ? 1 : 0,
                                          z, z0, grad);
  GS_CHECK_LAUNCH();
}

}  // namespace gs
