// Device helpers shared by the preprocessing and refinement kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace gs {
namespace dev {

constexpr double kRefPi = 3.1415926535;  // include/utils/constants.h:3

// floor(hypot(dx,dy)+0.5)-1 evaluated exactly on the integer squared radius.
__device__ __forceinline__ int rbin(int dx, int dy) {
  const int n = dx * dx + dy * dy;
  int k = (int)sqrtf((float)n);
  while (k * k > n) --k;
  while ((k + 1) * (k + 1) <= n) ++k;
  if (n > k * k + k) ++k;
  return k - 1;
}

// CTF_AST (src/kernels.cu:290-317): the signed value and, when dv is not null, its derivative
// with respect to a defocus offset in Angstrom added to the micrograph defocus (dfu and dfv
// both move by -offset*1e-4 um, so d gamma / d offset = PI * lambda * s^2).
// Frequency geometry of CTF_AST at (x1, y1): s^2 in 1/A^2 and the astigmatism term
// cos(2 (dfang - angle)).
__device__ __forceinline__ void ctf_geometry(int x1, int y1, int nx, int ny, const DevParams& p, double* ss,
                                             double* cs_ast) {
  y1 = x1 > nx / 2 ? ny - y1 : y1;
  x1 = min(x1, nx - x1);
  const double dx = min(x1, nx - x1);
  const double dy = y1 - ny / 2;
  *ss = (dx * dx / (double)(nx * nx) + dy * dy / (double)(ny * ny)) / ((double)p.apix * (double)p.apix);
  const double ag = atan2((double)(y1 - ny / 2), (double)x1);
  *cs_ast = cos(2 * (p.dfang * kRefPi / 180 - ag));
}

__device__ __forceinline__ double ctf_signed(int x1, int y1, int nx, int ny, const DevParams& p,
                                             double* dv) {
  double ss, cs_ast;
  ctf_geometry(x1, y1, nx, ny, p, &ss, &cs_ast);
  const double df_ast = 0.5 * (p.dfu + p.dfv + 2 * p.dfdiff * cs_ast);
  const double lam = p.lambda;
  const double gamma = -2 * kRefPi * (p.cs * 2.5e6 * lam * lam * lam * ss * ss + df_ast * 5000.0 * lam * ss);
  const double B = sqrt(1.0 - p.ampconst * p.ampconst);
  if (dv) *dv = (B * cos(gamma) - p.ampconst * sin(gamma)) * (kRefPi * lam * ss);
  return B * sin(gamma) + p.ampconst * cos(gamma);
}

// mode 0 sign / 2 magnitude / else signed.
__device__ __forceinline__ double ctf_ast(int x1, int y1, int nx, int ny, const DevParams& p, int mode) {
  const double v = ctf_signed(x1, y1, nx, ny, p, nullptr);
  if (mode == 0) return v > 0 ? 1.0 : -1.0;
  if (mode == 2) return fabs(v);
  return v;
}

__device__ __forceinline__ double band_template(int r, float hr, float lr) {  // k:263-275
  if (r < hr && r >= 0) return 1.0;
  if (r >= hr && r < hr + 8) return 0.5 * cos(kRefPi * (r - (double)hr) / 16) + 0.5;
  if (r >= (lr - 8) && r < lr && r >= 0) return 0.5 * cos(kRefPi * ((double)lr - r) / 16) + 0.5;
  return 0.0;
}

__device__ __forceinline__ double band_image(int r, float hr, float lr) {  // k:183-189
  if (r < hr && r >= lr) return 1.0;
  if (r >= hr && r < hr + 8) return 0.5 * cos(kRefPi * (r - (double)hr) / 16) + 0.5;
  if (r >= (lr - 8) && r < lr && r >= 0) return 0.5 * cos(kRefPi * ((double)lr - r) / 16) + 0.5;
  return 0.0;
}

// Block-wide sum of a double (blockDim.x a multiple of 64, at most 1024); valid in thread 0.
__device__ __forceinline__ double block_sum(double v) {
  __shared__ double red[16];
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const int w = threadIdx.x / 64, lane = threadIdx.x % 64;
  __syncthreads();
  if (lane == 0) red[w] = v;
  __syncthreads();
  double t = 0;
  if (threadIdx.x == 0)
    for (int i = 0; i < (int)(blockDim.x / 64); ++i) t += red[i];
  return t;
}

// Output tile of the template rotations: kRotTX x kRotTY pixels per workgroup of 256 threads, two
// pixels per thread (each wavefront a 16 x 8 part of the tile, as two 8 x 8 squares).  A compact tile keeps the rotated source footprint of a
// workgroup compact at every angle (a full-row strip fans out over many source rows near 45
// degrees).
constexpr int kRotTX = 32, kRotTY = 16;

// Bilinear sample geometry of rotate_subIMG (k:534-603), float arithmetic with explicit FMA.
struct Bil {
  int k0, k1, k2, k3;
  float t, u;
  bool ok;
};
__device__ __forceinline__ Bil bil_coords(int i, int j, int l, float cose, float sine) {
  Bil o;
  const float x = (float)(i - l / 2), y = (float)(j - l / 2);
  const float half = (float)(l / 2);
  const float x2 = __fmaf_rn(cose, x, sine * y) + half;
  const float y2 = __fmaf_rn(-sine, x, cose * y) + half;
  o.ok = !(x2 < 0 || x2 > l - 1.0f || y2 < 0 || y2 > l - 1.0f);
  if (!o.ok) return o;
  const int ii = (int)floorf(x2), jj = (int)floorf(y2);
  o.k0 = ii + jj * l;
  o.k1 = o.k0 + 1;
  o.k2 = o.k0 + l + 1;
  o.k3 = o.k0 + l;
  if (ii == l - 1) { o.k1 -= 1; o.k2 -= 1; }
  if (jj == l - 1) { o.k2 -= l; o.k3 -= l; }
  o.t = x2 - (float)ii;
  o.u = y2 - (float)jj;
  return o;
}

// Copies the N elements e = tid + 256 k of a 256-thread workgroup's tile in bursts of up to CH
// loads per thread, all issued before the first is waited for (a plain strided loop waits for
// each load in turn, which leaves the copy bound by memory latency).  src(e) returns element e,
// dst(e, v) stores it.
template <int N, int CH, typename Src, typename Dst>
__device__ __forceinline__ void burst_copy(int tid, Src&& src, Dst&& dst) {
  constexpr int NE = (N + 255) / 256;
#pragma unroll
  for (int k0 = 0; k0 < NE; k0 += CH) {
    float2 v[CH];
#pragma unroll
    for (int k = 0; k < CH; ++k) {
      const int e = tid + 256 * (k0 + k);
      if (k0 + k < NE && (N % 256 == 0 || e < N)) v[k] = src(e);
    }
#pragma unroll
    for (int k = 0; k < CH; ++k) {
      const int e = tid + 256 * (k0 + k);
      if (k0 + k < NE && (N % 256 == 0 || e < N)) dst(e, v[k]);
    }
  }
}

// N elements of a table held in registers of a 256-thread workgroup between a load and a store,
// so that the load's latency overlaps the other loads issued in between.
template <int N>
struct RegStage {
  float2 v[(N + 255) / 256];
  __device__ __forceinline__ void load(const float2* __restrict__ src, int tid) {
#pragma unroll
    for (int k = 0; k < (N + 255) / 256; ++k)
      if (N % 256 == 0 || tid + 256 * k < N) v[k] = src[tid + 256 * k];
  }
  __device__ __forceinline__ void store(float2* dst, int tid) const {
#pragma unroll
    for (int k = 0; k < (N + 255) / 256; ++k)
      if (N % 256 == 0 || tid + 256 * k < N) dst[tid + 256 * k] = v[k];
  }
};

// Soft disc of apply_mask (k:225-241) at (i, j), in float as the rotation kernel applies it.
__device__ __forceinline__ float mask_factor(int i, int j, int l, float rin, float rout, float ehw) {
  const float r = hypotf((float)(i - l / 2), (float)(j - l / 2));
  if (r > rout) return 0.f;
  if (r >= rin) return 0.5f * cosf((float)kRefPi * (r - rin) / (2.0f * ehw)) + 0.5f;
  return 1.f;
}

}  // namespace dev
}  // namespace gs
