// Correlation for any even window size l: the same inverse 2-D transform of
// Ihat_b * conj(That_t) and the same reductions as the tiled kernels of corr.hip (which exist
// for l = 320 and 720), built from the runtime-planned FFT.
//
// Reference: compute_corner_CCG (k:645-678) + cufftExecC2C INVERSE (src/norm.cu:536,568;
// src/nonorm.cu:478) + add_CCG_to_sum / update_CCG / get_peak_pos (k:681-793) and
// get_peak_and_SUM (k:725-763).  The reference accepts any window_size (src/norm.cu:93-116
// plans cuFFT for it); this path keeps that.
//
// Per launch of tc templates x nb blocks (H = l/2, hx = H + 1):
//   1. P[t][b][kx][ky] = Ihat[b][kx][ky] conj(That[t0 + t][kx][ky])      (half spectra, columns)
//   2. inverse length-l transforms of every column (contiguous in ky)
//   3. per row y, the real-to-complex packing of the inverse real row transform:
//      Zc(k) = (X(k) + conj X(H - k)) + i e^{2 pi i k/l} (X(k) - conj X(H - k)), k < H,
//      through an LDS tile that turns the column layout into rows
//   4. inverse length-H transforms of the packed rows: out(2m) + i out(2m + 1)
//   5. fold: record (py, px) of the block reads output ((py + l/2) mod l, (px + l/2) mod l)
//      (the (-1)^(x+y) checkerboard as a shift; its sign for odd l/2 and 1/l^2 are in Ihat),
//      norm_type=1: per-pixel sum, sum of squares, max and first argmax over the templates in
//      order; norm_type=0: per-plane peak (strictly positive, off the border, lowest index on
//      ties), sum and sum of squares.
// The arithmetic is the tiled kernels' in another order (other FFT factorisations), so results
// agree with them to float rounding; both are checked against the FP64 oracle.
#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <utility>
#include <vector>

#include "common.hpp"
#include "fft_generic.hpp"
#include "kernels.hpp"

namespace gs {

namespace {

constexpr int BS = 256;

__global__ void k_gen_product(const float2* __restrict__ that, const float2* __restrict__ ihat,
                              float2* __restrict__ P, int t0, int tc, int nb, long long plane) {
  const long long total = (long long)tc * nb * plane;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += stride) {
    const long long q = e / plane, r = e - q * plane;
    const int t = (int)(q / nb), b = (int)(q - (long long)t * nb);
    P[e] = c_mulconj(ihat[(long long)b * plane + r], that[(long long)(t0 + t) * plane + r]);
  }
}

// Step 3.  One workgroup = one plane's tile of 32 packed outputs k x 32 rows y; it stages the
// 32 columns k0.. and their 32 mirrors H - k0 - 31 .. H - k0 in LDS (loads along y, the
// contiguous direction of the column layout) and writes rows (contiguous in k).
constexpr int PT = 32;
__global__ void __launch_bounds__(256)
    k_gen_pack(const float2* __restrict__ P, float2* __restrict__ Zr, const float2* __restrict__ twl, int l,
               int tiles_k, int tiles_y) {
  const int H = l / 2, hx = H + 1;
  __shared__ float2 xa[PT][PT + 1], xm[PT][PT + 1];
  const long long tile = blockIdx.x;
  const long long q = tile / ((long long)tiles_k * tiles_y);
  const int rem = (int)(tile - q * tiles_k * tiles_y);
  const int k0 = (rem % tiles_k) * PT, y0 = (rem / tiles_k) * PT;
  const float2* Pq = P + q * hx * (long long)l;
  const int tx = threadIdx.x % PT, ty = threadIdx.x / PT;  // 32 x 8
  for (int i = ty; i < PT; i += 8) {
    const int y = y0 + tx;
    const int k = k0 + i, km = H - k;  // km in [0, H] whenever k in [0, H]
    if (y < l && k < H) {
      xa[i][tx] = Pq[(long long)k * l + y];
      xm[i][tx] = Pq[(long long)km * l + y];
    }
  }
  __syncthreads();
  for (int j = ty; j < PT; j += 8) {
    const int y = y0 + j, k = k0 + tx;
    if (y < l && k < H) {
      const float2 a = xa[tx][j], bm = c_conj(xm[tx][j]);
      const float2 sm = c_add(a, bm), df = c_sub(a, bm);
      const float2 w = c_conj(twl[k]);  // exp(+2 pi i k / l)
      const float2 wd = c_mul(w, df);
      Zr[(q * l + y) * H + k] = make_float2(sm.x - wd.y, sm.y + wd.x);
    }
  }
}

// Step 5, norm_type=1: one thread per block pixel, templates in order.
__global__ void k_gen_fold_pixel(const float* __restrict__ out, int tc, int nb, int l, double* __restrict__ S,
                                 double* __restrict__ SS, float* __restrict__ best, int32_t* __restrict__ arg,
                                 int arg_base) {
  const long long L2 = (long long)l * l;
  const long long gi = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gi >= (long long)nb * L2) return;
  const int b = (int)(gi / L2);
  const int p = (int)(gi - b * L2), py = p / l, px = p - py * l;
  const int y = (py + l / 2) % l, x = (px + l / 2) % l;
  const float* o = out + ((long long)b * l + y) * l + x;
  const long long tstride = (long long)nb * L2;
  float s = 0.f, ss = 0.f, bs = -INFINITY;
  int ag = -1;
  for (int t = 0; t < tc; ++t) {
    const float v = o[t * tstride];
    s += v;
    ss = fmaf(v, v, ss);
    if (v > bs) {  // strict: the first maximum in template order is kept
      bs = v;
      ag = t;
    }
  }
  S[gi] += (double)s;
  SS[gi] += (double)ss;
  if (bs > best[gi]) {
    best[gi] = bs;
    arg[gi] = arg_base + ag;
  }
}

struct PlanePeak {
  float v;
  int i;
  float s, s2;
};
__device__ __forceinline__ PlanePeak peak_combine(PlanePeak a, PlanePeak b) {
  const bool take = (b.v > a.v) || (b.v == a.v && b.i < a.i);
  PlanePeak o;
  o.v = take ? b.v : a.v;
  o.i = take ? b.i : a.i;
  o.s = a.s + b.s;
  o.s2 = a.s2 + b.s2;
  return o;
}

// Step 5, norm_type=0: one workgroup per (template, block) plane, one partial each.
__global__ void __launch_bounds__(BS)
    k_gen_fold_plane(const float* __restrict__ out, int l, float border, float4* __restrict__ partial) {
  const long long q = blockIdx.x;
  const float* o = out + q * l * (long long)l;
  PlanePeak pb{0.f, 0x7fffffff, 0.f, 0.f};
  for (int e = threadIdx.x; e < l * l; e += BS) {
    const int py = e / l, px = e - py * l;
    const int y = (py + l / 2) % l, x = (px + l / 2) % l;
    const float v = o[(long long)y * l + x];
    pb.s += v;
    pb.s2 = fmaf(v, v, pb.s2);
    const bool edge = px < border || px > l - border || py < border || py > l - border;
    if (!edge && v > 0.f && (v > pb.v || (v == pb.v && e < pb.i))) {
      pb.v = v;
      pb.i = e;
    }
  }
  __shared__ PlanePeak red[BS / 64];
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  for (int off = 32; off > 0; off >>= 1) {
    PlanePeak r;
    r.v = __shfl_down(pb.v, off, 64);
    r.i = __shfl_down(pb.i, off, 64);
    r.s = __shfl_down(pb.s, off, 64);
    r.s2 = __shfl_down(pb.s2, off, 64);
    if (lane + off < 64) pb = peak_combine(pb, r);
  }
  if (lane == 0) red[wave] = pb;
  __syncthreads();
  if (threadIdx.x == 0) {
    PlanePeak r = red[0];
    for (int w = 1; w < BS / 64; ++w) r = peak_combine(r, red[w]);
    partial[q] = make_float4(r.v, __int_as_float(r.i), r.s, r.s2);
  }
}

// Reference records: one wavefront per 256-element chunk of a (template, block) plane in the
// correlation index order p = y*l + x (the reference's CCG after compute_corner_CCG: p reads
// output ((y + l/2) mod l, (x + l/2) mod l)).  Lane j holds chunk elements j, j + 64, j + 128,
// j + 192; get_peak_pos's tree (k:766-793: at stride s = 128, 64, ..., 1 element i keeps its
// value unless element i + s is strictly greater) is followed exactly: the strides 128 and 64
// inside the lane, the strides 32 .. 1 across lanes.
__global__ void __launch_bounds__(BS)
    k_gen_chunk_records(const float* __restrict__ out, int tc, int nb, int l, const double* __restrict__ mu,
                        const double* __restrict__ sigma, int t0, int angle, float thresh,
                        const int* __restrict__ offx, const int* __restrict__ offy, int bx, int Wimg, int Himg,
                        float d_m, RefRec* __restrict__ rec, int* __restrict__ count, int max_out) {
  const long long L2 = (long long)l * l;
  const int chunks = (int)(L2 / 256);
  const long long wid = (long long)blockIdx.x * (BS / 64) + threadIdx.x / 64;
  const int lane = threadIdx.x % 64;
  if (wid >= (long long)tc * nb * chunks) return;  // uniform per wavefront
  const int c = (int)(wid % chunks);
  const long long q = wid / chunks;  // plane index t * nb + b
  const int b = (int)(q % nb), t = (int)(q / nb);
  const float* o = out + q * L2;
  float v[4];
  int pos[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int p = c * 256 + lane + 64 * k;
    const int py = p / l, px = p - py * l;
    const int y = (py + l / 2) % l, x = (px + l / 2) % l;
    const double m = mu[(long long)b * L2 + p], sd = sigma[(long long)b * L2 + p];
    v[k] = sd > 0 ? (float)(((double)o[(long long)y * l + x] - m) / sd) : 0.f;  // update_CCG (k:709-721)
    pos[k] = p;
  }
  // stride 128: element i vs i + 128 (lane values k = 0 vs 2, 1 vs 3); stride 64: k = 0 vs 1
  if (v[2] > v[0]) { v[0] = v[2]; pos[0] = pos[2]; }
  if (v[3] > v[1]) { v[1] = v[3]; pos[1] = pos[3]; }
  if (v[1] > v[0]) { v[0] = v[1]; pos[0] = pos[1]; }
  float z = v[0];
  int zp = pos[0];
  for (int st = 32; st > 0; st >>= 1) {
    const float oz = __shfl_down(z, st, 64);
    const int op = __shfl_down(zp, st, 64);
    if (lane < st && oz > z) {
      z = oz;
      zp = op;
    }
  }
  if (lane == 0 && z >= thresh) {  // src/norm.cu:583-590
    const int cx = offx[b % bx] + zp % l, cy = offy[b / bx] + zp / l;
    if (cx >= d_m && cx < Wimg - d_m && cy >= d_m && cy < Himg - d_m) {
      const int slot = atomicAdd(count, 1);
      if (slot < max_out) rec[slot] = RefRec{angle, b, t0 + t, c, zp, z};
    }
  }
}

// exp(-2 pi i e / l), e < l, per (device, l), built once (host double precision).
const float2* twiddles(int l) {
  static std::mutex mu;
  static std::map<std::pair<int, int>, float2*> tabs;
  int dev = 0;
  GS_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(mu);
  auto it = tabs.find({dev, l});
  if (it != tabs.end()) return it->second;
  std::vector<float2> h(l);
  for (int e = 0; e < l; ++e) {
    const double a = 2.0 * M_PI * (double)e / (double)l;
    h[e] = make_float2((float)std::cos(a), (float)-std::sin(a));
  }
  float2* d = nullptr;
  GS_HIP(hipMalloc(&d, sizeof(float2) * l));
  GS_HIP(hipMemcpy(d, h.data(), sizeof(float2) * l, hipMemcpyHostToDevice));
  tabs.emplace(std::make_pair(dev, l), d);
  return d;
}

unsigned grid_for(long long n) { return (unsigned)std::min<long long>((n + BS - 1) / BS, 1LL << 20); }

// Steps 1-4; the real outputs end in `data`'s twin buffer, returned.
const float* corr_generic_transform(FftCache& cache, const CorrArgs& a, int l, float2* P, float2* W,
                                    hipStream_t s) {
  const int H = l / 2, hx = H + 1;
  const long long planes = (long long)a.tc * a.nb;
  const long long plane = (long long)hx * l;
  k_gen_product<<<grid_for(planes * plane), BS, 0, s>>>(a.that, a.ihat, P, a.t0, a.tc, a.nb, plane);
  GS_CHECK_LAUNCH();
  fft_rows(cache, P, W, l, planes * hx, +1, 1.f, s);
  const int tk = (H + PT - 1) / PT, ty = (l + PT - 1) / PT;
  const long long nblk = planes * tk * ty;
  GS_REQUIRE(nblk < (1LL << 31), kInvalid, "correlation launch too large");
  k_gen_pack<<<(unsigned)nblk, 256, 0, s>>>(P, W, twiddles(l), l, tk, ty);
  GS_CHECK_LAUNCH();
  fft_rows(cache, W, P, H, planes * l, +1, 1.f, s);
  return reinterpret_cast<const float*>(W);
}

}  // namespace

size_t corr_generic_buffer(int l, int tc, int nb) { return (size_t)tc * nb * (size_t)(l / 2 + 1) * l; }

void corr_generic_prepare(FftCache& cache, int l) {
  (void)cache.plan(l);
  (void)cache.plan(l / 2);
  (void)twiddles(l);
}

void launch_corr_generic_pixel(FftCache& cache, const CorrArgs& a, int l, float2* P, float2* W, double* S,
                               double* SS, float* best, int32_t* arg, int arg_base, hipStream_t s) {
  GS_REQUIRE(l > 0 && l % 2 == 0, kInvalid, "window size must be even");
  const float* out = corr_generic_transform(cache, a, l, P, W, s);
  const long long n = (long long)a.nb * l * l;
  k_gen_fold_pixel<<<(unsigned)((n + BS - 1) / BS), BS, 0, s>>>(out, a.tc, a.nb, l, S, SS, best, arg, arg_base);
  GS_CHECK_LAUNCH();
}

void launch_corr_generic_records(FftCache& cache, const CorrArgs& a, int l, float2* P, float2* W,
                                 const double* mu, const double* sigma, int angle, float thresh,
                                 const int* offx, const int* offy, int bx, int Wimg, int Himg, float d_m,
                                 RefRec* rec, int* count, int max_out, hipStream_t s) {
  GS_REQUIRE(l > 0 && l % 2 == 0 && ((long long)l * l) % 256 == 0, kInvalid,
             "reference records need window_size * window_size divisible by 256");
  const float* out = corr_generic_transform(cache, a, l, P, W, s);
  const long long waves = (long long)a.tc * a.nb * ((long long)l * l / 256);
  const long long blocks = (waves + BS / 64 - 1) / (BS / 64);
  GS_REQUIRE(blocks < (1LL << 31), kInvalid, "correlation launch too large");
  k_gen_chunk_records<<<(unsigned)blocks, BS, 0, s>>>(out, a.tc, a.nb, l, mu, sigma, a.t0, angle, thresh, offx,
                                                     offy, bx, Wimg, Himg, d_m, rec, count, max_out);
  GS_CHECK_LAUNCH();
}

void launch_corr_generic_plane(FftCache& cache, const CorrArgs& a, int l, float2* P, float2* W, float4* partial,
                               float border, hipStream_t s) {
  GS_REQUIRE(l > 0 && l % 2 == 0, kInvalid, "window size must be even");
  const float* out = corr_generic_transform(cache, a, l, P, W, s);
  const long long planes = (long long)a.tc * a.nb;
  GS_REQUIRE(planes < (1LL << 31), kInvalid, "correlation launch too large");
  k_gen_fold_plane<<<(unsigned)planes, BS, 0, s>>>(out, l, border, partial);
  GS_CHECK_LAUNCH();
}

}  // namespace gs
