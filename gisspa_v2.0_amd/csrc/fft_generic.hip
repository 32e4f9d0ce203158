// Runtime-planned mixed-radix FFT (see fft_generic.hpp).
#include <algorithm>
#include <cmath>

#include "common.hpp"
#include "fft_generic.hpp"
#include "fft_tile.hpp"

namespace gs {

FftPlanGpu::~FftPlanGpu() {
  if (tw) (void)hipFree(tw);
}

const FftPlanGpu& FftCache::plan(int n) {
  auto it = plans_.find(n);
  if (it != plans_.end()) return *it->second;
  GS_REQUIRE(n >= 1, kInvalid, "FFT length must be positive");
  auto p = std::make_unique<FftPlanGpu>();
  p->n = n;
  int rad[FftTable::kMax];
  p->radices.assign(rad, rad + factor_radices(n, rad, FftTable::kMax));
  std::vector<float2> h(n);
  for (int e = 0; e < n; ++e) {
    const double a = 2.0 * M_PI * (double)e / (double)n;
    h[e] = make_float2((float)std::cos(a), (float)-std::sin(a));
  }
  GS_HIP(hipMalloc(&p->tw, sizeof(float2) * n));
  GS_HIP(hipMemcpy(p->tw, h.data(), sizeof(float2) * n, hipMemcpyHostToDevice));
  auto& ref = *p;
  plans_.emplace(n, std::move(p));
  return ref;
}

template <int SIGN>
__device__ __forceinline__ float2 twv(const float2* tw, long long e) {
  float2 w = tw[e];
  return SIGN < 0 ? w : make_float2(w.x, -w.y);
}

template <int R, int SIGN>
__device__ __forceinline__ void stage_fixed(const float2* __restrict__ src, float2* __restrict__ dst,
                                            const float2* __restrict__ tw, int n, int Ns, int j,
                                            float scale) {
  const int nb = n / R;
  const int jm = j % Ns;
  const int step = n / (Ns * R);
  float2 v[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    float2 a = src[j + r * nb];
    if (r > 0 && jm > 0) a = c_mul(a, twv<SIGN>(tw, (long long)r * jm * step));
    v[r] = a;
  }
  ct::Dft<R, SIGN>::run(v);
  const int idxD = (j / Ns) * Ns * R + jm;
#pragma unroll
  for (int r = 0; r < R; ++r) dst[idxD + r * Ns] = c_scale(v[r], scale);
}

template <int SIGN>
__device__ __forceinline__ void stage_direct(const float2* __restrict__ src, float2* __restrict__ dst,
                                             const float2* __restrict__ tw, int n, int R, int Ns,
                                             int j, float scale) {
  const int nb = n / R;
  const int jm = j % Ns;
  const int step = n / (Ns * R);
  const int idxD = (j / Ns) * Ns * R + jm;
  const int nR = n / R;  // table stride of W_R
  for (int k = 0; k < R; ++k) {
    float2 acc = make_float2(0.f, 0.f);
    for (int r = 0; r < R; ++r) {
      float2 a = src[j + r * nb];
      if (r > 0 && jm > 0) a = c_mul(a, twv<SIGN>(tw, (long long)r * jm * step));
      const float2 w = twv<SIGN>(tw, (long long)((r * k) % R) * nR);
      acc = c_add(acc, c_mul(a, w));
    }
    dst[idxD + k * Ns] = c_scale(acc, scale);
  }
}

template <int SIGN>
__global__ void k_fft_stage(const float2* __restrict__ src, float2* __restrict__ dst,
                            const float2* __restrict__ tw, int n, int R, int Ns, long long nrows,
                            float scale) {
  const int nb = n / R;
  const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= nrows * nb) return;
  const long long row = gid / nb;
  const int j = (int)(gid - row * nb);
  const float2* s = src + row * n;
  float2* d = dst + row * n;
  switch (R) {
    case 2: stage_fixed<2, SIGN>(s, d, tw, n, Ns, j, scale); break;
    case 3: stage_fixed<3, SIGN>(s, d, tw, n, Ns, j, scale); break;
    case 4: stage_fixed<4, SIGN>(s, d, tw, n, Ns, j, scale); break;
    case 5: stage_fixed<5, SIGN>(s, d, tw, n, Ns, j, scale); break;
    case 7: stage_fixed<7, SIGN>(s, d, tw, n, Ns, j, scale); break;
    default: stage_direct<SIGN>(s, d, tw, n, R, Ns, j, scale); break;
  }
}

// A stage of any other radix R (the primes 11, 31, 53, 101 ... of frames such as 4092, 3710 or
// 3838 px) as direct sums, one thread per output: a workgroup takes NB = 256 / R butterflies
// (consecutive in the row-major (row, j) order), stages their R pre-twiddled inputs in LDS
// (each input read and twiddled once) with the table W_R^k, and thread (b, k) forms output k of
// butterfly b as the sum over r in order of A[b][r] W_R^(r k mod R).  The same arithmetic, term
// for term, as one thread per butterfly, with R times the parallelism and R times fewer reads.
template <int SIGN>
__global__ void __launch_bounds__(256)
    k_fft_stage_direct(const float2* __restrict__ src, float2* __restrict__ dst, const float2* __restrict__ tw,
                       int n, int R, int Ns, long long nbfly, float scale) {
  extern __shared__ float2 sm[];
  const int NB = 256 / R;
  float2* const A = sm;           // [NB][R]
  float2* const WR = sm + NB * R;  // W_R^k, k < R
  const int nb = n / R, step = n / (Ns * R), nR = n / R;
  const long long b0 = (long long)blockIdx.x * NB;
  const int tid = threadIdx.x;
  for (int k = tid; k < R; k += 256) WR[k] = twv<SIGN>(tw, (long long)k * nR);
  for (int idx = tid; idx < NB * R; idx += 256) {
    const int bl = idx / R, r = idx - bl * R;
    const long long g = b0 + bl;
    if (g >= nbfly) continue;
    const long long row = g / nb;
    const int j = (int)(g - row * nb), jm = j % Ns;
    float2 a = src[row * n + j + (long long)r * nb];
    if (r > 0 && jm > 0) a = c_mul(a, twv<SIGN>(tw, (long long)r * jm * step));
    A[idx] = a;
  }
  __syncthreads();
  if (tid >= NB * R) return;  // no barrier follows
  const int bl = tid / R, k = tid - bl * R;
  const long long g = b0 + bl;
  if (g >= nbfly) return;
  const long long row = g / nb;
  const int j = (int)(g - row * nb), jm = j % Ns;
  const float2* Ab = A + bl * R;
  float2 acc = make_float2(0.f, 0.f);
  int e = 0;  // r k mod R
  for (int r = 0; r < R; ++r) {
    acc = c_add(acc, c_mul(Ab[r], WR[e]));
    e += k;
    if (e >= R) e -= R;
  }
  dst[row * n + (j / Ns) * Ns * R + jm + (long long)k * Ns] = c_scale(acc, scale);
}

__global__ void k_copy_scale(const float2* __restrict__ src, float2* __restrict__ dst, long long n,
                             float scale) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = c_scale(src[i], scale);
}

static inline unsigned nblk(long long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

int factor_radices(int n, int* out, int max_out) {
  int k = 0;
  auto push = [&](int r) {
    GS_REQUIRE(k < max_out, kInvalid, "FFT length has too many factors");
    out[k++] = r;
  };
  while (n % 4 == 0) { push(4); n /= 4; }
  while (n % 2 == 0) { push(2); n /= 2; }
  for (int p = 3; p * p <= n; p += 2)
    while (n % p == 0) { push(p); n /= p; }
  if (n > 1) push(n);
  return k;
}

void fft_rows_table(const int* radices, int nrad, const float2* tw, float2* data, float2* scratch, int n,
                    long long nrows, int sign, float scale, hipStream_t s) {
  if (nrows <= 0) return;
  if (n == 1) {
    if (scale != 1.f) {
      k_copy_scale<<<nblk(nrows, 256), 256, 0, s>>>(data, data, nrows, scale);
      GS_CHECK_LAUNCH();
    }
    return;
  }
  float2* src = data;
  float2* dst = scratch;
  int Ns = 1;
  for (int st = 0; st < nrad; ++st) {
    const int R = radices[st];
    const float sc = (st == nrad - 1) ? scale : 1.f;
    const long long work = nrows * (n / R);
    if (R > 7 && R <= 256) {  // direct sums, one thread per output
      const int NB = 256 / R;
      const size_t lds = sizeof(float2) * (size_t)(NB * R + R);
      const unsigned grid = nblk(work, NB);
      if (sign < 0)
        k_fft_stage_direct<-1><<<grid, 256, lds, s>>>(src, dst, tw, n, R, Ns, work, sc);
      else
        k_fft_stage_direct<+1><<<grid, 256, lds, s>>>(src, dst, tw, n, R, Ns, work, sc);
    } else if (sign < 0) {
      k_fft_stage<-1><<<nblk(work, 256), 256, 0, s>>>(src, dst, tw, n, R, Ns, nrows, sc);
    } else {
      k_fft_stage<+1><<<nblk(work, 256), 256, 0, s>>>(src, dst, tw, n, R, Ns, nrows, sc);
    }
    GS_CHECK_LAUNCH();
    std::swap(src, dst);
    Ns *= R;
  }
  if (src != data) {
    k_copy_scale<<<nblk((long long)n * nrows, 256), 256, 0, s>>>(src, data, (long long)n * nrows, 1.f);
    GS_CHECK_LAUNCH();
  }
}

void fft_rows(FftCache& cache, float2* data, float2* scratch, int n, long long nrows, int sign,
              float scale, hipStream_t s) {
  if (nrows <= 0) return;
  const FftPlanGpu& P = cache.plan(n);
  fft_rows_table(P.radices.data(), (int)P.radices.size(), P.tw, data, scratch, n, nrows, sign, scale, s);
}

// tw[e] = exp(-2 pi i e / n), evaluated in double precision and rounded to float (the values the
// host builds for the plan cache).
__global__ void k_fill_twiddles(float2* __restrict__ tw, int n) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const double a = 2.0 * M_PI * (double)e / (double)n;
  tw[e] = make_float2((float)cos(a), (float)-sin(a));
}

void launch_fill_twiddles(float2* tw, int n, hipStream_t s) {
  k_fill_twiddles<<<nblk(n, 256), 256, 0, s>>>(tw, n);
  GS_CHECK_LAUNCH();
}

// 32x32 tiles through LDS; grid (ceil(cols/32), ceil(rows/32), batch).
__global__ void k_transpose(const float2* __restrict__ in, float2* __restrict__ out, int rows,
                            int cols, int out_rows) {
  __shared__ float2 tile[32][33];
  const long long b = blockIdx.z;
  const float2* src = in + b * (long long)rows * cols;
  float2* dst = out + b * (long long)out_rows * rows;
  const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
  const int tx = threadIdx.x, ty = threadIdx.y;  // 32 x 8
  for (int k = 0; k < 32; k += 8) {
    const int r = r0 + ty + k, c = c0 + tx;
    if (r < rows && c < cols) tile[ty + k][tx] = src[(long long)r * cols + c];
  }
  __syncthreads();
  for (int k = 0; k < 32; k += 8) {
    const int c = c0 + ty + k, r = r0 + tx;  // output row c, column r
    if (c < out_rows && c < cols && r < rows) dst[(long long)c * rows + r] = tile[tx][ty + k];
  }
}

void transpose_batched(const float2* in, float2* out, int batch, int rows, int cols, int out_rows,
                       hipStream_t s) {
  if (batch <= 0) return;
  dim3 grid((unsigned)((std::min(cols, out_rows) + 31) / 32), (unsigned)((rows + 31) / 32),
            (unsigned)batch);
  k_transpose<<<grid, dim3(32, 8), 0, s>>>(in, out, rows, cols, out_rows);
  GS_CHECK_LAUNCH();
}

void fft2_table(const FftTable& rx, const FftTable& ry, float2* data, float2* scratch, int batch, int ny, int nx,
                int sign, float scale, hipStream_t s) {
  // rows (length nx), transpose to [nx][ny], rows (length ny), transpose back
  fft_rows_table(rx.radices, rx.nrad, rx.tw, data, scratch, nx, (long long)batch * ny, sign, 1.f, s);
  transpose_batched(data, scratch, batch, ny, nx, nx, s);
  fft_rows_table(ry.radices, ry.nrad, ry.tw, scratch, data, ny, (long long)batch * nx, sign, scale, s);
  transpose_batched(scratch, data, batch, nx, ny, ny, s);
}

void fft2_batched(FftCache& cache, float2* data, float2* scratch, int batch, int ny, int nx,
                  int sign, float scale, hipStream_t s) {
  const FftPlanGpu& px = cache.plan(nx);
  const FftPlanGpu& py = cache.plan(ny);
  FftTable rx{}, ry{};
  rx.nrad = (int)px.radices.size();
  ry.nrad = (int)py.radices.size();
  GS_REQUIRE(rx.nrad <= FftTable::kMax && ry.nrad <= FftTable::kMax, kInvalid, "FFT length has too many factors");
  std::copy(px.radices.begin(), px.radices.end(), rx.radices);
  std::copy(py.radices.begin(), py.radices.end(), ry.radices);
  rx.tw = px.tw;
  ry.tw = py.tw;
  fft2_table(rx, ry, data, scratch, batch, ny, nx, sign, scale, s);
}

}  // namespace gs
