// Runtime-planned mixed-radix FFT for the preprocessing transforms (any length).
//
// Batched rows of length n are transformed by one kernel launch per Stockham stage between
// two global buffers; radices 4, 2, 3, 5 use register butterflies and any other prime factor
// (11, 31, 53, 101 ... for frames like 4092 or 3838 px) a direct sum.  Twiddles come from a
// per-length table exp(-2 pi i e / n) computed on the host in double precision.  Columns are
// handled by tiled transposes.  These transforms run once per micrograph (whole frame,
// template preparation, block spectra) or once per angle on the template side, so generality
// matters more than peak speed here.
#pragma once
#include <hip/hip_runtime.h>

#include <map>
#include <memory>
#include <vector>

namespace gs {

struct FftPlanGpu {
  int n = 0;
  std::vector<int> radices;
  float2* tw = nullptr;  // device, n entries
  ~FftPlanGpu();
};

class FftCache {
 public:
  const FftPlanGpu& plan(int n);
  ~FftCache() = default;

 private:
  std::map<int, std::unique_ptr<FftPlanGpu>> plans_;
};

// A factorisation and twiddle table exp(-2 pi i e / n) (device, n entries) of one length,
// held in a fixed-size struct so that a transform can be set up without any allocation.
struct FftTable {
  static constexpr int kMax = 32;
  int radices[kMax];
  int nrad;
  const float2* tw;
};
// Radices of n (4s first, then 2, then odd primes) into out; returns their count.
int factor_radices(int n, int* out, int max_out);
// tw[e] = exp(-2 pi i e / n) for e < n, computed on the device.
void launch_fill_twiddles(float2* tw, int n, hipStream_t s);
// fft_rows with an explicit factorisation and twiddle table (no plan cache, nothing allocated).
void fft_rows_table(const int* radices, int nrad, const float2* tw, float2* data, float2* scratch, int n,
                    long long nrows, int sign, float scale, hipStream_t s);
// fft2_batched with explicit tables for the row (nx) and column (ny) lengths.
void fft2_table(const FftTable& rx, const FftTable& ry, float2* data, float2* scratch, int batch, int ny, int nx,
                int sign, float scale, hipStream_t s);

// In-place batched 1-D transform of nrows contiguous rows of length n, unnormalised, then
// multiplied by scale.  scratch must hold n*nrows elements.
void fft_rows(FftCache& cache, float2* data, float2* scratch, int n, long long nrows, int sign,
              float scale, hipStream_t s);

// out[b][c][r] = in[b][r][c] for r < rows, c < min(cols, out_rows): a batched transpose whose
// output may keep only its first out_rows rows (the half spectrum of a real transform).
void transpose_batched(const float2* in, float2* out, int batch, int rows, int cols, int out_rows,
                       hipStream_t s);

// In-place 2-D transform of a batch of ny x nx row-major planes; scratch holds one batch.
void fft2_batched(FftCache& cache, float2* data, float2* scratch, int batch, int ny, int nx,
                  int sign, float scale, hipStream_t s);

}  // namespace gs
