// FP64 mixed-radix FFT used by the oracle (test infrastructure only).
//
// Self-sorting Stockham iteration (Govindaraju et al., "High Performance Discrete Fourier
// Transforms on Graphics Processors", SC'08): n = R_1 R_2 ... R_k; stage s reads
// x[j + r n/R], multiplies by w^(r (j mod Ns)) with w = exp(sign 2 pi i / (Ns R)), applies a
// length-R DFT and writes y[(j/Ns) Ns R + (j mod Ns) + r Ns].  Radices 4, 2, 3, 5 have
// dedicated butterflies; any other prime factor uses a direct O(R^2) DFT.  Transforms are
// unnormalised, like cufftExecC2C (reference src/fft.cu:3-21).
#pragma once
#include <complex>
#include <vector>

namespace orc {
using cd = std::complex<double>;

struct FftPlan {
  int n = 0;
  std::vector<int> radices;
  std::vector<cd> wfwd;  // exp(-2 pi i e / n), e in [0, n)
  // per Stockham stage s (radix R, Ns = product of the earlier radices): w^(r jm) at
  // [jm (R - 1) + r - 1], w = exp(-2 pi i / (Ns R)), forward sign
  std::vector<std::vector<cd>> stage_tw;
};

const FftPlan& fft_plan(int n);
// In-place unnormalised DFT of x[0..n) (sign -1 forward, +1 inverse). work has n entries.
void fft(cd* x, cd* work, int n, int sign);
// Strided convenience wrapper (allocates its own scratch).
void fft_strided(cd* x, int n, int stride, int sign);
// Row-major ny x nx 2-D transform, OpenMP over rows/columns.
void fft2(cd* x, int ny, int nx, int sign);
}  // namespace orc
