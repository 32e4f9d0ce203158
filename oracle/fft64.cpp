// FP64 mixed-radix Stockham FFT (test infrastructure only).  See fft64.hpp.
#include "fft64.hpp"

#include <cmath>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

static std::vector<int> factorize(int n) {
  std::vector<int> f;
  while (n % 4 == 0) { f.push_back(4); n /= 4; }
  while (n % 2 == 0) { f.push_back(2); n /= 2; }
  for (int p = 3; p * p <= n; p += 2)
    while (n % p == 0) { f.push_back(p); n /= p; }
  if (n > 1) f.push_back(n);
  return f;
}

const FftPlan& fft_plan(int n) {
  static std::mutex mu;
  static std::map<int, std::unique_ptr<FftPlan>> cache;
  std::lock_guard<std::mutex> lock(mu);
  auto it = cache.find(n);
  if (it != cache.end()) return *it->second;
  if (n < 1) throw std::invalid_argument("fft size must be positive");
  auto p = std::make_unique<FftPlan>();
  p->n = n;
  p->radices = factorize(n);
  p->wfwd.resize(n);
  for (int e = 0; e < n; ++e) {
    // exact argument reduction: angle = 2 pi e / n with e/n in [0,1)
    const double a = 2.0 * M_PI * (double)e / (double)n;
    p->wfwd[e] = cd(std::cos(a), -std::sin(a));
  }
  auto& ref = *p;
  cache.emplace(n, std::move(p));
  return ref;
}

// w^k for the length-m sub-transform, from the length-n table (m divides n).
static inline cd tw(const FftPlan& P, long long k, int m, int sign) {
  long long e = (k % m) * (long long)(P.n / m);
  cd w = P.wfwd[(size_t)e];
  return sign < 0 ? w : std::conj(w);
}

static void dft_small(cd* v, int R, int sign, const FftPlan& P) {
  switch (R) {
    case 2: {
      cd a = v[0], b = v[1];
      v[0] = a + b;
      v[1] = a - b;
      return;
    }
    case 4: {
      cd a0 = v[0] + v[2], a1 = v[0] - v[2];
      cd b0 = v[1] + v[3], b1 = v[1] - v[3];
      // multiply b1 by sign*i
      cd jb1 = sign < 0 ? cd(b1.imag(), -b1.real()) : cd(-b1.imag(), b1.real());
      v[0] = a0 + b0;
      v[2] = a0 - b0;
      v[1] = a1 + jb1;
      v[3] = a1 - jb1;
      return;
    }
    case 3: {
      const double c = -0.5, s = (sign < 0 ? -1.0 : 1.0) * std::sqrt(3.0) / 2.0;
      cd t1 = v[1] + v[2], t2 = v[1] - v[2];
      cd m1 = v[0] + c * t1;
      cd m2 = cd(-s * t2.imag(), s * t2.real());
      v[0] = v[0] + t1;
      v[1] = m1 + m2;
      v[2] = m1 - m2;
      return;
    }
    default: {
      cd out[64];
      std::vector<cd> big;
      cd* o = out;
      if (R > 64) { big.resize(R); o = big.data(); }
      for (int k = 0; k < R; ++k) {
        cd s = 0;
        for (int r = 0; r < R; ++r) s += v[r] * tw(P, (long long)r * k, R, sign);
        o[k] = s;
      }
      for (int k = 0; k < R; ++k) v[k] = o[k];
      return;
    }
  }
}

void fft(cd* x, cd* work, int n, int sign) {
  if (n == 1) return;
  const FftPlan& P = fft_plan(n);
  cd* src = x;
  cd* dst = work;
  int Ns = 1;
  std::vector<cd> vbuf(64);
  for (int R : P.radices) {
    if ((int)vbuf.size() < R) vbuf.resize(R);
    cd* v = vbuf.data();
    const int nR = n / R;
    for (int j = 0; j < nR; ++j) {
      const int jm = j % Ns;
      for (int r = 0; r < R; ++r) {
        cd a = src[j + r * nR];
        if (r && jm) a *= tw(P, (long long)r * jm, Ns * R, sign);
        v[r] = a;
      }
      dft_small(v, R, sign, P);
      const int idxD = (j / Ns) * Ns * R + jm;
      for (int r = 0; r < R; ++r) dst[idxD + r * Ns] = v[r];
    }
    std::swap(src, dst);
    Ns *= R;
  }
  if (src != x)
    for (int i = 0; i < n; ++i) x[i] = src[i];
}

void fft_strided(cd* x, int n, int stride, int sign) {
  std::vector<cd> a(n), w(n);
  for (int i = 0; i < n; ++i) a[i] = x[(size_t)i * stride];
  fft(a.data(), w.data(), n, sign);
  for (int i = 0; i < n; ++i) x[(size_t)i * stride] = a[i];
}

void fft2(cd* x, int ny, int nx, int sign) {
  fft_plan(nx);
  fft_plan(ny);
#pragma omp parallel
  {
    std::vector<cd> w(std::max(nx, ny)), col(ny);
#pragma omp for schedule(static)
    for (int y = 0; y < ny; ++y) fft(x + (size_t)y * nx, w.data(), nx, sign);
#pragma omp for schedule(static)
    for (int xx = 0; xx < nx; ++xx) {
      for (int y = 0; y < ny; ++y) col[y] = x[(size_t)y * nx + xx];
      fft(col.data(), w.data(), ny, sign);
      for (int y = 0; y < ny; ++y) x[(size_t)y * nx + xx] = col[y];
    }
  }
}

}  // namespace orc
