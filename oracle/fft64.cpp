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

static const FftPlan* build_plan(int n) {
  static std::mutex mu;
  static std::map<int, std::unique_ptr<FftPlan>> cache;
  std::lock_guard<std::mutex> lock(mu);
  auto it = cache.find(n);
  if (it != cache.end()) return it->second.get();
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
  // per stage: w^(r jm) for jm < Ns, 1 <= r < R, w = exp(-2 pi i / (Ns R)), from the length-n table
  int Ns = 1;
  for (int R : p->radices) {
    std::vector<cd> t((size_t)Ns * (R - 1));
    const long long m = (long long)Ns * R, step = n / m;
    for (int jm = 0; jm < Ns; ++jm)
      for (int r = 1; r < R; ++r) t[(size_t)jm * (R - 1) + (r - 1)] = p->wfwd[(size_t)(((long long)r * jm % m) * step)];
    p->stage_tw.push_back(std::move(t));
    Ns *= R;
  }
  const FftPlan* ref = p.get();
  cache.emplace(n, std::move(p));
  return ref;
}

const FftPlan& fft_plan(int n) {
  // a per-thread map in front of the shared cache: no lock once a thread has seen the size
  thread_local std::map<int, const FftPlan*> seen;
  auto it = seen.find(n);
  if (it != seen.end()) return *it->second;
  const FftPlan* p = build_plan(n);
  seen.emplace(n, p);
  return *p;
}

// w^k for the length-m sub-transform, from the length-n table (m divides n).
static inline cd tw(const FftPlan& P, long long k, int m, int sign) {
  long long e = (k % m) * (long long)(P.n / m);
  cd w = P.wfwd[(size_t)e];
  return sign < 0 ? w : std::conj(w);
}

static inline cd mul_i(cd a, int sign) {  // a * (sign * i)
  return sign < 0 ? cd(a.imag(), -a.real()) : cd(-a.imag(), a.real());
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
      cd jb1 = mul_i(b1, sign);
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
    case 5: {
      // y1,4 = a1 -/+ i b1, y2,3 = a2 -/+ i b2 (forward), from the pairs v1 +- v4, v2 +- v3
      static const double c1 = std::cos(2 * M_PI / 5), c2 = std::cos(4 * M_PI / 5);
      static const double s1 = std::sin(2 * M_PI / 5), s2 = std::sin(4 * M_PI / 5);
      const cd t1 = v[1] + v[4], t2 = v[2] + v[3], t3 = v[1] - v[4], t4 = v[2] - v[3];
      const cd a1 = v[0] + c1 * t1 + c2 * t2, a2 = v[0] + c2 * t1 + c1 * t2;
      const cd b1 = mul_i(s1 * t3 + s2 * t4, sign), b2 = mul_i(s2 * t3 - s1 * t4, sign);
      v[0] = v[0] + t1 + t2;
      v[1] = a1 + b1;
      v[4] = a1 - b1;
      v[2] = a2 + b2;
      v[3] = a2 - b2;
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
  cd vsmall[64];
  std::vector<cd> vbig;
  for (size_t st = 0; st < P.radices.size(); ++st) {
    const int R = P.radices[st];
    cd* v = vsmall;
    if (R > 64) {
      vbig.resize(R);
      v = vbig.data();
    }
    const cd* T = P.stage_tw[st].data();
    const int nR = n / R;
    for (int j = 0; j < nR; ++j) {
      const int jm = j % Ns;
      const cd* tj = T + (size_t)jm * (R - 1);
      v[0] = src[j];
      for (int r = 1; r < R; ++r) {
        const cd a = src[j + r * nR];
        v[r] = jm ? a * (sign < 0 ? tj[r - 1] : std::conj(tj[r - 1])) : a;
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
