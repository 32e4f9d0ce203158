// Compile-time FFT building blocks for the l x l correlation transforms.
//
// * Dft<R, SIGN>: length-R DFT held in registers (explicit 2/3/4/5; composite radices by the
//   prime-factor algorithm where R splits into coprime factors, else by an in-register
//   Cooley-Tukey step; direct sums for other primes).  Twiddles are constexpr values, so every
//   butterfly compiles to straight-line FMA code.
// * stockham<N, T, SIGN, R...>: self-sorting Stockham stages over an LDS-resident transform
//   of length N shared by T threads (Govindaraju et al., SC'08), with an LDS slot map per stage
//   boundary.  Each stage reads all of its
//   inputs into registers, synchronises, writes its outputs in place and synchronises again, so
//   one LDS buffer per transform suffices.  Stages whose butterfly count divides T compile
//   without predication.
// SIGN = -1 is the forward transform exp(-2 pi i k n / N); +1 the unnormalised inverse.
#pragma once
#include "common.hpp"

namespace gs {
namespace ct {

// constexpr sin/cos of 2*pi*num/den for exact small-angle twiddles.
constexpr double kPi = 3.14159265358979323846264338327950288;
constexpr double cx_sin_red(double x) {  // |x| <= pi/4
  double term = x, sum = x;
  for (int i = 1; i < 14; ++i) {
    term *= -x * x / ((2.0 * i) * (2.0 * i + 1.0));
    sum += term;
  }
  return sum;
}
constexpr double cx_cos_red(double x) {
  double term = 1.0, sum = 1.0;
  for (int i = 1; i < 14; ++i) {
    term *= -x * x / ((2.0 * i - 1.0) * (2.0 * i));
    sum += term;
  }
  return sum;
}
// angle = 2 pi num/den; exact octant reduction on the rational argument.
constexpr void cx_sincos_frac(long long num, long long den, double& s, double& c) {
  num %= den;
  if (num < 0) num += den;
  // work in units of den/8 octants
  const long long n8 = num * 8;
  const long long oct = n8 / den;             // 0..7
  const double rem = (double)(n8 - oct * den) / (double)den;  // [0,1) of an octant
  double x = rem * (kPi / 4);               // [0, pi/4)
  double sx = cx_sin_red(x), cxv = cx_cos_red(x);
  // angle = oct*pi/4 + x
  double so = 0, co = 1;
  switch (oct) {
    case 0: so = sx; co = cxv; break;
    case 1: { double y = kPi / 4 - x; so = cx_cos_red(y); co = cx_sin_red(y); break; }
    case 2: so = cxv; co = -sx; break;
    case 3: { double y = kPi / 4 - x; so = cx_sin_red(y); co = -cx_cos_red(y); break; }
    case 4: so = -sx; co = -cxv; break;
    case 5: { double y = kPi / 4 - x; so = -cx_cos_red(y); co = -cx_sin_red(y); break; }
    case 6: so = -cxv; co = sx; break;
    default: { double y = kPi / 4 - x; so = -cx_sin_red(y); co = cx_cos_red(y); break; }
  }
  s = so;
  c = co;
}
constexpr double cx_cos_frac(long long num, long long den) {
  double s = 0, c = 0;
  cx_sincos_frac(num, den, s, c);
  return c;
}
constexpr double cx_sin_frac(long long num, long long den) {
  double s = 0, c = 0;
  cx_sincos_frac(num, den, s, c);
  return s;
}

// W_R^k with sign: exp(SIGN * 2 pi i k / R)
template <int R, int K, int SIGN>
struct Tw {
  static constexpr float re = (float)cx_cos_frac(K, R);
  static constexpr float im = (float)(SIGN * cx_sin_frac(K, R));
};

template <int R, int K, int SIGN>
__device__ __forceinline__ float2 twmul(float2 a) {
  constexpr int k = ((K % R) + R) % R;
  if constexpr (k == 0) {
    return a;
  } else if constexpr (4 * k == R) {  // SIGN*i
    return c_muli(a, SIGN);
  } else if constexpr (2 * k == R) {
    return make_float2(-a.x, -a.y);
  } else if constexpr (4 * k == 3 * R) {
    return c_muli(a, -SIGN);
  } else {
    constexpr float wr = Tw<R, k, SIGN>::re, wi = Tw<R, k, SIGN>::im;
    return make_float2(a.x * wr - a.y * wi, a.x * wi + a.y * wr);
  }
}

constexpr int smallest_factor(int n) {
  for (int p = 2; p * p <= n; ++p)
    if (n % p == 0) return p;
  return n;
}
// Largest power of n's smallest prime that divides n (n itself for prime powers).
constexpr int prime_power_part(int n) {
  const int p = smallest_factor(n);
  int q = 1;
  while (n % (q * p) == 0) q *= p;
  return q;
}
// Inverse of a modulo m (gcd(a, m) = 1, m > 1).
constexpr int inv_mod(int a, int m) {
  for (int x = 1; x < m; ++x)
    if ((a * x) % m == 1) return x;
  return 1;  // m == 1
}

template <int R, int SIGN>
struct Dft;

template <int SIGN>
struct Dft<1, SIGN> {
  __device__ __forceinline__ static void run(float2*) {}
};

template <int SIGN>
struct Dft<2, SIGN> {
  __device__ __forceinline__ static void run(float2* v) {
    float2 a = v[0], b = v[1];
    v[0] = c_add(a, b);
    v[1] = c_sub(a, b);
  }
};

template <int SIGN>
struct Dft<3, SIGN> {
  __device__ __forceinline__ static void run(float2* v) {
    constexpr float c = -0.5f;
    constexpr float s = (float)(SIGN * 0.86602540378443864676);
    float2 t1 = c_add(v[1], v[2]), t2 = c_sub(v[1], v[2]);
    float2 m1 = make_float2(v[0].x + c * t1.x, v[0].y + c * t1.y);
    float2 m2 = make_float2(-s * t2.y, s * t2.x);
    v[0] = c_add(v[0], t1);
    v[1] = c_add(m1, m2);
    v[2] = c_sub(m1, m2);
  }
};

template <int SIGN>
struct Dft<4, SIGN> {
  __device__ __forceinline__ static void run(float2* v) {
    float2 a0 = c_add(v[0], v[2]), a1 = c_sub(v[0], v[2]);
    float2 b0 = c_add(v[1], v[3]), b1 = c_muli(c_sub(v[1], v[3]), SIGN);
    v[0] = c_add(a0, b0);
    v[2] = c_sub(a0, b0);
    v[1] = c_add(a1, b1);
    v[3] = c_sub(a1, b1);
  }
};

template <int SIGN>
struct Dft<5, SIGN> {
  __device__ __forceinline__ static void run(float2* v) {
    constexpr float c1 = 0.30901699437494742410f, c2 = -0.80901699437494742410f;
    constexpr float s1 = (float)(SIGN * 0.95105651629515357212);
    constexpr float s2 = (float)(SIGN * 0.58778525229247312917);
    float2 a = c_add(v[1], v[4]), b = c_sub(v[1], v[4]);
    float2 c = c_add(v[2], v[3]), d = c_sub(v[2], v[3]);
    float2 x0 = v[0];
    float2 p1 = make_float2(x0.x + c1 * a.x + c2 * c.x, x0.y + c1 * a.y + c2 * c.y);
    float2 p2 = make_float2(x0.x + c2 * a.x + c1 * c.x, x0.y + c2 * a.y + c1 * c.y);
    // q1 = s1*b + s2*d ; q2 = s2*b - s1*d  (then multiplied by i)
    float2 q1 = make_float2(s1 * b.x + s2 * d.x, s1 * b.y + s2 * d.y);
    float2 q2 = make_float2(s2 * b.x - s1 * d.x, s2 * b.y - s1 * d.y);
    v[0] = make_float2(x0.x + a.x + c.x, x0.y + a.y + c.y);
    v[1] = make_float2(p1.x - q1.y, p1.y + q1.x);
    v[4] = make_float2(p1.x + q1.y, p1.y - q1.x);
    v[2] = make_float2(p2.x - q2.y, p2.y + q2.x);
    v[3] = make_float2(p2.x + q2.y, p2.y - q2.x);
  }
};

// Composite R or a prime handled by direct summation.  R = P * Q with coprime P, Q > 1 (P the
// prime-power part of the smallest prime) uses the prime-factor algorithm (Good-Thomas): input
// n = (Q n1 + P n2) mod R, output k = (Q (Q^-1 mod P) k1 + P (P^-1 mod Q) k2) mod R, so that
// X(k) = sum_n2 W_Q^(n2 k2) sum_n1 x(n) W_P^(n1 k1) with no twiddle factors between the two
// sets of DFTs (the index maps are register renames).  Prime powers R = A * B (A = smallest
// factor) take one Cooley-Tukey step with twiddles W_R^(a k2).
template <int R, int SIGN>
struct Dft {
  static constexpr int A = smallest_factor(R);
  static constexpr int B = R / A;
  static constexpr int P = prime_power_part(R);
  static constexpr int Q = R / P;

  template <int K, int N>
  __device__ __forceinline__ static void direct_acc(const float2* v, float2& acc) {
    if constexpr (N < R) {
      float2 t = twmul<R, (K * N) % R, SIGN>(v[N]);
      acc.x += t.x;
      acc.y += t.y;
      direct_acc<K, N + 1>(v, acc);
    }
  }
  template <int K>
  __device__ __forceinline__ static void direct_all(const float2* v, float2* o) {
    if constexpr (K < R) {
      float2 acc = make_float2(0.f, 0.f);
      direct_acc<K, 0>(v, acc);
      o[K] = acc;
      direct_all<K + 1>(v, o);
    }
  }
  template <int a, int k2>
  __device__ __forceinline__ static void twiddle_grid(float2 (&y)[A][B]) {
    if constexpr (a < A) {
      if constexpr (k2 < B) {
        y[a][k2] = twmul<R, a * k2, SIGN>(y[a][k2]);
        twiddle_grid<a, k2 + 1>(y);
      } else {
        twiddle_grid<a + 1, 0>(y);
      }
    }
  }

  __device__ __forceinline__ static void run(float2* v) {
    if constexpr (A == R) {  // prime
      float2 o[R];
      direct_all<0>(v, o);
#pragma unroll
      for (int k = 0; k < R; ++k) v[k] = o[k];
    } else if constexpr (Q > 1) {  // prime-factor algorithm
      constexpr int t1 = inv_mod(Q % P, P), t2 = inv_mod(P % Q, Q);
      float2 y[Q][P];
#pragma unroll
      for (int n2 = 0; n2 < Q; ++n2) {
#pragma unroll
        for (int n1 = 0; n1 < P; ++n1) y[n2][n1] = v[(Q * n1 + P * n2) % R];
        Dft<P, SIGN>::run(y[n2]);
      }
#pragma unroll
      for (int k1 = 0; k1 < P; ++k1) {
        float2 z[Q];
#pragma unroll
        for (int n2 = 0; n2 < Q; ++n2) z[n2] = y[n2][k1];
        Dft<Q, SIGN>::run(z);
#pragma unroll
        for (int k2 = 0; k2 < Q; ++k2) v[(Q * t1 * k1 + P * t2 * k2) % R] = z[k2];
      }
    } else {
      // X[k2 + B k1] = sum_a W_A^{a k1} W_R^{a k2} sum_b v[a + A b] W_B^{b k2}
      float2 y[A][B];
#pragma unroll
      for (int a = 0; a < A; ++a) {
#pragma unroll
        for (int b = 0; b < B; ++b) y[a][b] = v[a + A * b];
        Dft<B, SIGN>::run(y[a]);
      }
      twiddle_grid<0, 0>(y);
#pragma unroll
      for (int k2 = 0; k2 < B; ++k2) {
        float2 z[A];
#pragma unroll
        for (int a = 0; a < A; ++a) z[a] = y[a][k2];
        Dft<A, SIGN>::run(z);
#pragma unroll
        for (int k1 = 0; k1 < A; ++k1) v[k2 + B * k1] = z[k1];
      }
    }
  }
};

// Twiddle for stage inputs: exp(SIGN 2 pi i e / N) from an LDS table tw[e] = exp(-2 pi i e/N).
template <int SIGN>
__device__ __forceinline__ float2 tw_lookup(const float2* tw, int e) {
  float2 w = tw[e];
  return SIGN < 0 ? w : make_float2(w.x, -w.y);
}

// Ordering of LDS traffic between the read and write phases of a stage.  WAVE: all threads of
// a transform belong to one wavefront, whose LDS operations execute in order, so only compiler
// reordering must be prevented.  BLOCK: the transform spans wavefronts; use a workgroup barrier.
enum class Sync { BLOCK, WAVE };

template <Sync S>
__device__ __forceinline__ void lds_sync() {
  if constexpr (S == Sync::BLOCK) {
    __syncthreads();
  } else {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
}

// Index map of a transform's LDS slots: identity, or an XOR swizzle of the low bits (used to
// spread adjacent transforms over the LDS banks without padding; N must be a multiple of
// the mask range).
struct IdentityMap {
  __device__ __forceinline__ int operator()(int i) const { return i; }
};
struct XorMap {
  int mask;
  __device__ __forceinline__ int operator()(int i) const { return i ^ mask; }
};
// i ^ mask for mask < 16, written so that for i = lane part (< 16) + multiple of 16 the compiler
// keeps the lane part's XOR as one value and folds the multiple into the access offset.
struct LowXor16 {
  int mask;
  __device__ __forceinline__ int operator()(int i) const { return (i & ~15) | ((i & 15) ^ mask); }
};

template <int... RS>
struct RadixList {
  static constexpr int product() { return (RS * ... * 1); }
};

// Bank swizzles for the radix-4 prefixes of the Stockham plans (LDS slots of 8 bytes; 16 of
// them span the 32 four-byte banks).  A radix-4 stage writes its outputs 4 (first stage) or
// 16 (second stage) slots apart, so T consecutive lanes of one transform would share a few
// banks.  XOR-ing bits of the slot's block index into its low bits spreads those writes over
// distinct banks, while reads of T consecutive slots (one aligned block) stay a permutation of
// that block.  Both maps keep every slot inside its aligned block of T, so N % T == 0 suffices.
// Swz16: T = 16 lanes per transform, block t = bits 4-5, low bits ^= t | t << 2 (written as Swz8).
struct Swz16 {
  __device__ __forceinline__ int operator()(int i) const {
    const int t = (i >> 4) & 3;
    return (i & ~15) | ((i & 15) ^ (t | (t << 2)));
  }
};
// Swz8: T = 8 lanes per transform, block t = bits 3-4, low bits ^= t | (t & 2) << 1.  Written
// as (i & ~7) | ((i & 7) ^ g): for i = lane part (< 8) + multiple of 8 the block index and the
// XOR value are constants and the lane part's XOR is shared by all accesses with the same g.
struct Swz8 {
  __device__ __forceinline__ int operator()(int i) const {
    const int t = (i >> 3) & 3;
    return (i & ~7) | ((i & 7) ^ (t | ((t & 2) << 1)));
  }
};

// Per-stage twiddle layout of a radix plan: stage s > 0 (NS = product of the earlier radices,
// radix R) holds exp(-2 pi i r jm / (NS R)) for r in [1, R), jm in [0, NS) at offset
// OFF_s + (r - 1) NS + jm, so the lanes of a stage read consecutive entries instead of the
// stride-r entries r jm N / (NS R) of the flat table.
template <int... RS>
constexpr int stage_twiddle_count() {
  int ns = 1, n = 0;
  for (int r : {RS...}) {
    if (ns > 1) n += (r - 1) * ns;
    ns *= r;
  }
  return n;
}
template <int... RS>
constexpr int stage_twiddle_count(RadixList<RS...>) {
  return stage_twiddle_count<RS...>();
}

// Fills the per-stage layout from the flat table flat[e] = exp(-2 pi i e / N); threads tid of
// nt cooperate (the caller synchronises before use).
template <int N, int... RS>
__device__ __forceinline__ void fill_stage_twiddles(float2* dst, const float2* flat, int tid, int nt,
                                                    RadixList<RS...>) {
  constexpr int rs[] = {RS...};
  int off = 0, ns = 1;
#pragma unroll
  for (int s = 0; s < (int)sizeof...(RS); ++s) {
    const int R = rs[s];
    if (ns > 1) {
      const int cnt = (R - 1) * ns;
      for (int e = tid; e < cnt; e += nt) {
        const int r = 1 + e / ns, jm = e - (r - 1) * ns;
        dst[off + e] = flat[r * jm * (N / (ns * R))];
      }
      off += cnt;
    }
    ns *= R;
  }
}

// Twiddle sources of a Stockham transform.
//  FLAT:  tw[e] = exp(-2 pi i e / N), the flat table (global memory or an LDS copy).
//  STAGE: the per-stage layout of fill_stage_twiddles.
//  LANE:  the lane's own twiddles in registers (load_lane_twiddles): every stage with NS > 1
//         needs T % NS == 0, so that a lane meets a single jm = tid % NS in that stage.
//  RECUR: one base twiddle per lane and stage in registers (load_lane_bases), its powers formed
//         by repeated complex multiplication within the stage (w_r = w_(r-1) w_1: about r
//         roundings, under 2e-6 relative for r < 20); the same condition on T as LANE.
enum class TwSrc { FLAT, STAGE, LANE, RECUR };

template <int T, int... RS>
constexpr bool lane_twiddles_fit(RadixList<RS...>) {
  int ns = 1;
  for (int r : {RS...}) {
    if (ns > 1 && T % ns != 0) return false;
    ns *= r;
  }
  return true;
}
template <int... RS>
constexpr int lane_twiddle_count(RadixList<RS...>) {
  int ns = 1, n = 0;
  for (int r : {RS...}) {
    if (ns > 1) n += r - 1;
    ns *= r;
  }
  return n;
}

template <int... RS>
constexpr int lane_base_count(RadixList<RS...>) {
  int ns = 1, n = 0;
  for (int r : {RS...}) {
    if (ns > 1) n += 1;
    ns *= r;
  }
  return n;
}

template <int N, int T, int NS, int OFF, int R, int... REST>
__device__ __forceinline__ void load_lane_bases_impl(float2* w, const float2* flat, int tid) {
  if constexpr (NS > 1) {
    static_assert(T % NS == 0, "lane twiddles need one jm per lane and stage");
    w[OFF] = flat[(tid % NS) * (N / (NS * R))];
  }
  if constexpr (sizeof...(REST) > 0)
    load_lane_bases_impl<N, T, NS * R, OFF + (NS > 1 ? 1 : 0), REST...>(w, flat, tid);
}
// Loads exp(-2 pi i jm / (NS R)) for lane tid's jm = tid % NS of every stage with NS > 1, in
// stage order, into w[lane_base_count(plan)].
template <int N, int T, int... RS>
__device__ __forceinline__ void load_lane_bases(float2* w, const float2* flat, int tid, RadixList<RS...>) {
  load_lane_bases_impl<N, T, 1, 0, RS...>(w, flat, tid);
}

template <int N, int T, int NS, int OFF, int R, int... REST>
__device__ __forceinline__ void load_lane_twiddles_impl(float2* w, const float2* flat, int tid) {
  if constexpr (NS > 1) {
    static_assert(T % NS == 0, "lane twiddles need one jm per lane and stage");
    const int jm = tid % NS;
#pragma unroll
    for (int r = 1; r < R; ++r) w[OFF + r - 1] = flat[r * jm * (N / (NS * R))];
  }
  if constexpr (sizeof...(REST) > 0)
    load_lane_twiddles_impl<N, T, NS * R, OFF + (NS > 1 ? R - 1 : 0), REST...>(w, flat, tid);
}
// Loads the twiddles lane tid (of T per transform) multiplies by in every stage, in stage
// order, into w[lane_twiddle_count(plan)] (a register array once inlined).
template <int N, int T, int... RS>
__device__ __forceinline__ void load_lane_twiddles(float2* w, const float2* flat, int tid, RadixList<RS...>) {
  load_lane_twiddles_impl<N, T, 1, 0, RS...>(w, flat, tid);
}

// Slot maps of a transform's LDS buffer per layout: layout 0 is what the first stage reads,
// layout s (0 < s < number of stages) what stage s - 1 writes and stage s reads, and the last
// layout what the last stage writes (the layout its consumer reads).  Maps must be bijections
// of the transform's slots.  InOutMaps: one map for every layout but the last.
template <typename M, typename MF, int NST>
struct InOutMaps {
  M map;
  MF fmap;
  template <int S>
  __device__ __forceinline__ int at(int i) const {
    if constexpr (S == NST)
      return fmap(i);
    else
      return map(i);
  }
};

// One Stockham stage (radix R, NS = product of earlier radices) over a length-N transform in
// buf, shared by T threads (tid in [0,T)); then recurse into the remaining radices.  OFF is the
// offset of this stage's entries in the STAGE or LANE twiddle source; SI is the stage's index:
// it reads through maps.at<SI> and writes through maps.at<SI + 1>.
template <int N, int T, int SIGN, int NS, Sync S, typename MS, TwSrc TS, int OFF, int SI, int R, int... REST>
__device__ __forceinline__ void stockham(float2* buf, const float2* tw, int tid, const MS& maps) {
  constexpr int NB = N / R;
  constexpr int IT = (NB + T - 1) / T;
  constexpr bool even = (NB % T) == 0;
  float2 v[IT][R];
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const int j = tid + it * T;
    if (even || j < NB) {
      const int jm = j % NS;
      float2 wr = make_float2(1.f, 0.f);  // RECUR: w_1^r
#pragma unroll
      for (int r = 0; r < R; ++r) {
        float2 a = buf[maps.template at<SI>(j + r * NB)];
        if constexpr (NS > 1) {
          if (r > 0) {
            if constexpr (TS == TwSrc::RECUR) {
              const float2 w1 = tw_lookup<SIGN>(tw, OFF);
              wr = r == 1 ? w1 : c_mul(wr, w1);
              a = c_mul(a, wr);
            } else if constexpr (TS == TwSrc::LANE)
              a = c_mul(a, tw_lookup<SIGN>(tw, OFF + r - 1));
            else if constexpr (TS == TwSrc::STAGE)
              a = c_mul(a, tw_lookup<SIGN>(tw, OFF + (r - 1) * NS + jm));
            else
              a = c_mul(a, tw_lookup<SIGN>(tw, r * jm * (N / (NS * R))));
          }
        }
        v[it][r] = a;
      }
      Dft<R, SIGN>::run(v[it]);
    }
  }
  lds_sync<S>();
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const int j = tid + it * T;
    if (even || j < NB) {
      const int jm = j % NS;
      const int idxD = (j / NS) * NS * R + jm;
#pragma unroll
      for (int r = 0; r < R; ++r) buf[maps.template at<SI + 1>(idxD + r * NS)] = v[it][r];
    }
  }
  lds_sync<S>();
  constexpr int OFF2 =
      OFF + (NS > 1 ? (TS == TwSrc::LANE ? R - 1 : (TS == TwSrc::RECUR ? 1 : (R - 1) * NS)) : 0);
  if constexpr (sizeof...(REST) > 0)
    stockham<N, T, SIGN, NS * R, S, MS, TS, OFF2, SI + 1, REST...>(buf, tw, tid, maps);
}

template <int N, int T, int SIGN, Sync S, TwSrc TS, typename MS, int... RS>
__device__ __forceinline__ void run_stockham_maps(float2* buf, const float2* tw, int tid, RadixList<RS...>,
                                                  const MS& maps) {
  static_assert(RadixList<RS...>::product() == N, "radix plan must multiply to N");
  // WAVE: the caller places every transform inside one wavefront (with 64 % T != 0 it maps the
  // floor(64 / T) transforms of a wave to lanes [T j, T j + T) itself)
  static_assert(S == Sync::BLOCK || T <= 64, "wave-level sync needs transforms inside one wavefront");
  static_assert((TS != TwSrc::LANE && TS != TwSrc::RECUR) || lane_twiddles_fit<T>(RadixList<RS...>{}),
                "lane twiddles need T % NS == 0");
  stockham<N, T, SIGN, 1, S, MS, TS, 0, 0, RS...>(buf, tw, tid, maps);
}

template <int N, int T, int SIGN, Sync S = Sync::BLOCK, typename M = IdentityMap,
          TwSrc TS = TwSrc::FLAT, int... RS>
__device__ __forceinline__ void run_stockham(float2* buf, const float2* tw, int tid,
                                             RadixList<RS...> plan, M map = M{}) {
  run_stockham_maps<N, T, SIGN, S, TS>(buf, tw, tid, plan, InOutMaps<M, M, sizeof...(RS)>{map, map});
}

// As run_stockham, with the last stage writing through fmap instead of map.
template <int N, int T, int SIGN, Sync S, typename M, typename MF, TwSrc TS, int... RS>
__device__ __forceinline__ void run_stockham_fm(float2* buf, const float2* tw, int tid, RadixList<RS...> plan, M map,
                                                MF fmap) {
  run_stockham_maps<N, T, SIGN, S, TS>(buf, tw, tid, plan, InOutMaps<M, MF, sizeof...(RS)>{map, fmap});
}

}  // namespace ct
}  // namespace gs
