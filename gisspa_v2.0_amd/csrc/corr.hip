// Correlation hot path: the inverse 2-D transform of Ihat_b * conj(That_t) for every (template,
// block) plane, fused with the reductions the scoring needs.
//
// Reference: compute_corner_CCG (k:645-678) + cufftExecC2C INVERSE (src/norm.cu:536,568;
// src/nonorm.cu:478) + add_CCG_to_sum / update_CCG / get_peak_pos (k:681-793) and
// get_peak_and_SUM (k:725-763).
//
// Layout (half spectra of real planes, H = L/2):
//   That[t][kx 0..H][ky 0..L)   Ihat[b][kx 0..H][ky 0..L)   (column-major: a column is contiguous)
//   Z[t][b][y][kx' 0..H)        inter-pass buffer; kx'=0 packs the real DC and Nyquist columns
//                               as (re, im), rows are already shifted by L/2.
// Column pass: one workgroup = COLS adjacent packed columns of one (t, b) plane.  It loads
// Ihat*conj(That) (the DC and Nyquist columns are Hermitian along ky, so their inverse
// transforms are real and share one complex transform), runs length-L inverse FFTs in LDS and
// writes rows of Z.  The (-1)^(x+y) checkerboard of the reference is a shift by L/2 of the
// output; rows shift here, columns in the row pass, and the extra sign for odd L/2 is folded
// into the spectra by the caller.
// Row pass: one workgroup = ROWS rows of one block, looping over the templates of the launch,
// TPI at a time: each row becomes a length-H complex sequence (real-to-complex packing of the
// inverse real transform), one LDS FFT gives the L real outputs, and every thread folds the
// pixels it owns into registers -- sum, sum of squares, max and argmax for norm_type=1 (the
// per-pixel best record), or per-plane peak/sum partials for norm_type=0.
#include <cmath>
#include <mutex>
#include <type_traits>
#include <vector>

#include "common.hpp"
#include "device_common.hpp"
#include "fft_tile.hpp"
#include "kernels.hpp"

namespace gs {

template <int L>
struct CorrCfg;

template <>
struct CorrCfg<320> {
  static constexpr int H = 160;
  static constexpr int COLS = 16, CT = 256, TPC = 16;
  using ColPlan = ct::RadixList<4, 4, 4, 5>;
  static constexpr int ROWS = 8, TPI = 4, RT = 256, TPT = 8;
  using RowPlan = ct::RadixList<4, 4, 10>;
};

template <>
struct CorrCfg<720> {
  static constexpr int H = 360;
  static constexpr int COLS = 8, CT = 256, TPC = 32;
  using ColPlan = ct::RadixList<10, 6, 12>;
  static constexpr int ROWS = 4, TPI = 2, RT = 360, TPT = 45;
  using RowPlan = ct::RadixList<4, 9, 10>;
};

namespace {

// Column-pass radix plans: Plan = 0 the configuration default, 1 the alternative with one LDS
// round trip fewer where one exists (320 = 4*4*20: 16 lanes x 20 points every stage).
template <int L, int Plan>
struct ColPlanOf {
  using type = typename CorrCfg<L>::ColPlan;
};
template <>
struct ColPlanOf<320, 1> {
  using type = ct::RadixList<4, 4, 20>;
};

// Real-to-complex packing of the row transform for one Hermitian pair of a row: from X(k) = a
// and X(H - k) = b, Zc(k) = (a + conj b) + i w (a - conj b) with w = exp(+2 pi i k / L), whose
// length-H inverse transform gives the row's real outputs two at a time.  zpack2 forms both
// packed values of a pair k != H - k: with sm = a + conj b, df = a - conj b, wd = w df,
// Zc(k) = sm + i wd, and since exp(+2 pi i (H - k) / L) = -conj w, Zc(H - k) = conj sm + i conj wd.
__device__ __forceinline__ void zpack2(float2 a, float2 b, float2 w, float2& zk, float2& zm) {
  const float2 bc = c_conj(b);
  const float2 sm = c_add(a, bc), df = c_sub(a, bc);
  const float2 wd = c_mul(w, df);
  zk = make_float2(sm.x - wd.y, sm.y + wd.x);
  zm = make_float2(sm.x + wd.y, wd.x - sm.y);
}

// LDS layouts of the compact column buffer for the radix 4*4*20 plan with 16 lanes per column,
// one per stage boundary (all XOR the slot's low four bits with bits of its 16-slot block, so
// each stays a bijection of the column's 320 slots):
//  0 (loads -> stage 1): natural order.  Stage 1 reads 16 consecutive slots per lane group.
//  1 (stage 1 -> 2): i ^ ((i >> 4) & 3).  Stage 1 writes 4 t + r (+ 64 k) from lane t: bits 4-5
//    of the slot are t >> 2, so the write lands on 4 (t & 3) + (r ^ (t >> 2)), sixteen distinct
//    bank pairs; stage 2 reads t + 16 k, i.e. t ^ (k & 3) within the block.
//  2 (stage 2 -> 3): i ^ (((i >> 4) & 3) << 2).  Stage 2 writes 16 (t >> 2) + (t & 3) + 4 r:
//    (t & 3) + 4 (r ^ (t >> 2)) within the block, again sixteen distinct; stage 3 reads t + 16 k.
//  3 (stage 3 -> write-out): LowXor16 by the column's mask (columns of a row gather on distinct
//    banks).
// Without layouts 1 and 2 the radix-4 writes of stages 1 and 2 fall four lanes to a bank pair.
struct ColStageMaps {
  int mask;
  template <int S>
  __device__ __forceinline__ int at(int i) const {
    if constexpr (S == 0)
      return i;
    else if constexpr (S == 1)
      return (i & ~15) | ((i & 15) ^ ((i >> 4) & 3));
    else if constexpr (S == 2)
      return (i & ~15) | ((i & 15) ^ (((i >> 4) & 3) << 2));
    else
      return (i & ~15) | ((i & 15) ^ mask);
  }
};

// Column LDS layouts (LAY):
//  0: column stride L + 1, slots in natural order, flat twiddle table in LDS.
//  1: column stride L + 17, slots permuted by the radix-4 bank swizzle Swz16 and twiddles per
//     lane in registers (or in the per-stage layout where the plan does not allow that): with 16
//     lanes per column every Stockham read and write, every twiddle read and the row-wise gather
//     of the write-out touch 16 distinct bank pairs.
//  2: compact: columns unpadded, element y of column c at slot y ^ (c & 15) (adjacent columns on
//     different banks for the write-out gather), twiddles read from the global table: exactly
//     40 KB at L = 320, so that four workgroups fit a CU.
//  3: compact as 2, twiddles formed per lane by recurrence from one base per stage (registers;
//     no table reads in the transform), and the stage layouts of ColStageMaps.
//  4: as 3, writing the row-packed layout Zp of the inter-pass buffer: the 16 columns of group g
//     are 8 mirrored pairs (k, H - k) for k in [8g, 8g + 8) (k = 0 pairs the DC/Nyquist column
//     with column H/2), and the write-out applies the real-to-complex packing of the row
//     transform (zpack2 above) to each pair and stores the two packed values next to each other:
//     Zp[..][y][2j + 0/1] = packed column j / H - j (H/2 for j = 0).  The row pass then starts
//     from packed rows.
//  5: as 4 with every twiddle of the lane (22 for the 4*4*20 plan: they depend only on the lane's
//     place in its column) loaded from the global table into registers with the spectra, instead
//     of formed by recurrence in every stage; more registers, which the LDS-bound occupancy (four
//     workgroups of four wavefronts per CU) leaves unused.
// Z is written with streaming stores: it is read back only by the next kernel, and keeping it
// out of L2 keeps the spectra columns that neighbouring workgroups are still reading there.
// BT x TT: tile of the logical order (blocks x templates).  At 512 templates per launch, tiles
// of 32 x 16 or 32 x 32 fetch 1.8 GB per launch from beyond L2 instead of 6.4 GB but ran 4 %
// slower on MI355X; 16 x 16 and 16 x 32 ran as 8 x 16.
// CGF: the column group is the fastest index of the logical order instead of the slowest, so
// that the H / COLS workgroups of a plane run together and write its rows as whole runs of
// 8 H bytes rather than one COLS-column piece per row at a time (at the cost of reading each
// tile's spectra columns once per plane group instead of once per tile).
template <int L, int Plan = 0, int LAY = 0, int BT = 8, int TT = 16, bool CGF = false>
__global__ void __launch_bounds__(CorrCfg<L>::CT)
    k_colpass(const float2* __restrict__ that, const float2* __restrict__ ihat,
              const float2* __restrict__ twL, float2* __restrict__ Z, int t0, int tc, int nb) {
  using C = CorrCfg<L>;
  using P = typename ColPlanOf<L, Plan>::type;
  constexpr bool BK = LAY == 1, CM = LAY >= 2, RC = LAY == 3 || LAY == 4, PK = LAY == 4 || LAY == 5;
  constexpr bool LW = LAY == 5;
  constexpr int H = C::H, COLS = C::COLS, NT = C::CT, TPC = C::TPC;
  constexpr int LP = BK ? L + 17 : (CM ? L : L + 1);
  static_assert(COLS * TPC == NT, "column pass thread mapping");
  static_assert(H % COLS == 0, "column groups");
  static_assert(LAY == 0 || L % 16 == 0, "swizzle stays inside the column");
  static_assert(!BK || ct::stage_twiddle_count(P{}) <= L, "per-stage twiddles fit the table slot");
  // twiddles: each lane's own in registers where the plan allows it (BK), else a table in LDS
  // (per-stage layout for BK, flat otherwise), or the global table (CM)
  constexpr bool LT = (BK || RC || LW) && ct::lane_twiddles_fit<TPC>(P{});
  static_assert(!(RC || LW) || LT, "per-lane twiddles need T % NS == 0");
  constexpr ct::TwSrc TS = RC ? ct::TwSrc::RECUR : (LT ? ct::TwSrc::LANE : (BK ? ct::TwSrc::STAGE : ct::TwSrc::FLAT));
  constexpr long long PLANE = (long long)(H + 1) * L;
  __shared__ float2 smem[COLS * LP + (LT || CM ? 0 : L)];
  float2* const buf = smem;
  float2* const tws = smem + COLS * LP;  // LDS twiddles (unless LT or CM)
  // Slot of element y of column c: as loaded and inside the transform (slot_in), and as the last
  // stage leaves it for the write-out (slot_out).  The compact layouts swizzle only the output:
  // XOR by c & 15 (PK: 2 (c & 7), as PK gathers 8 pairs x 2 rows per 16 lanes) keeps the columns
  // a 16-lane group gathers on different banks.
  auto cmask = [](int c) { return PK ? 2 * (c & 7) : (c & 15); };
  auto slot_in = [](int c, int y) { return BK ? ct::Swz16{}(y) : y; };
  auto slot_out = [&](int c, int y) { return BK ? ct::Swz16{}(y) : (CM ? ct::LowXor16{cmask(c)}(y) : y); };
  // spectrum column held in LDS slot c (PK: the pair layout of group cg)
  auto column = [](int cg, int c) {
    if constexpr (PK) {
      const int j = 8 * cg + (c & 7);
      return c < 8 ? j : (j == 0 ? H / 2 : H - j);
    } else {
      return cg * COLS + c;
    }
  };

  // Workgroups are dispatched round-robin over the 8 XCDs; give each XCD a contiguous range of
  // logical work.  Logical order tiles the (block, template) product GEMM-style per column
  // group: a tile of BT blocks x TT templates touches BT + TT column groups of spectra (about
  // 1 MB), which its XCD's L2 keeps while the tile's BT*TT planes are produced.
  const unsigned G = gridDim.x, bid = blockIdx.x;
  const unsigned xcd = bid % 8u, idx = bid / 8u, per = G / 8u, rem = G % 8u;
  unsigned logical = (xcd < rem ? xcd * (per + 1u) : rem * (per + 1u) + (xcd - rem) * per) + idx;
  const unsigned ntb = (unsigned)(nb + BT - 1) / BT, ntt = (unsigned)(tc + TT - 1) / TT;
  int cgf = 0;
  if constexpr (CGF) {
    cgf = (int)(logical % (unsigned)(H / COLS));
    logical /= (unsigned)(H / COLS);
  }
  const int bi = (int)(logical % BT);
  logical /= BT;
  const int ti = (int)(logical % TT);
  logical /= TT;
  const int tt = (int)(logical % ntt);
  logical /= ntt;
  const int btile = (int)(logical % ntb);
  const int cg = CGF ? cgf : (int)(logical / ntb);
  const int b = btile * BT + bi;
  const int tl = tt * TT + ti;
  if (b >= nb || tl >= tc) return;  // padding of the last tiles (uniform per workgroup)
  const int tid = threadIdx.x;

  const float2* Tp = that + (long long)(t0 + tl) * PLANE;
  const float2* Ip = ihat + (long long)b * PLANE;
  const int kx0 = cg * COLS;
  // each wavefront loads and transforms its own CPW columns (contiguous in the spectra)
  constexpr int CPW = 64 / TPC;
  const int wave = tid / 64, lane = tid % 64;
  const int cw0 = wave * CPW;
  // issue every load of the wave's columns before using any, for memory-level parallelism
  constexpr int NLD = (CPW * L + 63) / 64;
  constexpr int NDC = (L + 63) / 64;  // loads covering column 0 (q < L)
  float2 tv[NLD], iv[NLD];
#pragma unroll
  for (int i = 0; i < NLD; ++i) {
    const int q = lane + 64 * i;
    if ((CPW * L) % 64 == 0 || q < CPW * L) {
      const int cl = q / L, ky = q - cl * L;
      const long long e = (long long)column(cg, cw0 + cl) * L + ky;
      tv[i] = Tp[e];
      iv[i] = Ip[e];
    }
  }
  const bool dc = (kx0 + cw0 == 0);  // this wave holds column 0: pack the Nyquist column
  float2 tn[NDC], in_[NDC];
  if (dc) {
#pragma unroll
    for (int i = 0; i < NDC; ++i) {
      const int ky = lane + 64 * i;
      if (ky < L) {
        tn[i] = Tp[(long long)H * L + ky];
        in_[i] = Ip[(long long)H * L + ky];
      }
    }
  }
  // twiddles after the spectra loads, so that their latencies overlap
  float2 wl[LT ? (RC ? ct::lane_base_count(P{}) : ct::lane_twiddle_count(P{})) : 1];
  if constexpr (RC) {
    ct::load_lane_bases<L, TPC>(wl, twL, tid % TPC, P{});
  } else if constexpr (LT) {
    ct::load_lane_twiddles<L, TPC>(wl, twL, tid % TPC, P{});
  } else if constexpr (!CM) {
    if constexpr (BK)
      ct::fill_stage_twiddles<L>(tws, twL, tid, NT, P{});
    else
      for (int i = tid; i < L; i += NT) tws[i] = twL[i];
    __syncthreads();
  }
  const float2* const tw = LT ? wl : (CM ? twL : tws);
#pragma unroll
  for (int i = 0; i < NLD; ++i) {
    const int q = lane + 64 * i;
    if ((CPW * L) % 64 == 0 || q < CPW * L) {
      const int cl = q / L, ky = q - cl * L;
      float2 p = c_mulconj(iv[i], tv[i]);
      if (i < NDC && dc && cl == 0) {  // column 0 is q < L, i.e. load i < NDC with ky = q
        const float2 pn = c_mulconj(in_[i < NDC ? i : 0], tn[i < NDC ? i : 0]);
        p = make_float2(p.x - pn.y, p.y + pn.x);
      }
      buf[(cw0 + cl) * LP + slot_in(cw0 + cl, ky)] = p;
    }
  }
  ct::lds_sync<ct::Sync::WAVE>();
  const int col = tid / TPC;
  if constexpr (BK)
    ct::run_stockham<L, TPC, +1, ct::Sync::WAVE, ct::Swz16, TS>(buf + col * LP, tw, tid % TPC, P{});
  else if constexpr (RC || LW) {
    static_assert(std::is_same_v<P, ct::RadixList<4, 4, 20>> && TPC == 16, "stage layouts of the 4*4*20 plan");
    ct::run_stockham_maps<L, TPC, +1, ct::Sync::WAVE, LW ? ct::TwSrc::LANE : ct::TwSrc::RECUR>(
        buf + col * LP, tw, tid % TPC, P{}, ColStageMaps{cmask(col)});
  }
  else if constexpr (CM)
    ct::run_stockham_fm<L, TPC, +1, ct::Sync::WAVE, ct::IdentityMap, ct::LowXor16, ct::TwSrc::FLAT>(
        buf + col * LP, tw, tid % TPC, P{}, ct::IdentityMap{}, ct::LowXor16{cmask(col)});
  else
    ct::run_stockham<L, TPC, +1, ct::Sync::WAVE>(buf + col * LP, tw, tid % TPC, P{});
  __syncthreads();  // rows of Z gather all COLS columns of the workgroup

  if constexpr (!PK) {
    // write-out: lane (y0, cc) moves rows y0, y0 + RPI, ... of column cc; every LDS read is
    // issued before the first store
    static_assert(NT % COLS == 0, "write-out mapping");
    constexpr int RPI = NT / COLS, NWO = (L + RPI - 1) / RPI;
    const int cc = tid % COLS, y0 = tid / COLS;
    float2* const Zp = Z + ((long long)tl * nb + b) * (long long)L * H + kx0 + cc;
    float2 ov[NWO];
#pragma unroll
    for (int i = 0; i < NWO; ++i)
      if (L % RPI == 0 || y0 + i * RPI < L) ov[i] = buf[cc * LP + slot_out(cc, y0 + i * RPI)];
#pragma unroll
    for (int i = 0; i < NWO; ++i) {
      const int y = y0 + i * RPI;
      if (L % RPI == 0 || y < L) st_stream(Zp + (long long)(y < L / 2 ? y + L / 2 : y - L / 2) * H, ov[i]);
    }
  } else {
    // packed write-out: lane (y0, pair jl) packs rows y0, y0 + RPP, ... of the pair (j, H - j)
    // and stores both values as one 16-byte word; 8 lanes cover the group's 128 bytes of a row
    static_assert(NT % 8 == 0 && L % (NT / 8) == 0 && COLS == 16, "packed write-out mapping");
    constexpr int RPP = NT / 8, NWP = L / RPP;
    const int jl = tid % 8, y0 = tid / 8;
    const int j = 8 * cg + jl;
    const float2 w1 = c_conj(twL[j]);  // exp(+2 pi i j / L)
    float4* const Zq = reinterpret_cast<float4*>(Z + ((long long)tl * nb + b) * (long long)L * H + 16 * cg) + jl;
    float2 av[NWP], bv[NWP];
#pragma unroll
    for (int i = 0; i < NWP; ++i) {
      av[i] = buf[jl * LP + slot_out(jl, y0 + i * RPP)];
      bv[i] = buf[(8 + jl) * LP + slot_out(8 + jl, y0 + i * RPP)];
    }
    // pair (j, H - j): both packed values from one sum, difference and product.  Group 0 (a
    // uniform branch) also holds, in lane 0, the DC/Nyquist column's real pair (X(0), X(H)) and
    // the self-paired column H/2, whose packed value (a = b, w = exp(+2 pi i (H/2) / L) = i) is
    // exactly (2 b.x, -2 b.y).
    auto pack_rows = [&](auto first) {
#pragma unroll
      for (int i = 0; i < NWP; ++i) {
        const int y = y0 + i * RPP;
        float2 z1, z2;
        zpack2(av[i], bv[i], w1, z1, z2);
        if constexpr (decltype(first)::value) {
          if (jl == 0) {
            z1 = make_float2(av[i].x + av[i].y, av[i].x - av[i].y);
            z2 = make_float2(2.f * bv[i].x, -2.f * bv[i].y);
          }
        }
        st_stream(Zq + (long long)(y < L / 2 ? y + L / 2 : y - L / 2) * (H / 2), make_float4(z1.x, z1.y, z2.x, z2.y));
      }
    };
    if (cg == 0)
      pack_rows(std::true_type{});
    else
      pack_rows(std::false_type{});
  }
}

// Column pass looping over TL templates per workgroup (l = 320, the row-packed layout of variant
// 8): workgroup (block, column group, chunk of TL templates).  Every plane's transform uses the
// same 22 lane twiddles of the 4 * 4 * 20 plan (they depend only on the lane's place in its
// column), so they are loaded once per workgroup into registers instead of being formed by
// recurrence for every plane (about a quarter of the transform's VALU instructions); the block's
// spectrum columns are read through L2 again for each template.  Arithmetic otherwise as
// k_colpass<320, 1, 4>: product with the Nyquist column packed into column 0, radix 4 * 4 * 20
// over the ColStageMaps layouts, real-to-complex pair packing in the write-out, streaming stores.
// Logical order: tiles of BT blocks x TT template chunks per column group, XCD-contiguous.
// PF: the block's columns are loaded once and held in registers, and the next template's
// columns are loaded into the registers the product has just consumed, so that they arrive
// during this template's transform and write-out.
template <int L, int TL, int BT, int TT, bool PF = false>
__global__ void __launch_bounds__(CorrCfg<L>::CT)
    k_colpass_tl(const float2* __restrict__ that, const float2* __restrict__ ihat,
                 const float2* __restrict__ twL, float2* __restrict__ Z, int t0, int tc, int nb) {
  using C = CorrCfg<L>;
  using P = ct::RadixList<4, 4, 20>;
  constexpr int H = C::H, COLS = C::COLS, NT = C::CT, TPC = C::TPC, LP = L;
  static_assert(L == 320 && COLS == 16 && TPC == 16 && COLS * TPC == NT, "the 4*4*20 plan over 16 lanes");
  constexpr int NLT = ct::lane_twiddle_count(P{});
  constexpr long long PLANE = (long long)(H + 1) * L;
  __shared__ float2 buf[COLS * LP];
  auto cmask = [](int c) { return 2 * (c & 7); };
  auto column = [](int cg, int c) {
    const int j = 8 * cg + (c & 7);
    return c < 8 ? j : (j == 0 ? H / 2 : H - j);
  };
  const unsigned G = gridDim.x, bid = blockIdx.x;
  const unsigned xcd = bid % 8u, idx = bid / 8u, per = G / 8u, rem = G % 8u;
  unsigned logical = (xcd < rem ? xcd * (per + 1u) : rem * (per + 1u) + (xcd - rem) * per) + idx;
  const int ntc = (tc + TL - 1) / TL;  // template chunks
  const unsigned ntb = (unsigned)(nb + BT - 1) / BT, ntt = (unsigned)(ntc + TT - 1) / TT;
  const int bi = (int)(logical % BT);
  logical /= BT;
  const int ti = (int)(logical % TT);
  logical /= TT;
  const int tt = (int)(logical % ntt);
  logical /= ntt;
  const int btile = (int)(logical % ntb);
  const int cg = (int)(logical / ntb);
  const int b = btile * BT + bi;
  const int tch = tt * TT + ti;
  if (b >= nb || tch >= ntc) return;  // padding of the last tiles (uniform per workgroup)
  const int tid = threadIdx.x;
  constexpr int CPW = 64 / TPC;
  const int wave = tid / 64, lane = tid % 64;
  const int cw0 = wave * CPW;
  constexpr int NLD = (CPW * L + 63) / 64;
  constexpr int NDC = (L + 63) / 64;
  static_assert((CPW * L) % 64 == 0, "whole loads");
  const float2* Ip = ihat + (long long)b * PLANE;
  const bool dc = (cg == 0 && cw0 == 0);
  float2 wl[NLT];
  ct::load_lane_twiddles<L, TPC>(wl, twL, tid % TPC, P{});
  const int col = tid / TPC;
  // packed write-out mapping (as k_colpass with LAY 4)
  constexpr int RPP = NT / 8, NWP = L / RPP;
  const int jl = tid % 8, y0 = tid / 8;
  const int j = 8 * cg + jl;
  const float2 w1 = c_conj(twL[j]);  // exp(+2 pi i j / L)
  const int tl1 = min(tc, (tch + 1) * TL);
  float2 tv[NLD], iv[NLD], tn[NDC], in_[NDC];
  auto load_t = [&](int tl) {
    const float2* Tp = that + (long long)(t0 + tl) * PLANE;
#pragma unroll
    for (int i = 0; i < NLD; ++i) {
      const int q = lane + 64 * i;
      const int cl = q / L, ky = q - cl * L;
      tv[i] = Tp[(long long)column(cg, cw0 + cl) * L + ky];
    }
    if (dc) {
#pragma unroll
      for (int i = 0; i < NDC; ++i) {
        const int ky = lane + 64 * i;
        if (ky < L) tn[i] = Tp[(long long)H * L + ky];
      }
    }
  };
  auto load_i = [&]() {
#pragma unroll
    for (int i = 0; i < NLD; ++i) {
      const int q = lane + 64 * i;
      const int cl = q / L, ky = q - cl * L;
      iv[i] = Ip[(long long)column(cg, cw0 + cl) * L + ky];
    }
    if (dc) {
#pragma unroll
      for (int i = 0; i < NDC; ++i) {
        const int ky = lane + 64 * i;
        if (ky < L) in_[i] = Ip[(long long)H * L + ky];
      }
    }
  };
  if constexpr (PF) {
    load_i();
    load_t(tch * TL);  // tch * TL < tl1: the workgroup's chunk is not empty
  }
  for (int tl = tch * TL; tl < tl1; ++tl) {
    if constexpr (!PF) {
      load_t(tl);
      load_i();
    }
#pragma unroll
    for (int i = 0; i < NLD; ++i) {
      const int q = lane + 64 * i;
      const int cl = q / L, ky = q - cl * L;
      float2 p = c_mulconj(iv[i], tv[i]);
      if (i < NDC && dc && cl == 0) {
        const float2 pn = c_mulconj(in_[i < NDC ? i : 0], tn[i < NDC ? i : 0]);
        p = make_float2(p.x - pn.y, p.y + pn.x);
      }
      buf[(cw0 + cl) * LP + ky] = p;
    }
    if constexpr (PF) {
      if (tl + 1 < tl1) load_t(tl + 1);  // in flight during the transform and write-out
    }
    ct::lds_sync<ct::Sync::WAVE>();
    ct::run_stockham_maps<L, TPC, +1, ct::Sync::WAVE, ct::TwSrc::LANE>(buf + col * LP, wl, tid % TPC, P{},
                                                                        ColStageMaps{cmask(col)});
    __syncthreads();  // rows of Z gather all COLS columns of the workgroup
    float4* const Zq = reinterpret_cast<float4*>(Z + ((long long)tl * nb + b) * (long long)L * H + 16 * cg) + jl;
    float2 av[NWP], bv[NWP];
#pragma unroll
    for (int i = 0; i < NWP; ++i) {
      av[i] = buf[jl * LP + ColStageMaps{cmask(jl)}.template at<3>(y0 + i * RPP)];
      bv[i] = buf[(8 + jl) * LP + ColStageMaps{cmask(8 + jl)}.template at<3>(y0 + i * RPP)];
    }
#pragma unroll
    for (int i = 0; i < NWP; ++i) {
      const int y = y0 + i * RPP;
      float2 z1, z2;
      zpack2(av[i], bv[i], w1, z1, z2);
      if (cg == 0 && jl == 0) {
        z1 = make_float2(av[i].x + av[i].y, av[i].x - av[i].y);
        z2 = make_float2(2.f * bv[i].x, -2.f * bv[i].y);
      }
      st_stream(Zq + (long long)(y < L / 2 ? y + L / 2 : y - L / 2) * (H / 2), make_float4(z1.x, z1.y, z2.x, z2.y));
    }
    __syncthreads();  // the next template's columns overwrite the buffer
  }
}

// Column pass, persistent over the templates of the launch: one workgroup owns one (block,
// column group) and loops over the tc templates, so the twiddle table is staged once and the
// image columns stay cache-resident.  PF: the next template's columns are loaded while the
// current one is transformed.  Each wavefront transforms its own CPW columns (wave-level sync);
// the workgroup meets once per template to write full 128-byte rows of Z.
template <int L, bool PF>
__global__ void __launch_bounds__(CorrCfg<L>::CT)
    k_colpass_pers(const float2* __restrict__ that, const float2* __restrict__ ihat,
                   const float2* __restrict__ twL, float2* __restrict__ Z, int t0, int tc, int nb) {
  using C = CorrCfg<L>;
  constexpr int H = C::H, COLS = C::COLS, NT = C::CT, TPC = C::TPC, LP = L + 1;
  constexpr int NCG = H / COLS;
  constexpr int CPW = 64 / TPC;
  constexpr int NLD = (CPW * L + 63) / 64;
  constexpr int NDC = (L + 63) / 64;
  constexpr long long PLANE = (long long)(H + 1) * L;
  static_assert(COLS * TPC == NT, "column pass thread mapping");
  __shared__ float2 buf[COLS * LP];
  __shared__ float2 tw[L];

  const int tid = threadIdx.x, wave = tid / 64, lane = tid % 64;
  const int cg = (int)(blockIdx.x % NCG);
  const int b = (int)(blockIdx.x / NCG);
  for (int i = tid; i < L; i += NT) tw[i] = twL[i];
  __syncthreads();
  const int kx0 = cg * COLS, cw0 = wave * CPW;
  const bool dc = (kx0 + cw0 == 0);
  const float2* Iw = ihat + (long long)b * PLANE + (long long)(kx0 + cw0) * L;
  const float2* In = ihat + (long long)b * PLANE + (long long)H * L;

  float2 tv[NLD], iv[NLD], tn[NDC], in_[NDC];
  auto load = [&](int tl) {
    const float2* Tw = that + (long long)(t0 + tl) * PLANE + (long long)(kx0 + cw0) * L;
#pragma unroll
    for (int i = 0; i < NLD; ++i) {
      const int q = lane + 64 * i;
      if ((CPW * L) % 64 == 0 || q < CPW * L) {
        tv[i] = Tw[q];
        iv[i] = Iw[q];
      }
    }
    if (dc) {
      const float2* Tn = that + (long long)(t0 + tl) * PLANE + (long long)H * L;
#pragma unroll
      for (int i = 0; i < NDC; ++i) {
        const int ky = lane + 64 * i;
        if (ky < L) {
          tn[i] = Tn[ky];
          in_[i] = In[ky];
        }
      }
    }
  };

  if constexpr (PF) load(0);
  for (int tl = 0; tl < tc; ++tl) {
    if constexpr (!PF) load(tl);
#pragma unroll
    for (int i = 0; i < NLD; ++i) {
      const int q = lane + 64 * i;
      if ((CPW * L) % 64 == 0 || q < CPW * L) {
        const int cl = q / L, ky = q - cl * L;
        float2 p = c_mulconj(iv[i], tv[i]);
        if (i < NDC && dc && cl == 0) {  // pack the Nyquist column as the imaginary part
          const float2 pn = c_mulconj(in_[i < NDC ? i : 0], tn[i < NDC ? i : 0]);
          p = make_float2(p.x - pn.y, p.y + pn.x);
        }
        buf[(cw0 + cl) * LP + ky] = p;
      }
    }
    if constexpr (PF) {
      if (tl + 1 < tc) load(tl + 1);
    }
    ct::lds_sync<ct::Sync::WAVE>();
    ct::run_stockham<L, TPC, +1, ct::Sync::WAVE>(buf + (tid / TPC) * LP, tw, tid % TPC,
                                                  typename C::ColPlan{});
    __syncthreads();
    float2* Zp = Z + ((long long)tl * nb + b) * (long long)L * H;
    for (int q = tid; q < COLS * L; q += NT) {
      const int y = q / COLS, cc = q - y * COLS;
      int py = y + L / 2;
      if (py >= L) py -= L;
      st_stream(Zp + (long long)py * H + kx0 + cc, buf[cc * LP + y]);
    }
    __syncthreads();
  }
}

struct PlaneBest {
  float v;
  int i;
  float s, s2;
};
__device__ __forceinline__ PlaneBest plane_combine(PlaneBest a, PlaneBest b) {
  PlaneBest o;
  const bool take = (b.v > a.v) || (b.v == a.v && b.i < a.i);
  o.v = take ? b.v : a.v;
  o.i = take ? b.i : a.i;
  o.s = a.s + b.s;
  o.s2 = a.s2 + b.s2;
  return o;
}

template <int L, bool PLANE>
__global__ void __launch_bounds__(CorrCfg<L>::RT)
    k_rowpass(const float2* __restrict__ Z, const float2* __restrict__ twL,
              const float2* __restrict__ twH, int tc, int nb, double* __restrict__ S,
              double* __restrict__ SS, float* __restrict__ best, int32_t* __restrict__ arg,
              int arg_base, float4* __restrict__ partial, float border) {
  using C = CorrCfg<L>;
  constexpr int H = C::H, ROWS = C::ROWS, TPI = C::TPI, NT = C::RT, TPT = C::TPT;
  constexpr int NTR = ROWS * TPI;
  constexpr int PPT = ROWS * L / NT;
  static_assert(PPT * NT == ROWS * L, "pixels per thread");
  static_assert(NTR * TPT == NT, "row pass thread mapping");
  constexpr int NW = (NT + 63) / 64;  // wavefronts per workgroup
  __shared__ float2 buf[NTR * H];
  __shared__ float2 twh[H];
  __shared__ float2 twc[H];

  const int rg = blockIdx.x, b = blockIdx.y, tid = threadIdx.x;
  const int groups = L / ROWS;
  for (int i = tid; i < H; i += NT) {
    twh[i] = twH[i];
    const float2 w = twL[i];
    twc[i] = make_float2(w.x, -w.y);  // exp(+2 pi i k / L)
  }
  const int y0 = rg * ROWS;

  float s[PPT], ss[PPT], bs[PPT];
  int ag[PPT];
#pragma unroll
  for (int i = 0; i < PPT; ++i) {
    s[i] = 0.f;
    ss[i] = 0.f;
    bs[i] = -INFINITY;
    ag[i] = -1;
  }
  __syncthreads();

  for (int t = 0; t < tc; t += TPI) {
    // Real inverse transform of each row as a half-length complex one:
    // Zc(k) = (X(k) + conj X(H-k)) + i e^{2 pi i k/L} (X(k) - conj X(H-k)),
    // whose inverse gives out(2m) + i out(2m+1).
    for (int e = tid; e < NTR * H; e += NT) {
      const int k = e % H, tr = e / H;
      const int ti = tr / ROWS, row = tr - ti * ROWS;
      float2 zc = make_float2(0.f, 0.f);
      if (t + ti < tc) {
        const float2* zr = Z + ((((long long)(t + ti) * nb + b) * L) + y0 + row) * H;
        float2 A, B;
        if (k == 0) {
          const float2 z0 = zr[0];
          A = make_float2(z0.x, 0.f);
          B = make_float2(z0.y, 0.f);
        } else {
          A = zr[k];
          const float2 m = zr[H - k];
          B = make_float2(m.x, -m.y);
        }
        const float2 sum = c_add(A, B), dif = c_sub(A, B);
        const float2 wd = c_mul(twc[k], dif);
        zc = make_float2(sum.x - wd.y, sum.y + wd.x);
      }
      buf[tr * H + k] = zc;
    }
    __syncthreads();
    ct::run_stockham<H, TPT, +1>(buf + (tid / TPT) * H, twh, tid % TPT, typename C::RowPlan{});

    const float* out = reinterpret_cast<const float*>(buf);
    if constexpr (!PLANE) {
#pragma unroll
      for (int i = 0; i < PPT; ++i) {
        const int q = tid + i * NT;
        const int row = q / L, x = q - row * L;
#pragma unroll
        for (int ti = 0; ti < TPI; ++ti) {
          if (t + ti < tc) {
            const float v = out[(ti * ROWS + row) * L + x];
            s[i] += v;
            ss[i] = fmaf(v, v, ss[i]);
            if (v > bs[i]) {
              bs[i] = v;
              ag[i] = t + ti;
            }
          }
        }
      }
    } else {
      for (int ti = 0; ti < TPI; ++ti) {
        PlaneBest pb{0.f, 0x7fffffff, 0.f, 0.f};
        if (t + ti < tc) {
#pragma unroll
          for (int i = 0; i < PPT; ++i) {
            const int q = tid + i * NT;
            const int row = q / L, x = q - row * L;
            const float v = out[(ti * ROWS + row) * L + x];
            pb.s += v;
            pb.s2 = fmaf(v, v, pb.s2);
            int px = x + L / 2;
            if (px >= L) px -= L;
            const int py = y0 + row;
            const bool edge = px < border || px > L - border || py < border || py > L - border;
            const int id = py * L + px;
            if (!edge && v > 0.f && (v > pb.v || (v == pb.v && id < pb.i))) {
              pb.v = v;
              pb.i = id;
            }
          }
        }
        // per-wavefront reduction; one partial per (row group, wavefront), merged by
        // k_finalize_plane (the combine rule is order-independent)
        const int wave = tid / 64, lane = tid % 64;
        const int nact = min(64, NT - wave * 64);  // lanes present in this wavefront
        for (int o = 32; o > 0; o >>= 1) {
          PlaneBest q;
          q.v = __shfl_down(pb.v, o, 64);
          q.i = __shfl_down(pb.i, o, 64);
          q.s = __shfl_down(pb.s, o, 64);
          q.s2 = __shfl_down(pb.s2, o, 64);
          if (lane + o < nact) pb = plane_combine(pb, q);
        }
        if (lane == 0 && t + ti < tc)
          partial[(((long long)(t + ti) * nb + b) * groups + rg) * NW + wave] =
              make_float4(pb.v, __int_as_float(pb.i), pb.s, pb.s2);
      }
    }
    __syncthreads();
  }

  if constexpr (!PLANE) {
#pragma unroll
    for (int i = 0; i < PPT; ++i) {
      const int q = tid + i * NT;
      const int row = q / L, x = q - row * L;
      int px = x + L / 2;
      if (px >= L) px -= L;
      const long long gi = ((long long)b * L + (y0 + row)) * L + px;
      S[gi] += (double)s[i];
      SS[gi] += (double)ss[i];
      if (bs[i] > best[gi]) {
        best[gi] = bs[i];
        arg[gi] = arg_base + ag[i];
      }
    }
  }
}

// Column pass of the norm_type=0 correlation at l = 720, one column per wavefront: workgroup
// (block, group of 8 columns) with 8 waves loops over the templates of the launch.  Each wave
// forms the product of its spectrum column with the template's (column 0 packs the Nyquist
// column as its imaginary part), transforms it in its own LDS slice (64 lanes, radix 10 * 6 * 12)
// and the workgroup writes the group's 8 columns row-shifted by l/2 into the column-grouped
// inter-pass layout Zg[t][b][group][y][8], one contiguous 46 KB region per plane.  The block's
// columns are loaded once per workgroup and the next template's columns are loaded while the
// current one is transformed and written: with one column per wave the LDS holds at most 28
// columns per CU, too few waves to hide the loads' latency otherwise.
template <int L>
struct ColGroupCfg;
template <>
struct ColGroupCfg<720> {
  static constexpr int H = 360, COLS = 8;
  using Plan = ct::RadixList<10, 6, 12>;
};

template <int L>
__global__ void __launch_bounds__(64 * ColGroupCfg<L>::COLS) __attribute__((amdgpu_waves_per_eu(4)))
    k_colpass_grouped(const float2* __restrict__ that, const float2* __restrict__ ihat,
                      const float2* __restrict__ twL, float2* __restrict__ Z, int t0, int tc, int nb) {
  using C = ColGroupCfg<L>;
  constexpr int H = C::H, COLS = C::COLS, NT = 64 * COLS;
  constexpr int NCG = H / COLS;
  constexpr long long PLANE = (long long)(H + 1) * L;
  constexpr int NLD = (L + 63) / 64;
  static_assert(H % COLS == 0 && COLS % 2 == 0, "column groups");
  __shared__ float2 buf[COLS * L + L];
  float2* const tw = buf + COLS * L;
  // XCD-aware order: each XCD takes a contiguous range of (block, group), blocks fastest, so an
  // XCD's L2 holds the group's template columns while the blocks stream past
  const unsigned G = gridDim.x, bid = blockIdx.x;
  const unsigned xcd = bid % 8u, idx = bid / 8u, per = G / 8u, rem = G % 8u;
  const unsigned lg = (xcd < rem ? xcd * (per + 1u) : rem * (per + 1u) + (xcd - rem) * per) + idx;
  const int b = (int)(lg % (unsigned)nb);
  const int cg = (int)(lg / (unsigned)nb);
  const int tid = threadIdx.x, wave = tid / 64, lane = tid % 64;
  const int kx = cg * COLS + wave;
  const bool dc = kx == 0;  // column 0 carries the Nyquist column as its imaginary part
  const float2* Ip = ihat + (long long)b * PLANE;
  // registers hold the block's column and the template's (the next template's once consumed);
  // the Nyquist columns of the one wave holding column 0 are read when used, so that no wave
  // reserves registers for them
  // Addresses come from an opaque copy of the lane index taken where they are used: hoisted out
  // of the template loop, the 64-bit address of every unrolled load stayed live across it
  // (about 120 registers in all), which left room for one workgroup per CU.
  auto opaque_lane = [&]() {
    int ln = lane;
    asm volatile("" : "+v"(ln));
    return ln & 63;
  };
  float2 iv[NLD], tv[NLD];
  auto load_t = [&](int tl) {
    const float2* Tp = that + (long long)(t0 + tl) * PLANE + (long long)kx * L + opaque_lane();
#pragma unroll
    for (int i = 0; i < NLD; ++i)
      if (L % 64 == 0 || opaque_lane() + 64 * i < L) tv[i] = Tp[64 * i];
  };
  {
    const float2* Ic = Ip + (long long)kx * L + opaque_lane();
#pragma unroll
    for (int i = 0; i < NLD; ++i)
      if (L % 64 == 0 || opaque_lane() + 64 * i < L) iv[i] = Ic[64 * i];
  }
  load_t(0);
  for (int i = tid; i < L; i += NT) tw[i] = twL[i];
  float2* const col = buf + wave * L;
  for (int tl = 0; tl < tc; ++tl) {
    if (tl > 0) __syncthreads();  // the previous template's write-out has read every column
    {
      const int ln = opaque_lane();
#pragma unroll
      for (int i = 0; i < NLD; ++i) {
        const int ky = ln + 64 * i;
        if (L % 64 == 0 || ky < L) {
          float2 p = c_mulconj(iv[i], tv[i]);
          if (dc) {
            const float2 pn = c_mulconj(Ip[(long long)H * L + ky], that[(t0 + tl) * PLANE + (long long)H * L + ky]);
            p = make_float2(p.x - pn.y, p.y + pn.x);
          }
          col[ky] = p;
        }
      }
    }
    if (tl + 1 < tc) load_t(tl + 1);  // in flight during this template's transform and write-out
    if (tl == 0)
      __syncthreads();  // the twiddle table
    else
      ct::lds_sync<ct::Sync::WAVE>();
    ct::run_stockham<L, 64, +1, ct::Sync::WAVE>(col, tw, lane, typename C::Plan{});
    __syncthreads();  // rows of the group gather all its columns
    float4* const Zg = reinterpret_cast<float4*>(Z + (((long long)tl * nb + b) * NCG + cg) * (long long)L * COLS);
    constexpr int NQ = L * COLS / 2;  // 16-byte words of the group's region
    const int wtid = 64 * (tid / 64) + opaque_lane();  // tid, opaque as above
#pragma unroll
    for (int e0 = 0; e0 < NQ; e0 += NT) {
      const int e = e0 + wtid;
      if (NQ % NT == 0 || e < NQ) {
        const int yd = e / (COLS / 2), cp = e % (COLS / 2);
        const int y = yd < L / 2 ? yd + L / 2 : yd - L / 2;  // output row yd holds transform row y
        const float2 v0 = buf[(2 * cp) * L + y], v1 = buf[(2 * cp + 1) * L + y];
        st_stream(Zg + e, make_float4(v0.x, v0.y, v1.x, v1.y));
      }
    }
  }
}

// Row pass, norm_type=0, wavefront-private (l = 720): one wavefront per (template, block, chunk of
// RCH rows) of the correlation planes, read from the column-grouped layout of k_colpass_grouped.  Each wave iteration transforms TPW rows (one length-H
// inverse transform of the packed row per TPT lanes, radix 6 * 6 * 10), with the next
// iteration's Hermitian pairs already loaded, and folds the rows' outputs into per-lane
// records: the largest value off the border of width d_m/4 and above 0 (lowest plane index among
// equal values), the sum and the sum of squares (float per row, double across rows).  One
// partial per wave, reduced over the lanes with the order-independent rule of plane_combine.
// With one transform per wave (TPT = 64) a wave needs 3 KB of LDS and few registers, so that
// many waves per SIMD hide the latency of each transform's LDS round trips (five transforms of
// 12 lanes per wave ran at 2 waves per SIMD and half the speed).
// Replaces get_peak_and_SUM (k:725-763) over the planes of src/nonorm.cu:478-509.
template <int L>
struct PlaneWaveCfg;
template <>
struct PlaneWaveCfg<720> {
  static constexpr int H = 360, TPT = 64, TPW = 1, RCH = 20, WAVES = 4;
  using Plan = ct::RadixList<6, 6, 10>;
};

template <int L>
__global__ void __launch_bounds__(64 * PlaneWaveCfg<L>::WAVES)
    k_rowpass_plane_wave(const float2* __restrict__ Z, const float2* __restrict__ twL,
                         const float2* __restrict__ twH, int tc, int nb, float border,
                         float4* __restrict__ partial) {
  using C = PlaneWaveCfg<L>;
  constexpr int H = C::H, TPT = C::TPT, TPW = C::TPW, RCH = C::RCH, WAVES = C::WAVES;
  constexpr int NCH = L / RCH;                  // chunks (partials) per plane
  constexpr int ITER = RCH / TPW;               // wave iterations per chunk
  constexpr int NP = (H / 2 + TPT - 1) / TPT;   // Hermitian pairs per lane and row (last guarded)
  constexpr int NO = (H + TPT - 1) / TPT;       // output slots per lane and row (last guarded)
  constexpr bool DUMMY = TPW * TPT < 64;        // lanes past the last transform run a spare slice
  // transforms HS slots apart (with 12 lanes per transform, 2 * (H + 4) = 728 dwords puts the
  // wave's transforms 24 banks apart)
  constexpr int HS = H + 4;
  constexpr int SLICE = (TPW + (DUMMY ? 1 : 0)) * HS;
  static_assert(L % RCH == 0 && RCH % TPW == 0 && TPW * TPT <= 64, "row mapping");
  static_assert(!DUMMY || TPW * TPT + TPT > 64, "the spare slice covers the remaining lanes");
  static_assert(L % 4 == 0, "the half-row shift is a whole number of output slots");
  __shared__ float2 smem[WAVES * SLICE + H + H / 2 + 1];
  float2* const twh = smem + WAVES * SLICE;  // exp(-2 pi i e / H)
  float2* const twc = twh + H;               // exp(+2 pi i k / L), k <= H/2
  const int tid = threadIdx.x, wave = tid / 64, lane = tid % 64;
  for (int i = tid; i < H; i += 64 * WAVES) twh[i] = twH[i];
  for (int i = tid; i <= H / 2; i += 64 * WAVES) {
    const float2 w = twL[i];
    twc[i] = make_float2(w.x, -w.y);
  }
  __syncthreads();
  const long long wid = (long long)blockIdx.x * WAVES + wave;
  if (wid >= (long long)tc * nb * NCH) return;  // whole wavefronts; no barrier follows
  const int ch = (int)(wid % NCH);
  const long long q = wid / NCH;  // plane t * nb + b
  const float2* const Zq = Z + q * (long long)L * H;
  const int y0 = ch * RCH;
  const int tr = lane / TPT, sub = lane % TPT;
  const bool act = tr < TPW;  // lanes of the spare slice load nothing and fold nothing
  float2* const dst = smem + wave * SLICE + tr * HS;

  float2 pa[NP], pb[NP], p0 = make_float2(0.f, 0.f);
#pragma unroll
  for (int m = 0; m < NP; ++m) pa[m] = pb[m] = make_float2(0.f, 0.f);
  // element k of row y in the column-grouped layout Zg[..][k / 8][y][k % 8] of k_colpass_grouped
  constexpr int GC = ColGroupCfg<L>::COLS;
  auto at = [](const float2* zq, int y, int k) { return zq[((long long)(k / GC) * L + y) * GC + k % GC]; };
  auto prefetch = [&](int it) {
    if (act) {
      const int y = y0 + it * TPW + tr;
      if (sub == 0) p0 = at(Zq, y, 0);
#pragma unroll
      for (int m = 0; m < NP; ++m) {
        const int k = 1 + sub + TPT * m;  // 1 .. H/2
        if ((H / 2) % TPT == 0 || k <= H / 2) {
          pa[m] = at(Zq, y, k);
          pb[m] = at(Zq, y, H - k);
        }
      }
    }
  };
  float vbest = 0.f;
  int ibest = 0x7fffffff;
  double ds = 0.0, ds2 = 0.0;
  prefetch(0);
  for (int it = 0; it < ITER; ++it) {
    // real-to-complex packing of the row's inverse transform (X(0) and X(H) share slot 0)
    if (sub == 0) dst[0] = make_float2(p0.x + p0.y, p0.x - p0.y);
#pragma unroll
    for (int m = 0; m < NP; ++m) {
      const int k = 1 + sub + TPT * m;
      if ((H / 2) % TPT == 0 || k <= H / 2) {
        float2 zk, zm;
        zpack2(pa[m], pb[m], twc[k], zk, zm);
        dst[k] = zk;
        if (k != H - k) dst[H - k] = zm;
      }
    }
    if (it + 1 < ITER) prefetch(it + 1);
    ct::lds_sync<ct::Sync::WAVE>();
    ct::run_stockham<H, TPT, +1, ct::Sync::WAVE>(dst, twh, sub, typename C::Plan{});
    if (act) {
      const int py = y0 + it * TPW + tr;
      const bool yedge = py < border || py > L - border;
      float rs = 0.f, rs2 = 0.f;
      // Lane element e holds the record pixels px = 2e, 2e + 1: they sit in output slot
      // e + L/4 (mod L/2), the half-row shift of the correlation read as a rotation of the slots.
      // A lane thus meets its pixels in ascending index (rows ascend across iterations), so a
      // strict > keeps the lowest index among equal values, and vbest >= 0 makes v > vbest imply
      // the reference's v > 0.
#pragma unroll
      for (int i = 0; i < NO; ++i) {
        const int e = sub + TPT * i;
        if (H % TPT == 0 || e < H) {
          const float2 v2 = dst[e < H / 2 ? e + H / 2 : e - H / 2];
#pragma unroll
          for (int h = 0; h < 2; ++h) {
            const float v = h ? v2.y : v2.x;
            const int px = 2 * e + h;
            rs += v;
            rs2 = fmaf(v, v, rs2);
            const bool ok = !yedge && !(px < border || px > L - border) && v > vbest;
            vbest = ok ? v : vbest;
            ibest = ok ? py * L + px : ibest;
          }
        }
      }
      ds += (double)rs;
      ds2 += (double)rs2;
    }
    ct::lds_sync<ct::Sync::WAVE>();  // the next packing overwrites the slice
  }
  for (int o = 32; o > 0; o >>= 1) {
    const float ov = __shfl_down(vbest, o, 64);
    const int oi = __shfl_down(ibest, o, 64);
    ds += __shfl_down(ds, o, 64);
    ds2 += __shfl_down(ds2, o, 64);
    if (lane + o < 64 && (ov > vbest || (ov == vbest && oi < ibest))) {
      vbest = ov;
      ibest = oi;
    }
  }
  if (lane == 0) partial[q * NCH + ch] = make_float4(vbest, __int_as_float(ibest), (float)ds, (float)ds2);
}

// Row pass, norm_type=1, wavefront-private rows: each wavefront owns RW rows of the block and
// its own LDS slice, so the template loop needs no workgroup barrier.  Per iteration a wave
// transforms RW rows x TPI templates (TPT lanes per length-H transform); each lane folds the
// same PPL pixels of its rows into registers every iteration.  The real-to-complex packing
// loads each Hermitian pair X(k), X(H-k) once and writes both packed values.  PF: load the next
// iteration's pairs before transforming the current one (more registers, fewer stalls).
// TPI_: templates per wave iteration (rows per wave = transforms per wave / TPI_).
// BK: bank-conflict-free LDS layout: transforms H + 8 slots apart (adjacent transforms of a
// 16-lane group fall on opposite halves of the banks), slots permuted by the radix-4 swizzle
// Swz8 and the length-H twiddles in the per-stage layout; Z is read with streaming loads (each
// value is read once, after the column pass wrote it with streaming stores).
// PKR: Z holds row-packed pairs (column pass layout 4): each lane loads 16-byte pairs and writes
// them to their two slots, with no packing arithmetic here.
// LZ: the running maximum starts from the pixel's record of the earlier launches, and an
// iteration's TPI values take the per-template strict-maximum update only where their maximum
// exceeds it (per pixel: the sums, one max chain and one compare instead of a compare and two
// selects per template).  Once a few thousand hypotheses have been folded, new maxima are rare
// and most waves skip the update: the result is the same first strict maximum in template order.
// CR: instead of the per-pixel statistics, the reference's records (second pass of
// src/norm.cu:560-600 with mu and sigma known): z per (pixel, template) as update_CCG (k:709-721),
// and for every template whose z reaches the threshold somewhere in a wavefront's row, the best
// value at or above the threshold of each part of a 256-element chunk the row holds (a row of
// 320 pixels meets two chunks; the workgroup's 4 rows are 5 whole chunks), ties to the smaller
// bit-reversed offset in the chunk (the order get_peak_pos's tree keeps, k:766-793).  Each part
// is appended to ChunkRecArgs::rec; the host merges the parts of a chunk and applies the centre
// range to the winner.  A part with no value at or above the threshold cannot win a chunk whose
// maximum reaches it, so the parts skipped lose nothing.
struct ChunkRecArgs {
  const double* mu;
  const double* sigma;
  float thresh;
  int angle, t0, max_out;
  RefRec* rec;
  int* count;
};

template <int L, bool PF, int TPI_ = CorrCfg<L>::TPI, bool BK = false, bool PKR = false, bool LZ = false,
          bool CR = false>
__global__ void __launch_bounds__(256)
    k_rowpass_pix_wave(const float2* __restrict__ Z, const float2* __restrict__ twL,
                       const float2* __restrict__ twH, int tc, int nb, double* __restrict__ S,
                       double* __restrict__ SS, float* __restrict__ best, int32_t* __restrict__ arg,
                       int arg_base, ChunkRecArgs cr = {}) {
  using C = CorrCfg<L>;
  constexpr int H = C::H, TPI = TPI_, TPT = C::TPT;
  constexpr int TPW = 64 / TPT;  // transforms per wave
  constexpr int RW = TPW / TPI;  // rows per wave
  constexpr int WAVES = 4;
  constexpr int ROWS = RW * WAVES;
  constexpr int PPL = RW * L / 64;
  constexpr int HS = BK ? H + 8 : H;  // LDS slots per transform
  static_assert(TPW % TPI == 0 && RW * L % 64 == 0 && L % 64 == 0, "wave row mapping");
  static_assert((TPT & (TPT - 1)) == 0 && H % TPT == 0, "mirror slots: power-of-two lanes per transform");
  static_assert((H / 2) % TPT == 0, "paired packing");
  static_assert(!BK || (TPT == 8 && H % 8 == 0 && ct::stage_twiddle_count(typename C::RowPlan{}) <= H),
                "bank-conflict-free layout: 8 lanes per transform");
  // wave slices, then the twiddles exp(-2 pi i k/H) (flat or per-stage) and exp(+2 pi i k/L)
  __shared__ float2 smem[WAVES * TPW * HS + 2 * H];
  float2* const twf = smem + WAVES * TPW * HS;
  float2* const twc = twf + H;
  auto map = [](int i) { return BK ? ct::Swz8{}(i) : i; };

  const int tid = threadIdx.x, wave = tid / 64, lane = tid % 64;
  if constexpr (BK) ct::fill_stage_twiddles<H>(twf, twH, tid, 256, typename C::RowPlan{});
  for (int i = tid; i < H; i += 256) {
    if constexpr (!BK) twf[i] = twH[i];
    const float2 w = twL[i];
    twc[i] = make_float2(w.x, -w.y);  // exp(+2 pi i k / L)
  }
  __syncthreads();
  const int b = blockIdx.y;
  const int y0 = blockIdx.x * ROWS + wave * RW;
  float2* wb = smem + wave * TPW * HS;
  const int ti = (lane / TPT) % TPI;  // template of this lane's transform within an iteration

  float s[PPL], ss[PPL], bs[PPL];
  int ag[PPL];
  // record index of the lane's pixel i: lane + 64 i of the wave's rows, x shifted by L/2
  auto record = [&](int i) {
    const int q = lane + 64 * i;
    const int r2 = q / L, x = q - r2 * L;
    int px = x + L / 2;
    if (px >= L) px -= L;
    return ((long long)b * L + (y0 + r2)) * L + px;
  };
  static_assert(!CR || (RW == 1 && (4 * L) % 256 == 0), "chunk records: one row per wave, whole chunks per workgroup");
  double cmu[CR ? PPL : 1], cis[CR ? PPL : 1];  // mu and 1/sigma (0 where sigma <= 0: z = 0)
  int cpix[CR ? PPL : 1], ckey[CR ? PPL : 1];
  bool cseg[CR ? PPL : 1];
  if constexpr (CR) {
    const int part0 = (wave * L) / 256;  // first chunk part of this row, in workgroup chunks
#pragma unroll
    for (int i = 0; i < PPL; ++i) {
      const long long gi = record(i);
      const double sg = cr.sigma[gi];
      cmu[i] = cr.mu[gi];
      cis[i] = sg > 0 ? 1.0 / sg : 0.0;
      cpix[i] = (int)(gi - (long long)b * L * L);  // y * L + x in the block
      const int q = wave * L + cpix[i] % L;         // index in the workgroup's 4 * L
      ckey[i] = (int)(__builtin_bitreverse32((unsigned)(q % 256)) >> 24);
      cseg[i] = q / 256 != part0;
    }
  }
#pragma unroll
  for (int i = 0; i < PPL; ++i) {
    s[i] = 0.f;
    ss[i] = 0.f;
    bs[i] = LZ ? best[record(i)] : -INFINITY;
    ag[i] = -1;
  }

  // Z values of the next iteration are loaded into registers while the current one is
  // transformed and folded (software pipelining of the only HBM stream of this kernel).
  constexpr int NP = H / 2 / TPT;
  float2 pa[PKR ? 1 : NP], pb[PKR ? 1 : NP], p0 = make_float2(0.f, 0.f);
  float4 pq[PKR ? NP : 1];
  // Lane part of the Z offsets (template ti of the iteration, row zrow of the wave's rows, pair
  // group zsub), computed once: per iteration only the uniform template base moves.
  const int zsub = lane % TPT, zrow = lane / TPT / TPI;
  const long long zlane = ((long long)ti * nb * L + zrow) * H;
  auto prefetch = [&](int t) {
    if (t + ti < tc) {
      const float2* zr = Z + (((long long)t * nb + b) * L + y0) * H + zlane;
      if constexpr (PKR) {
#pragma unroll
        for (int m = 0; m < NP; ++m) pq[m] = ld_stream(reinterpret_cast<const float4*>(zr) + zsub + TPT * m);
      } else {
        if (zsub == 0) p0 = BK ? ld_stream(zr) : zr[0];
#pragma unroll
        for (int m = 0; m < NP; ++m) {
          const int k = 1 + zsub + TPT * m;  // 1 .. H/2
          pa[m] = BK ? ld_stream(zr + k) : zr[k];
          pb[m] = BK ? ld_stream(zr + H - k) : zr[H - k];
        }
      }
    }
  };
  // LDS addresses are derived from an opaque copy of the lane index inside the loop: hoisted out
  // of it, the swizzled address of every unrolled access would stay live in a register across
  // the loop (about 70 registers), halving the number of resident wavefronts.
  auto opaque_lane = [&]() {
    int ln = lane;
    asm volatile("" : "+v"(ln));
    return ln & 63;  // restores the known range (non-negative, below 64) for the index arithmetic
  };
  // One wave iteration over templates t .. t + TPI - 1 (FULL: all of them below tc).  Every
  // address below is a lane part shared by many accesses plus a compile-time offset.
  auto iteration = [&](int t, auto full) {
    constexpr bool FULL = decltype(full)::value;
    if constexpr (!PF) prefetch(t);
    const int ln = opaque_lane();
    const int sub = ln % TPT;
    float2* const dst = wb + (ln / TPT) * HS;
    if (FULL || t + ti < tc) {
      if constexpr (PKR) {
        // pair j = sub + TPT m: columns j and H - j (H/2 for j = 0).  For sub != 0 the mirror
        // H - j is TPT (H/TPT - 1 - m) + lp with lp = TPT - sub, a lane part below TPT on a
        // compile-time block; for sub == 0 it is a constant.
        const int lp = (TPT - sub) & (TPT - 1);
#pragma unroll
        for (int m = 0; m < NP; ++m) {
          dst[map(sub + TPT * m)] = make_float2(pq[m].x, pq[m].y);
          const int base = TPT * (H / TPT - 1 - m);
          const int s1 = map(base + lp) - base;
          const int s0 = map(m == 0 ? H / 2 : H - TPT * m) - base;
          dst[base + (sub != 0 ? s1 : s0)] = make_float2(pq[m].z, pq[m].w);
        }
      } else {
        if (sub == 0) dst[0] = make_float2(p0.x + p0.y, p0.x - p0.y);  // X(0), X(H) real, slot 0
#pragma unroll
        for (int m = 0; m < NP; ++m) {
          const int k = 1 + sub + TPT * m;
          const float2 A = pa[m], Bv = pb[m];
          {
            const float2 Bc = c_conj(Bv);
            const float2 sm = c_add(A, Bc), df = c_sub(A, Bc);
            const float2 wd = c_mul(twc[k], df);
            dst[map(k)] = make_float2(sm.x - wd.y, sm.y + wd.x);
          }
          if (k != H - k) {
            const float2 Ac = c_conj(A);
            const float2 sm = c_add(Bv, Ac), df = c_sub(Bv, Ac);
            const float2 wd = c_mul(twc[H - k], df);
            dst[map(H - k)] = make_float2(sm.x - wd.y, sm.y + wd.x);
          }
        }
      }
    }
    if constexpr (PF) prefetch(t + TPI);
    ct::lds_sync<ct::Sync::WAVE>();
    if constexpr (BK)
      ct::run_stockham<H, TPT, +1, ct::Sync::WAVE, ct::Swz8, ct::TwSrc::STAGE>(dst, twf, sub, typename C::RowPlan{});
    else
      ct::run_stockham<H, TPT, +1, ct::Sync::WAVE>(dst, twf, sub, typename C::RowPlan{});
    const float* out = reinterpret_cast<const float*>(wb);
    // Fold of the iteration's outputs into the lane's pixels, branch-free (the argmax update is a
    // select; written with an `if`, every (pixel, template) update became a divergent branch
    // with its LDS read waited on alone).  Pixel i of the lane is x = ln + 64 (i % (L/64)) of
    // row i / (L/64); its real output sits in complex slot x / 2 = ln / 2 + 32 (i % (L/64)), and
    // map(a + 32 c) = map(a) + 32 c, so one lane part serves every pixel.
    const int lx = 2 * map(ln >> 1) + (ln & 1);
    if constexpr (CR) {
#pragma unroll
      for (int u = 0; u < TPI; ++u) {
        if (!(FULL || t + u < tc)) continue;
        float z[PPL];
        bool hit = false;
#pragma unroll
        for (int i = 0; i < PPL; ++i) {
          const float v = out[u * 2 * HS + 64 * i + lx];
          z[i] = cis[i] != 0.0 ? (float)(((double)v - cmu[i]) * cis[i]) : 0.f;
          hit = hit || z[i] >= cr.thresh;
        }
        if (!__any(hit)) continue;  // no part of this template reaches the threshold (uniform)
#pragma unroll
        for (int sg = 0; sg < 2; ++sg) {
          float bz = -INFINITY;
          int bk = 1 << 16, bp = 0;
#pragma unroll
          for (int i = 0; i < PPL; ++i) {
            const bool take = cseg[i] == (sg == 1) && z[i] >= cr.thresh &&
                              (z[i] > bz || (z[i] == bz && ckey[i] < bk));
            bz = take ? z[i] : bz;
            bk = take ? ckey[i] : bk;
            bp = take ? cpix[i] : bp;
          }
          for (int o = 32; o > 0; o >>= 1) {
            const float oz = __shfl_xor(bz, o, 64);
            const int ok = __shfl_xor(bk, o, 64), op = __shfl_xor(bp, o, 64);
            const bool take = oz > bz || (oz == bz && ok < bk);
            bz = take ? oz : bz;
            bk = take ? ok : bk;
            bp = take ? op : bp;
          }
          if (ln == 0 && bz >= cr.thresh) {
            const int slot = atomicAdd(cr.count, 1);
            if (slot < cr.max_out) cr.rec[slot] = RefRec{cr.angle, b, cr.t0 + t + u, bp / 256, bp, bz};
          }
        }
      }
    }
#pragma unroll
    for (int i = 0; i < (CR ? 0 : PPL); ++i) {
      const int r2 = i / (L / 64), xo = 64 * (i % (L / 64)) + lx;
      if constexpr (LZ) {
        float v[TPI];
        float m = bs[i];
#pragma unroll
        for (int u = 0; u < TPI; ++u) {
          v[u] = (FULL || t + u < tc) ? out[(r2 * TPI + u) * 2 * HS + xo] : -INFINITY;
          if (FULL || t + u < tc) {
            s[i] += v[u];
            ss[i] = fmaf(v[u], v[u], ss[i]);
          }
          m = fmaxf(m, v[u]);
        }
        if (m > bs[i]) {  // a value of this iteration beats the running maximum: locate the first
#pragma unroll
          for (int u = 0; u < TPI; ++u) {
            const bool gt = v[u] > bs[i];
            bs[i] = gt ? v[u] : bs[i];
            ag[i] = gt ? t + u : ag[i];
          }
        }
      } else {
#pragma unroll
        for (int u = 0; u < TPI; ++u) {
          if (FULL || t + u < tc) {
            const float v = out[(r2 * TPI + u) * 2 * HS + xo];
            s[i] += v;
            ss[i] = fmaf(v, v, ss[i]);
            const bool gt = v > bs[i];  // strict: the first maximum in template order is kept
            bs[i] = gt ? v : bs[i];
            ag[i] = gt ? t + u : ag[i];
          }
        }
      }
    }
    ct::lds_sync<ct::Sync::WAVE>();
  };
  if constexpr (PF) prefetch(0);
  // full iterations, then the launch's partial last one (if any) on the checked path
  const int tfull = tc - tc % TPI;
  for (int t = 0; t < tfull; t += TPI) iteration(t, std::true_type{});
  if (tfull < tc) iteration(tfull, std::false_type{});

#pragma unroll
  for (int i = 0; i < (CR ? 0 : PPL); ++i) {
    const long long gi = record(i);
    S[gi] += (double)s[i];
    SS[gi] += (double)ss[i];
    // LZ: bs started at best[gi], so it moved only if a template of this launch beat it
    if (LZ ? ag[i] >= 0 : bs[i] > best[gi]) {
      best[gi] = bs[i];
      arg[gi] = arg_base + ag[i];
    }
  }
}

__global__ void k_init_acc(double* S, double* SS, float* best, int32_t* arg, long long count) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  S[i] = 0.0;
  SS[i] = 0.0;
  best[i] = -INFINITY;
  arg[i] = -1;
}

// set_CCG_mean + update_CCG + the host threshold of PickParticles (k:697-721,
// src/norm.cu:583-597) applied to the per-pixel best record.
__global__ void k_finalize_norm(const double* __restrict__ S, const double* __restrict__ SS,
                                const float* __restrict__ best, const int32_t* __restrict__ arg,
                                int nb, int l, double total_n, float thresh,
                                const int* __restrict__ offx, const int* __restrict__ offy, int bx,
                                int W, int H, float d_m, double* __restrict__ mu,
                                double* __restrict__ sigma, Cand* __restrict__ cand,
                                int* __restrict__ count, int max_cand) {
  const long long L2 = (long long)l * l;
  const long long gi = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gi >= L2 * nb) return;
  const int b = (int)(gi / L2);
  const int k = (int)(gi - (long long)b * L2);
  const double m = S[gi] / total_n;
  const double sd = sqrt(SS[gi] / total_n - m * m);
  const double z = sd > 0 ? ((double)best[gi] - m) / sd : 0.0;
  if (mu) mu[gi] = m;
  if (sigma) sigma[gi] = sd;
  if (z >= (double)thresh) {
    const int cx = offx[b % bx] + k % l, cy = offy[b / bx] + k / l;
    if (cx >= d_m && cx < W - d_m && cy >= d_m && cy < H - d_m) {
      const int slot = atomicAdd(count, 1);
      if (slot < max_cand) cand[slot] = Cand{(float)cx, (float)cy, (float)z, arg[gi], b, k};
    }
  }
}

// Per (template, block): merge row-group partials (largest peak, lowest index among equal peaks
// -- an order-independent rule; sums in double).  One wavefront per (template, block).
__global__ void k_finalize_plane(const float4* __restrict__ partial, int tc, int nb, int groups,
                                 long long tstride, double* __restrict__ out) {
  const int i = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
  const int lane = threadIdx.x % 64;
  if (i >= tc * nb) return;  // uniform per wavefront
  float v = 0.f;
  int id = 0x7fffffff;
  double s = 0, s2 = 0;
  for (int g = lane; g < groups; g += 64) {
    const float4 p = partial[(long long)i * groups + g];
    const int pi = __float_as_int(p.y);
    if (p.x > v || (p.x == v && pi < id)) {
      v = p.x;
      id = pi;
    }
    s += p.z;
    s2 += p.w;
  }
  for (int o = 32; o > 0; o >>= 1) {
    const float ov = __shfl_down(v, o, 64);
    const int oid = __shfl_down(id, o, 64);
    s += __shfl_down(s, o, 64);
    s2 += __shfl_down(s2, o, 64);
    if (ov > v || (ov == v && oid < id)) {
      v = ov;
      id = oid;
    }
  }
  if (lane == 0) {
    double* o = out + 4 * ((long long)(i / nb) * tstride + i % nb);
    o[0] = v;
    o[1] = (id == 0x7fffffff) ? 0.0 : (double)id;
    o[2] = s;
    o[3] = s2;
  }
}

}  // namespace

// Twiddle tables exp(-2 pi i e/N) for the compile-time transforms, built once per process.
struct TwTables {
  float2* t320 = nullptr;
  float2* t160 = nullptr;
  float2* t720 = nullptr;
  float2* t360 = nullptr;
};
static TwTables& tw_tables(int device);

static float2* make_table(int n) {
  std::vector<float2> h(n);
  for (int e = 0; e < n; ++e) {
    const double a = 2.0 * M_PI * (double)e / (double)n;
    h[e] = make_float2((float)std::cos(a), (float)-std::sin(a));
  }
  float2* d = nullptr;
  GS_HIP(hipMalloc(&d, sizeof(float2) * n));
  GS_HIP(hipMemcpy(d, h.data(), sizeof(float2) * n, hipMemcpyHostToDevice));
  return d;
}

static TwTables& tw_tables(int device) {
  static TwTables tabs[64];
  static bool ready[64] = {};
  static std::mutex mu;
  std::lock_guard<std::mutex> lock(mu);
  GS_REQUIRE(device >= 0 && device < 64, kInvalid, "device index out of range");
  if (!ready[device]) {
    tabs[device].t320 = make_table(320);
    tabs[device].t160 = make_table(160);
    tabs[device].t720 = make_table(720);
    tabs[device].t360 = make_table(360);
    ready[device] = true;
  }
  return tabs[device];
}

static int current_device() {
  int d = 0;
  GS_HIP(hipGetDevice(&d));
  return d;
}

bool corr_supported(int L) { return L == 320 || L == 720; }

void corr_prepare() { (void)tw_tables(current_device()); }

// ---- batched 2-D transforms of full L x L complex planes (row-major) ----
// Two passes, each one read and one write of the planes: rows (RPB contiguous rows per
// workgroup, TPC lanes per row transform, all transforms of a row set inside one wavefront) and
// columns (256/TPC adjacent columns per workgroup, staged transposed in LDS with a padded
// stride).  The column radix plan of the correlation is reused for both.
namespace {

// Input rows (and, in the column pass, input elements) y with min(y, L - y) > band are known to
// be zero: they are not read (band >= L/2: every row is read).
__device__ __forceinline__ bool zero_row(int y, int L, int band) { return min(y, L - y) > band; }

// Rows of `in` transformed into `out` (the same buffer for an in-place transform: every row a
// workgroup reads is staged in LDS before it writes any, so the pointers carry no restrict).  A
// workgroup whose rows are all zero by `band` writes nothing: those rows of `out` are left as
// they were and must only be read through the column pass with the same band.  PAIRS writes
// each row in mirror-pair order (pair_column below) for k_plane_cols_pairs.
template <int L>
__device__ __forceinline__ int pair_column(int s) {
  // slot 0: column 0, slot 1: column L/2, slots 2k and 2k + 1: columns k and L - k (0 < k < L/2)
  const int k = s >> 1;
  return k == 0 ? ((s & 1) ? L / 2 : 0) : ((s & 1) ? L - k : k);
}

template <int L, int SIGN, bool PAIRS = false>
__global__ void __launch_bounds__(256)
    k_plane_rows(const float2* in, float2* out, long long nrows, const float2* __restrict__ twL, float scale,
                 int band) {
  using C = CorrCfg<L>;
  constexpr int T = C::TPC, RPW = 64 / T, RPB = 4 * RPW;
  static_assert(L % RPB == 0, "a workgroup's rows lie in one plane");
  __shared__ float2 buf[RPB * L];
  __shared__ float2 tw[L];
  const int tid = threadIdx.x;
  const long long r0 = (long long)blockIdx.x * RPB;
  const int nr = (int)min((long long)RPB, nrows - r0);
  const float2* src = in + r0 * L;
  float2* dst = out + r0 * L;
  const int y0 = (int)(r0 % L);
  if (zero_row(y0, L, band) && zero_row(y0 + RPB - 1, L, band) && y0 + RPB - 1 < L - band)
    return;  // uniform per workgroup: zero rows transform to zero rows, which the column pass
             // (same band) takes as zero without reading them
  dev::RegStage<L> twr;
  twr.load(twL, tid);
  dev::burst_copy<RPB * L, 24>(
      tid,
      [&](int e) { return (e < nr * L && !zero_row(y0 + e / L, L, band)) ? src[e] : make_float2(0.f, 0.f); },
      [&](int e, float2 v) { buf[e] = v; });
  twr.store(tw, tid);
  __syncthreads();
  const int wave = tid / 64, lane = tid % 64;
  ct::run_stockham<L, T, SIGN, ct::Sync::WAVE>(buf + (wave * RPW + lane / T) * L, tw, lane % T,
                                               typename C::ColPlan{});
  __syncthreads();
  for (int e = tid; e < nr * L; e += 256) {
    int src_e = e;
    if constexpr (PAIRS) {
      const int r = e / L;
      src_e = r * L + pair_column<L>(e - r * L);
    }
    const float2 v = buf[src_e];
    dst[e] = make_float2(v.x * scale, v.y * scale);
  }
}

// Columns of planes whose rows are already transformed; input rows y that are zero by `band`, or
// outside [ylo, yhi] (all rows for the plane transforms), are taken as zero without being read.
// Workgroups whose columns all lie outside [clo, chi] do nothing (those columns are left as the
// row pass wrote them).
template <int L, int SIGN>
__global__ void __launch_bounds__(256)
    k_plane_cols(float2* __restrict__ data, const float2* __restrict__ twL, float scale, int band, int ylo,
                 int yhi, int clo, int chi) {
  using C = CorrCfg<L>;
  constexpr int T = C::TPC, COLS = 256 / T, LP = L + 1;
  static_assert(L % COLS == 0, "column groups");
  __shared__ float2 buf[COLS * LP];
  __shared__ float2 tw[L];
  const int tid = threadIdx.x;
  const int c0 = blockIdx.x * COLS;
  if (c0 > chi || c0 + COLS - 1 < clo) return;  // no column the caller needs (uniform per workgroup)
  dev::RegStage<L> twr;
  twr.load(twL, tid);
  float2* P = data + (long long)blockIdx.y * L * L;
  dev::burst_copy<COLS * L, 24>(
      tid,
      [&](int e) {
        const int y = e / COLS, c = e % COLS;
        return (zero_row(y, L, band) || y < ylo || y > yhi) ? make_float2(0.f, 0.f) : P[(long long)y * L + c0 + c];
      },
      [&](int e, float2 v) { buf[(e % COLS) * LP + e / COLS] = v; });
  twr.store(tw, tid);
  __syncthreads();
  ct::run_stockham<L, T, SIGN, ct::Sync::WAVE>(buf + (tid / T) * LP, tw, tid % T, typename C::ColPlan{});
  __syncthreads();
  for (int e = tid; e < COLS * L; e += 256) {
    const int y = e / COLS, c = e % COLS;
    const float2 v = buf[c * LP + y];
    P[(long long)y * L + c0 + c] = make_float2(v.x * scale, v.y * scale);
  }
}

// Column pass of the forward transform of real-pair planes (item 2j in the real part, 2j + 1 in
// the imaginary part) fused with their split into half spectra.  The row pass wrote each row in
// mirror-pair order (pair_column), so workgroup (group of HC column pairs, pair plane p = j *
// stride + g) reads one contiguous run of 2 HC slots per row: columns k and L - k for k in
// [kx0, kx0 + HC) (group 0: columns 0 and L/2, each its own mirror, instead of k = 0 and L).  It
// transforms them, forms A = (Z(k) + conj Z(-k)) / 2 and B = (Z(k) - conj Z(-k)) / 2i per
// (kx, ky) and writes out[o][kx][ky] (column-major, ky contiguous) for items i = 2j, 2j + 1 < n
// at o = i * stride + g, times scale; with stats the template standardisation (mean at the DC
// term times L^2, division by the standard deviation) as in k_unpack_pair_spectra.  The
// transformed columns never leave LDS.
template <int L>
__global__ void __launch_bounds__(256)
    k_plane_cols_pairs(const float2* __restrict__ data, const float2* __restrict__ twL, int n, int stride,
                       const float* __restrict__ stats, float scale, float2* __restrict__ out) {
  using C = CorrCfg<L>;
  constexpr int T = C::TPC, COLS = 256 / T, HC = COLS / 2, LP = L + 1, HX = L / 2 + 1;
  static_assert((L / 2) % HC == 0, "pair groups");
  __shared__ float2 buf[COLS * LP];
  __shared__ float2 tw[L];
  const int tid = threadIdx.x;
  const int kx0 = blockIdx.x * HC;
  const int p = blockIdx.y, jp = p / stride, g = p - jp * stride;
  dev::RegStage<L> twr;
  twr.load(twL, tid);
  const float2* P = data + (long long)p * L * L + 2 * kx0;
  // slot 2 jc + m of the run -> LDS column jc (m = 0) or HC + jc (m = 1)
  dev::burst_copy<COLS * L, 24>(
      tid, [&](int e) { return P[(long long)(e / COLS) * L + e % COLS]; },
      [&](int e, float2 v) {
        const int c = e % COLS;
        buf[((c & 1) * HC + (c >> 1)) * LP + e / COLS] = v;
      });
  twr.store(tw, tid);
  __syncthreads();
  ct::run_stockham<L, T, -1, ct::Sync::WAVE>(buf + (tid / T) * LP, tw, tid % T, typename C::ColPlan{});
  __syncthreads();
  float mean[2] = {0.f, 0.f}, sd[2] = {1.f, 1.f};
  if (stats) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int o = min(2 * jp + h, n - 1) * stride + g;
      mean[h] = stats[2 * o];
      const float var = stats[2 * o + 1];
      sd[h] = var != 0 ? sqrtf(var) : 1.f;
    }
  }
  // output column kx from LDS column cu and its mirror (row -ky) from LDS column cw
  auto emit = [&](int kx, int cu, int cw, int ky) {
    const float2 u = buf[cu * LP + ky], w = buf[cw * LP + (ky ? L - ky : 0)];
    const float2 v[2] = {make_float2(0.5f * (u.x + w.x), 0.5f * (u.y - w.y)),
                         make_float2(0.5f * (u.y + w.y), 0.5f * (w.x - u.x))};
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int i = 2 * jp + h;
      if (i >= n) break;
      float2 x = v[h];
      if (kx == 0 && ky == 0 && mean[h] != 0) x.x -= mean[h] * (float)((long long)L * L);
      x.x = x.x / sd[h] * scale;
      x.y = x.y / sd[h] * scale;
      out[((long long)(i * stride + g) * HX + kx) * L + ky] = x;
    }
  };
  for (int e = tid; e < HC * L; e += 256) {
    const int jc = e / L, ky = e - jc * L, kx = kx0 + jc;
    emit(kx, jc, kx == 0 ? 0 : HC + jc, ky);
  }
  if (kx0 == 0)  // column L/2, held in LDS column HC of group 0
    for (int ky = tid; ky < L; ky += 256) emit(L / 2, HC, HC, ky);
}

template <int L>
void plane_fft2_impl(const float2* in, float2* data, int batch, int sign, float scale, const float2* tw,
                     int band, int clo, int chi, hipStream_t s) {
  using C = CorrCfg<L>;
  constexpr int RPB = 4 * (64 / C::TPC), COLS = 256 / C::TPC;
  const long long nrows = (long long)batch * L;
  const long long gr = (nrows + RPB - 1) / RPB;
  GS_REQUIRE(batch > 0 && gr < (1LL << 31) && batch < 65536, kInvalid, "plane transform batch out of range");
  const dim3 gc(L / COLS, (unsigned)batch);
  if (sign < 0) {
    k_plane_rows<L, -1><<<(unsigned)gr, 256, 0, s>>>(in, data, nrows, tw, 1.f, band);
    k_plane_cols<L, -1><<<gc, 256, 0, s>>>(data, tw, scale, band, 0, L - 1, clo, chi);
  } else {
    k_plane_rows<L, +1><<<(unsigned)gr, 256, 0, s>>>(in, data, nrows, tw, 1.f, band);
    k_plane_cols<L, +1><<<gc, 256, 0, s>>>(data, tw, scale, band, 0, L - 1, clo, chi);
  }
  GS_CHECK_LAUNCH();
}

}  // namespace

bool launch_plane_fft2_from(const float2* in, float2* out, int batch, int L, int sign, float scale, hipStream_t s,
                            int band, int clo, int chi) {
  TwTables& T = tw_tables(current_device());
  if (band < 0 || band > L / 2) band = L / 2;
  if (L == 320) {
    plane_fft2_impl<320>(in, out, batch, sign, scale, T.t320, band, clo, chi < 0 ? L - 1 : chi, s);
    return true;
  }
  if (L == 720) {
    plane_fft2_impl<720>(in, out, batch, sign, scale, T.t720, band, clo, chi < 0 ? L - 1 : chi, s);
    return true;
  }
  return false;
}

bool launch_plane_fft2(float2* data, int batch, int L, int sign, float scale, hipStream_t s) {
  return launch_plane_fft2_from(data, data, batch, L, sign, scale, s);
}

namespace {
template <int L>
void plane_fft2_pairs_impl(const float2* in, float2* work, int npairs, int n, int stride, const float* stats,
                           float scale, float2* out, const float2* tw, hipStream_t s) {
  using C = CorrCfg<L>;
  constexpr int RPB = 4 * (64 / C::TPC), HC = 256 / C::TPC / 2;
  const long long nrows = (long long)npairs * L;
  const long long gr = (nrows + RPB - 1) / RPB;
  GS_REQUIRE(npairs > 0 && gr < (1LL << 31) && npairs < 65536 && stride >= 1 &&
                 (long long)((n + 1) / 2) * stride == npairs,
             kInvalid, "pair transform batch out of range");
  k_plane_rows<L, -1, true><<<(unsigned)gr, 256, 0, s>>>(in, work, nrows, tw, 1.f, L / 2);
  k_plane_cols_pairs<L><<<dim3((unsigned)(L / 2 / HC), (unsigned)npairs), 256, 0, s>>>(work, tw, n, stride, stats,
                                                                                        scale, out);
  GS_CHECK_LAUNCH();
}
}  // namespace

bool launch_plane_fft2_pairs(const float2* in, float2* work, int npairs, int L, int n, int stride,
                             const float* stats, float scale, float2* out, hipStream_t s) {
  TwTables& T = tw_tables(current_device());
  if (L == 320) {
    plane_fft2_pairs_impl<320>(in, work, npairs, n, stride, stats, scale, out, T.t320, s);
    return true;
  }
  if (L == 720) {
    plane_fft2_pairs_impl<720>(in, work, npairs, n, stride, stats, scale, out, T.t720, s);
    return true;
  }
  return false;
}

const float2* plane_twiddles(int L) {
  TwTables& T = tw_tables(current_device());
  return L == 320 ? T.t320 : (L == 720 ? T.t720 : nullptr);
}

// Partials per (template, block) written by the norm_type=0 row pass: one per row group and
// wavefront.
int rowpass_groups(int L) {
  if (L == 320) return 320 / CorrCfg<320>::ROWS * ((CorrCfg<320>::RT + 63) / 64);
  if (L == 720) return 720 / PlaneWaveCfg<720>::RCH;
  return 0;
}

template <int L, int Plan, int LAY, int BT, int TT, bool CGF = false>
static void colpass_tiled(const CorrArgs& a, const float2* twL, hipStream_t s) {
  using C = CorrCfg<L>;
  const long long G = (long long)(C::H / C::COLS) * ((a.tc + TT - 1) / TT * TT) * ((a.nb + BT - 1) / BT * BT);
  GS_REQUIRE(G > 0 && G < (1LL << 31), kInvalid, "column pass grid out of range");
  k_colpass<L, Plan, LAY, BT, TT, CGF><<<(unsigned)G, C::CT, 0, s>>>(a.that, a.ihat, twL, a.Z, a.t0, a.tc, a.nb);
}

template <int L>
static void colpass_impl(const CorrArgs& a, const float2* twL, int variant, hipStream_t s) {
  using C = CorrCfg<L>;
  if (variant == 0 || variant >= 3) {
    switch (variant) {
      case 3: colpass_tiled<L, 1, 0, 8, 16>(a, twL, s); break;
      case 4: colpass_tiled<L, 1, 1, 8, 16>(a, twL, s); break;
      case 5: colpass_tiled<L, 0, 1, 8, 16>(a, twL, s); break;
      case 6: colpass_tiled<L, 1, 2, 8, 16>(a, twL, s); break;
      case 7:  // recurrence twiddles need 16 lanes per column and a plan with T % NS == 0 (l = 320)
      case 8:  // the same with the row-packed layout (the row pass must be variant 7 or 8)
      case 10:  // 8 with the lane twiddles loaded into registers
      case 12:  // 8 with the column group the fastest index (4 x 4 tiles)
      case 13:  // 8 with the column group the fastest index (2 x 8 tiles)
        if constexpr (ct::lane_twiddles_fit<CorrCfg<L>::TPC>(typename ColPlanOf<L, 1>::type{}) &&
                      CorrCfg<L>::COLS == 16) {
          if (variant == 12)
            colpass_tiled<L, 1, 4, 4, 4, true>(a, twL, s);
          else if (variant == 13)
            colpass_tiled<L, 1, 4, 2, 8, true>(a, twL, s);
          else if (variant == 10)
            colpass_tiled<L, 1, 5, 8, 16>(a, twL, s);
          else if (variant == 8)
            colpass_tiled<L, 1, 4, 8, 16>(a, twL, s);
          else
            colpass_tiled<L, 1, 3, 8, 16>(a, twL, s);
        } else {
          colpass_tiled<L, 1, 0, 8, 16>(a, twL, s);
        }
        break;
      default: colpass_tiled<L, 0, 0, 8, 16>(a, twL, s); break;
    }
  } else {
    const unsigned G = (unsigned)((C::H / C::COLS) * a.nb);
    if (variant == 2)
      k_colpass_pers<L, true><<<G, C::CT, 0, s>>>(a.that, a.ihat, twL, a.Z, a.t0, a.tc, a.nb);
    else
      k_colpass_pers<L, false><<<G, C::CT, 0, s>>>(a.that, a.ihat, twL, a.Z, a.t0, a.tc, a.nb);
  }
  GS_CHECK_LAUNCH();
}

bool launch_colpass_plane(int L, const CorrArgs& a, int variant, hipStream_t s) {
  if (L != 720) return launch_colpass(L, a, variant, s);
  TwTables& T = tw_tables(current_device());
  using C = ColGroupCfg<720>;
  const long long G = (long long)(C::H / C::COLS) * a.nb;
  GS_REQUIRE(G > 0 && G < (1LL << 31), kInvalid, "column pass grid out of range");
  k_colpass_grouped<720><<<(unsigned)G, 64 * C::COLS, 0, s>>>(a.that, a.ihat, T.t720, a.Z, a.t0, a.tc, a.nb);
  GS_CHECK_LAUNCH();
  return true;
}

bool launch_colpass(int L, const CorrArgs& a, int variant, hipStream_t s) {
  TwTables& T = tw_tables(current_device());
  if (L == 320 && (variant == 9 || variant == 11)) {
    // 8 templates per workgroup, lane twiddles in registers (11: block columns held, next
    // template's columns prefetched)
    constexpr int TL = 8, BT = 8, TT = 8;
    const long long ntc = (a.tc + TL - 1) / TL;
    const long long G = (long long)(CorrCfg<320>::H / CorrCfg<320>::COLS) * ((ntc + TT - 1) / TT * TT) *
                        ((a.nb + BT - 1) / BT * BT);
    GS_REQUIRE(G > 0 && G < (1LL << 31), kInvalid, "column pass grid out of range");
    if (variant == 11)
      k_colpass_tl<320, TL, BT, TT, true><<<(unsigned)G, CorrCfg<320>::CT, 0, s>>>(a.that, a.ihat, T.t320, a.Z,
                                                                                 a.t0, a.tc, a.nb);
    else
      k_colpass_tl<320, TL, BT, TT><<<(unsigned)G, CorrCfg<320>::CT, 0, s>>>(a.that, a.ihat, T.t320, a.Z, a.t0,
                                                                           a.tc, a.nb);
    GS_CHECK_LAUNCH();
    return true;
  }
  if (L == 320) {
    colpass_impl<320>(a, T.t320, variant, s);
    return true;
  }
  if (L == 720) {
    colpass_impl<720>(a, T.t720, variant, s);
    return true;
  }
  return false;
}

template <int L, bool PLANE>
static void rowpass_impl(const float2* Z, int tc, int nb, double* S, double* SS, float* best,
                         int32_t* arg, int arg_base, float4* partial, float border,
                         const float2* twL, const float2* twH, hipStream_t s) {
  using C = CorrCfg<L>;
  dim3 grid(L / C::ROWS, nb);
  k_rowpass<L, PLANE><<<grid, C::RT, 0, s>>>(Z, twL, twH, tc, nb, S, SS, best, arg, arg_base,
                                             partial, border);
  GS_CHECK_LAUNCH();
}

bool launch_rowpass_pixel(int L, const float2* Z, int tc, int nb, double* S, double* SS,
                          float* best, int32_t* arg, int arg_base, int variant, hipStream_t s) {
  TwTables& T = tw_tables(current_device());
  if (L == 320) {
    // variants 0/1: 2 rows x 4 templates per wave iteration without/with prefetch; 2/3: 1 row x
    // 8 templates with/without prefetch; 4/5/6: as 1/0/2 with the bank-conflict-free LDS layout;
    // 7: as 6 reading the row-packed layout (column pass variant 8); 8 and 9: 7 and 6 with the
    // lazy argmax
    const int tpi = (variant == 2 || variant == 3 || (variant >= 6 && variant <= 9)) ? 8 : CorrCfg<320>::TPI;
    const int rows = 4 * (64 / CorrCfg<320>::TPT) / tpi;
    dim3 grid(320 / rows, nb);
    if (variant == 1)
      k_rowpass_pix_wave<320, true><<<grid, 256, 0, s>>>(Z, T.t320, T.t160, tc, nb, S, SS, best, arg, arg_base);
    else if (variant == 2)
      k_rowpass_pix_wave<320, true, 8><<<grid, 256, 0, s>>>(Z, T.t320, T.t160, tc, nb, S, SS, best, arg, arg_base);
    else if (variant == 3)
      k_rowpass_pix_wave<320, false, 8><<<grid, 256, 0, s>>>(Z, T.t320, T.t160, tc, nb, S, SS, best, arg, arg_base);
    else if (variant == 4)
      k_rowpass_pix_wave<320, true, 4, true><<<grid, 256, 0, s>>>(Z, T.t320, T.t160, tc, nb, S, SS, best, arg, arg_base);
    else if (variant == 5)
      k_rowpass_pix_wave<320, false, 4, true><<<grid, 256, 0, s>>>(Z, T.t320, T.t160, tc, nb, S, SS, best, arg, arg_base);
    else if (variant == 6)
      k_rowpass_pix_wave<320, true, 8, true><<<grid, 256, 0, s>>>(Z, T.t320, T.t160, tc, nb, S, SS, best, arg, arg_base);
    else if (variant == 7)
      k_rowpass_pix_wave<320, true, 8, true, true><<<grid, 256, 0, s>>>(Z, T.t320, T.t160, tc, nb, S, SS, best, arg, arg_base);
    else if (variant == 8)
      k_rowpass_pix_wave<320, true, 8, true, true, true><<<grid, 256, 0, s>>>(Z, T.t320, T.t160, tc, nb, S, SS, best, arg, arg_base);
    else if (variant == 9)
      k_rowpass_pix_wave<320, true, 8, true, false, true><<<grid, 256, 0, s>>>(Z, T.t320, T.t160, tc, nb, S, SS, best, arg, arg_base);
    else
      k_rowpass_pix_wave<320, false><<<grid, 256, 0, s>>>(Z, T.t320, T.t160, tc, nb, S, SS, best, arg, arg_base);
    GS_CHECK_LAUNCH();
    return true;
  }
  if (L == 720) {
    rowpass_impl<720, false>(Z, tc, nb, S, SS, best, arg, arg_base, nullptr, 0.f, T.t720, T.t360, s);
    return true;
  }
  return false;
}

bool launch_rowpass_chunk_records(int L, const float2* Z, int tc, int nb, const double* mu, const double* sigma,
                                  float thresh, int angle, int t0, RefRec* rec, int* count, int max_out,
                                  hipStream_t s) {
  if (L != 320) return false;
  TwTables& T = tw_tables(current_device());
  constexpr int rows = 4 * (64 / CorrCfg<320>::TPT) / 8;
  dim3 grid(320 / rows, nb);
  const ChunkRecArgs cr{mu, sigma, thresh, angle, t0, max_out, rec, count};
  k_rowpass_pix_wave<320, true, 8, true, true, false, true><<<grid, 256, 0, s>>>(
      Z, T.t320, T.t160, tc, nb, nullptr, nullptr, nullptr, nullptr, 0, cr);
  GS_CHECK_LAUNCH();
  return true;
}

bool launch_rowpass_plane(int L, const float2* Z, int tc, int nb, float4* partial, float border,
                          hipStream_t s) {
  TwTables& T = tw_tables(current_device());
  if (L == 320) {
    rowpass_impl<320, true>(Z, tc, nb, nullptr, nullptr, nullptr, nullptr, 0, partial, border, T.t320, T.t160, s);
    return true;
  }
  if (L == 720) {
    using C = PlaneWaveCfg<720>;
    const long long waves = (long long)tc * nb * (720 / C::RCH);
    const long long blocks = (waves + C::WAVES - 1) / C::WAVES;
    GS_REQUIRE(blocks > 0 && blocks < (1LL << 31), kInvalid, "row pass grid out of range");
    k_rowpass_plane_wave<720><<<(unsigned)blocks, 64 * C::WAVES, 0, s>>>(Z, T.t720, T.t360, tc, nb, border, partial);
    GS_CHECK_LAUNCH();
    return true;
  }
  return false;
}

void launch_finalize_plane(const float4* partial, int tc, int nb, int groups, long long tstride, double* out,
                           hipStream_t s) {
  const int n = tc * nb;  // one wavefront each, four per workgroup
  k_finalize_plane<<<(n + 3) / 4, 256, 0, s>>>(partial, tc, nb, groups, tstride, out);
  GS_CHECK_LAUNCH();
}

void launch_init_accumulators(double* S, double* SS, float* best, int32_t* arg, long long count,
                              hipStream_t s) {
  k_init_acc<<<(unsigned)((count + 255) / 256), 256, 0, s>>>(S, SS, best, arg, count);
  GS_CHECK_LAUNCH();
}

void launch_finalize_norm(const double* S, const double* SS, const float* best, const int32_t* arg,
                          int nb, int l, double total_n, float thresh, const int* offx,
                          const int* offy, int bx, int W, int H, float d_m, double* mu,
                          double* sigma, Cand* cand, int* count, int max_cand, hipStream_t s) {
  const long long n = (long long)nb * l * l;
  k_finalize_norm<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(S, SS, best, arg, nb, l, total_n, thresh,
                                                              offx, offy, bx, W, H, d_m, mu, sigma,
                                                              cand, count, max_cand);
  GS_CHECK_LAUNCH();
}

}  // namespace gs
