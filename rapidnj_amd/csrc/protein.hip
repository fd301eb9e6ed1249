// HIP kernels for protein alignments on gfx950: the residue census, packing into code bit
// planes, and the pair-tile kernel with the fused Kimura-protein / Jukes-Cantor (alpha = 20)
// correction.
//
// The reference compares residues byte by byte after mapping the gap characters
// - . X x z Z b B J j ? to '-' (bitDistanceProtein.cpp:31-61 over dataloader::resolveChar):
// a site is a mismatch when neither character is a gap and they differ, and the estimator's
// third count is the sites where neither is a gap. Here every byte value present in the
// alignment gets a code (gaps 0, the other values 1..S in byte order), stored bit-sliced:
//   planes[0][w][t]      gap bits: 1 where site 32w + b of taxon t is a gap
//   planes[1 + p][w][t]  bit p of the site's code, p < B (B = 5, 6 or 8 for S < 32, 64, 256)
// plus gaps[t], the taxon's gap sites. Padding (sites at or past L, taxa at or past n) is code 0
// without the gap bit. For a pair and 32 sites the kernel counts
//   x = OR_p (ci_p ^ cj_p)   sites whose codes differ (mismatches, and one-sided gaps)
//   g = gi & gj              sites where both are gaps
// in B + 3 VALU instructions (v_xor, B - 1 v_bitop3, v_and, 2 v_bcnt), and the epilogue recovers
//   mismatches = #x - (gaps_i + gaps_j - 2 #g),  residue pairs = L - gaps_i - gaps_j + #g.
// There is no matrix-core work: the counting is bitwise and integer.
#include "device_common.hpp"

#include <algorithm>

namespace rnj {
namespace {

// The Dayhoff PAM distances for p = 0.750, 0.751, .., 0.930 as the reference tabulates them
// (KimuraDistance.cpp:30-49, integers in PAM/100 units), converted as the reference converts
// them, (float)(pam / 100.0). The values have to be these to give the reference's distances.
#define RNJ_PAM(x) (float)((x) / 100.0)
__constant__ float kDayhoff[181] = {
    RNJ_PAM(195), RNJ_PAM(196), RNJ_PAM(197), RNJ_PAM(198), RNJ_PAM(199), RNJ_PAM(200),
    RNJ_PAM(200), RNJ_PAM(201), RNJ_PAM(202), RNJ_PAM(203), RNJ_PAM(204), RNJ_PAM(205),
    RNJ_PAM(206), RNJ_PAM(207), RNJ_PAM(208), RNJ_PAM(209), RNJ_PAM(209), RNJ_PAM(210),
    RNJ_PAM(211), RNJ_PAM(212), RNJ_PAM(213), RNJ_PAM(214), RNJ_PAM(215), RNJ_PAM(216),
    RNJ_PAM(217), RNJ_PAM(218), RNJ_PAM(219), RNJ_PAM(220), RNJ_PAM(221), RNJ_PAM(222),
    RNJ_PAM(223), RNJ_PAM(224), RNJ_PAM(226), RNJ_PAM(227), RNJ_PAM(228), RNJ_PAM(229),
    RNJ_PAM(230), RNJ_PAM(231), RNJ_PAM(232), RNJ_PAM(233), RNJ_PAM(234), RNJ_PAM(236),
    RNJ_PAM(237), RNJ_PAM(238), RNJ_PAM(239), RNJ_PAM(240), RNJ_PAM(241), RNJ_PAM(243),
    RNJ_PAM(244), RNJ_PAM(245), RNJ_PAM(246), RNJ_PAM(248), RNJ_PAM(249), RNJ_PAM(250),
    RNJ_PAM(252), RNJ_PAM(253), RNJ_PAM(254), RNJ_PAM(255), RNJ_PAM(257), RNJ_PAM(258),
    RNJ_PAM(260), RNJ_PAM(261), RNJ_PAM(262), RNJ_PAM(264), RNJ_PAM(265), RNJ_PAM(267),
    RNJ_PAM(268), RNJ_PAM(270), RNJ_PAM(271), RNJ_PAM(273), RNJ_PAM(274), RNJ_PAM(276),
    RNJ_PAM(277), RNJ_PAM(279), RNJ_PAM(281), RNJ_PAM(282), RNJ_PAM(284), RNJ_PAM(285),
    RNJ_PAM(287), RNJ_PAM(289), RNJ_PAM(291), RNJ_PAM(292), RNJ_PAM(294), RNJ_PAM(296),
    RNJ_PAM(298), RNJ_PAM(299), RNJ_PAM(301), RNJ_PAM(303), RNJ_PAM(305), RNJ_PAM(307),
    RNJ_PAM(309), RNJ_PAM(311), RNJ_PAM(313), RNJ_PAM(315), RNJ_PAM(317), RNJ_PAM(319),
    RNJ_PAM(321), RNJ_PAM(323), RNJ_PAM(325), RNJ_PAM(328), RNJ_PAM(330), RNJ_PAM(332),
    RNJ_PAM(335), RNJ_PAM(337), RNJ_PAM(339), RNJ_PAM(342), RNJ_PAM(344), RNJ_PAM(347),
    RNJ_PAM(349), RNJ_PAM(352), RNJ_PAM(354), RNJ_PAM(357), RNJ_PAM(360), RNJ_PAM(362),
    RNJ_PAM(365), RNJ_PAM(368), RNJ_PAM(371), RNJ_PAM(374), RNJ_PAM(377), RNJ_PAM(380),
    RNJ_PAM(383), RNJ_PAM(386), RNJ_PAM(389), RNJ_PAM(393), RNJ_PAM(396), RNJ_PAM(399),
    RNJ_PAM(403), RNJ_PAM(407), RNJ_PAM(410), RNJ_PAM(414), RNJ_PAM(418), RNJ_PAM(422),
    RNJ_PAM(426), RNJ_PAM(430), RNJ_PAM(434), RNJ_PAM(438), RNJ_PAM(442), RNJ_PAM(447),
    RNJ_PAM(451), RNJ_PAM(456), RNJ_PAM(461), RNJ_PAM(466), RNJ_PAM(471), RNJ_PAM(476),
    RNJ_PAM(482), RNJ_PAM(487), RNJ_PAM(493), RNJ_PAM(498), RNJ_PAM(504), RNJ_PAM(511),
    RNJ_PAM(517), RNJ_PAM(524), RNJ_PAM(531), RNJ_PAM(538), RNJ_PAM(545), RNJ_PAM(553),
    RNJ_PAM(560), RNJ_PAM(569), RNJ_PAM(577), RNJ_PAM(586), RNJ_PAM(595), RNJ_PAM(605),
    RNJ_PAM(615), RNJ_PAM(626), RNJ_PAM(637), RNJ_PAM(649), RNJ_PAM(661), RNJ_PAM(675),
    RNJ_PAM(688), RNJ_PAM(703), RNJ_PAM(719), RNJ_PAM(736), RNJ_PAM(754), RNJ_PAM(775),
    RNJ_PAM(796), RNJ_PAM(819), RNJ_PAM(845), RNJ_PAM(874), RNJ_PAM(907), RNJ_PAM(945),
    RNJ_PAM(988)};
#undef RNJ_PAM

// The tables the corrections read, staged in LDS by each work-group: the log table and the
// Dayhoff distances. Every lane reads its own entry, so an LDS copy serves the divergent
// lookups that a read of the constant table would turn into a branch around a global load.
struct ProteinTables {
  LogEntry log[128];
  float day[184];
};

__device__ __forceinline__ void stage_protein_tables(ProteinTables& t, unsigned tid,
                                                     unsigned nthreads) {
  stage_log_table(t.log, tid, nthreads);
  for (unsigned k = tid; k < 181; k += nthreads) t.day[k] = kDayhoff[k];
}

// ---- corrections ---------------------------------------------------------------------------
// mism and valid are the reference's data[0] (data[1] is 0 for protein) and data[2], valid being
// its 64-bit count. Proportions are float quotients of the float-converted counts, as the
// reference's `total / (distType) data[2]` computes them. kExact takes the library log; the
// default log_table rounds to the same float except in rare ties (see log_table).
#pragma clang fp contract(off)
// (float)mism / (float)valid. Below 2^24 both convert exactly and the float division is the
// correctly rounded quotient of the integers, which the default path forms from the FP64
// reciprocal as proportions2 does (same result, without the division sequence).
template <bool kExact>
__device__ __forceinline__ float protein_proportion(uint32_t mism, unsigned long long valid) {
  if (!kExact && valid < (1u << 24) && mism < (1u << 24)) {
    const double dv = (double)valid;
    double r = __builtin_amdgcn_rcp(dv);
    r = fma(r, fma(-dv, r, 1.0), r);
    r = fma(r, fma(-dv, r, 1.0), r);
    return (float)((double)mism * r);
  }
  return (float)mism / (float)valid;
}

// KimuraDistance.cpp:322-338: -1 when valid is 0, -log(1 - p - 0.2 p^2) below 0.75 (p * p in
// float, the rest in double), the Dayhoff table up to 0.930 and 10 above. Straight-line (every
// case computed, the result selected) so that a caller can interleave several pairs.
// kTable = false leaves the Dayhoff part out: for callers that know p < 0.75.
template <bool kExact, bool kTable = true>
__device__ __forceinline__ float kimura_protein_of_p(float p, const LogConsts& K,
                                                    const float* day) {
  const bool formula = p < 0.75f && p > 0.0f;
  const float pp = p * p;
  const double arg = formula ? (1.0 - (double)p) - (double)pp * 0.20 : 1.0;
  const double lg = kExact ? log(arg) : log_table_one(arg, K);
  const float low = formula ? (float)(-lg) : p;
  if constexpr (!kTable) return low;
  int idx = (int)((double)p * 1000.0 - 750.0);
  idx = idx < 0 ? 0 : (idx > 180 ? 180 : idx);
  float tab = day[idx];
  asm volatile("" : "+v"(tab));  // read for every pair: kept out of a branch on p
  const float table = (double)p > 0.930 ? 10.0f : tab;
  return p < 0.75f ? low : table;
}

template <bool kExact>
__device__ __forceinline__ float kimura_protein_value(uint32_t mism, unsigned long long valid,
                                                     const LogConsts& K, const float* day) {
  return kimura_protein_of_p<kExact>(valid != 0 ? protein_proportion<kExact>(mism, valid) : -1.0f,
                                     K, day);
}

// JCdistance.cpp:170-183 with alpha = 20: p = 0 when valid is 0, -1 when p / 0.95 >= 1 (the
// quotient in double, as the reference forms it, so that the saturated pairs are the same).
// The default path tests p >= b instead (the same test: b = 0.95 rounded to double is no float,
// so a float p above it gives a quotient of at least 1 and one below it a quotient below
// 1 - 2^-25) and forms 1 - p/b as fma(-p, 1/b, 1), within 2^-52 absolute, as the DNA JC does.
template <bool kExact>
__device__ __forceinline__ float jc_protein_of_p(float p, const LogConsts& K) {
  const double b = 19.0 / 20.0;  // (alpha - 1) / alpha
  if (kExact) {
    const double q = (double)p / b;
    const bool sat = q >= 1.0;
    return sat ? -1.0f : (float)(-b * log(sat ? 1.0 : 1.0 - q));
  }
  const bool sat = (double)p >= b;
  const double lg = log_table_one(fma(-(double)p, 20.0 / 19.0, 1.0), K);
  return sat ? -1.0f : (float)(-b * lg);
}

template <bool kExact>
__device__ __forceinline__ float jc_protein_value(uint32_t mism, unsigned long long valid,
                                                 const LogConsts& K) {
  return jc_protein_of_p<kExact>(valid != 0 ? protein_proportion<kExact>(mism, valid) : 0.0f, K);
}
#pragma clang fp contract(on)

// ---- epilogue ------------------------------------------------------------------------------
// Thread (ty, tx) of tile (bi, bj) holds x (differing sites) and g (gap-gap sites) of its
// MI x MJ pairs. The reference walks 128 * (L/128 + 1) positions and counts the padding past L
// as gaps, then subtracts from L rounded up to 128 (bitDistanceProtein.cpp:31-61): its valid
// count is the residue pairs, less 128 when L is a multiple of 128 (modulo 2^64). J.bias_rule
// selects that shortfall (it is 0 for the -f path's Hamming counter, which counts the residue
// pairs, simpleDistanceCalculator.cpp:40-55); a bootstrap replicate's length is read from
// J.len_dev.
template <int TILE, int MI, int MJ, bool kExact>
__device__ __forceinline__ void protein_epilogue(const JobDesc& J, const ProteinTables& T,
                                                 unsigned tile, unsigned bi,
                                                 unsigned bj, unsigned ty, unsigned tx,
                                                 unsigned tid,
                                                 const uint32_t (&acc_x)[MI][MJ],
                                                 const uint32_t (&acc_g)[MI][MJ]) {
  const unsigned n = J.n;
  const unsigned long long L = J.len_dev ? *J.len_dev : J.L;
  const unsigned long long bias = (J.bias_rule && L % 128u == 0) ? 128ull : 0ull;
  uint32_t gi[MI], gj[MJ];
#pragma unroll
  for (int a = 0; a < MI; ++a) gi[a] = J.gaps[bi * TILE + ty * MI + a];
#pragma unroll
  for (int b = 0; b < MJ; ++b) gj[b] = J.gaps[bj * TILE + tx * MJ + b];
  auto counts = [&](int a, int b, uint32_t& mism, unsigned long long& valid) {
    const unsigned long long gsum = (unsigned long long)gi[a] + gj[b];
    mism = (uint32_t)(acc_x[a][b] + 2ull * acc_g[a][b] - gsum);
    valid = L - gsum + acc_g[a][b] - bias;
  };
  if (J.model == kCounts) {
#pragma unroll
    for (int a = 0; a < MI; ++a) {
      const unsigned i = bi * TILE + ty * MI + a;
#pragma unroll
      for (int b = 0; b < MJ; ++b) {
        const unsigned j = bj * TILE + tx * MJ + b;
        if (j >= n || i >= j) continue;
        uint32_t mism;
        unsigned long long valid;
        counts(a, b, mism, valid);
        const unsigned long long p = tri_index(i, j, n);
        J.out_ts[p] = mism;
        J.out_tv[p] = 0;
        J.out_valid[p] = (uint32_t)valid;
      }
    }
    return;
  }
  float lmax = 0.0f;
  uint32_t sat = 0;  // saturated pairs of the micro-tile, bit a * MJ + b
  const LogConsts K = log_consts(T.log);
  // Inner tiles of an upper-triangle output with fewer than 2^24 sites: the counts in 32-bit
  // arithmetic and each row's corrections as straight-line code, stored from one triangle index.
  // The valid count is then s = L - bias - (gaps_i + gaps_j - #g) in int32; the reference's
  // 64-bit count wraps when s < 0 (fewer residue pairs than the 128 of the bias), and its float
  // is 2^64, so the proportion is mism * 2^-64 exactly (as the float division gives).
  if (!kExact && J.layout == kTriangle && bi < bj && (bj + 1) * TILE <= n && L < (1u << 24)) {
    const int lb = (int)(L - bias);
    const unsigned j0 = bj * TILE + tx * MJ;
    auto rows = [&](auto model) {
#pragma unroll
      for (int a = 0; a < MI; ++a) {
        const unsigned i = bi * TILE + ty * MI + a;
        float d[MJ];
#pragma unroll
        for (int b = 0; b < MJ; ++b) {
          const uint32_t gsum = gi[a] + gj[b];
          const uint32_t mism = acc_x[a][b] + 2u * acc_g[a][b] - gsum;
          const int s = lb - (int)gsum + (int)acc_g[a][b];
          float pq = (float)((double)mism * reciprocal(s > 0 ? (uint32_t)s : 1u));
          // formed for every pair: kept out of a branch on s, so the row's pairs interleave
          asm volatile("" : "+v"(pq));
          const float pw = __builtin_ldexpf((float)mism, -64);
          if constexpr (decltype(model)::value == kJC)
            d[b] = jc_protein_of_p<false>(s == 0 ? 0.0f : (s < 0 ? pw : pq), K);
          else
            d[b] = s == 0 ? -1.0f : (s < 0 ? pw : pq);  // p, corrected below
        }
        if constexpr (decltype(model)::value != kJC) {
          // the Dayhoff part only in waves where some pair of the row has p >= 0.75
          bool far = false;
#pragma unroll
          for (int b = 0; b < MJ; ++b) far |= d[b] >= 0.75f;
          if (__ballot(far) != 0ull) {
#pragma unroll
            for (int b = 0; b < MJ; ++b) d[b] = kimura_protein_of_p<false>(d[b], K, T.day);
          } else {
#pragma unroll
            for (int b = 0; b < MJ; ++b) d[b] = kimura_protein_of_p<false, false>(d[b], K, T.day);
          }
        }
        const gptr_mut<float> row = (gptr_mut<float>)J.out + tri_index(i, j0, n);
#pragma unroll
        for (int b = 0; b < MJ; ++b) {
          row[b] = d[b];
          if (d[b] == -1.0f) sat |= 1u << (a * MJ + b);
          lmax = fmaxf(lmax, d[b]);  // a saturated pair's -1 leaves it as it is
        }
      }
    };
    if (J.model == kJC) rows(std::integral_constant<int, kJC>());
    else rows(std::integral_constant<int, kK2P>());
    wave_record(J, tile * ((TILE / MI) * (TILE / MJ) / 64) + tid / 64, bi << 16 | bj, sat, lmax);
    return;
  }
#pragma unroll
  for (int a = 0; a < MI; ++a) {
    const unsigned i = bi * TILE + ty * MI + a;
    float d[MJ];
#pragma unroll
    for (int b = 0; b < MJ; ++b) {
      uint32_t mism;
      unsigned long long valid;
      counts(a, b, mism, valid);
      d[b] = J.model == kJC ? jc_protein_value<kExact>(mism, valid, K)
                            : kimura_protein_value<kExact>(mism, valid, K, T.day);
    }
#pragma unroll
    for (int b = 0; b < MJ; ++b) {
      const unsigned j = bj * TILE + tx * MJ + b;
      if (j >= n) continue;
      if (i < j) {
        if (d[b] == -1.0f) sat |= 1u << (a * MJ + b);
        lmax = fmaxf(lmax, d[b]);  // a saturated pair's -1 leaves it as it is
        store_pair(J, i, j, d[b]);  // protein jobs are never kRows (their counts are symmetric)
      } else if (i == j) {
        finish_diagonal(J, i);
      }
    }
  }
  wave_record(J, tile * ((TILE / MI) * (TILE / MJ) / 64) + tid / 64, bi << 16 | bj, sat, lmax);
}

// ---- the pair-tile kernel ------------------------------------------------------------------
// Waves per SIMD the register allocation aims for: 4 for 64-pair tiles with 8-word chunks (the
// epilogue's interleaved corrections would otherwise take it to 130 VGPRs and 3 waves: 1.95 ms
// against 1.82 ms at 20000 x 1250, for a 12-byte spill), otherwise no constraint.
template <int TILE, int CW, bool kExact>
constexpr int kProteinWaves = (TILE == 64 && CW == 8 && !kExact) ? 4 : 1;

// The DNA tile kernel's schedule (kernels.hip) over 1 + B planes: one work-group per TILE x TILE
// block of pairs and range of words, chunks of CW words of both blocks staged in LDS as
// [plane][word][taxon], the next chunk prefetched into registers while this one is counted.
template <int TILE, int MI, int MJ, int CW, int B, bool kExact>
__global__ __launch_bounds__((TILE / MI) * (TILE / MJ))
__attribute__((amdgpu_waves_per_eu(kProteinWaves<TILE, CW, kExact>))) void protein_tile_kernel(
    const JobDesc* __restrict__ jobs, unsigned njobs) {
  constexpr int P = 1 + B;
  constexpr int NTX = TILE / MJ;
  constexpr int NT = (TILE / MI) * NTX;
  constexpr int Q = TILE / 4;
  constexpr int E = P * CW * Q;
  static_assert(MI % 4 == 0 && MJ % 4 == 0, "micro-tiles are read as uint4");
  constexpr int EP = (E + NT - 1) / NT;

  __shared__ v4u sI[E];
  __shared__ v4u sJ[E];
  __shared__ ProteinTables sTab;

  const unsigned item = xcd_item(blockIdx.x, gridDim.x);
  const unsigned k = find_job(jobs, njobs, item);
  const JobDesc& J = jobs[k];
  const unsigned local = item - J.item_base;
  const unsigned tile = local / J.splits, split = local % J.splits;
  unsigned bi, bj;
  patch_coords(tile, J.nblk, bi, bj);

  const unsigned tid = threadIdx.x;
  const unsigned ty = tid / NTX, tx = tid % NTX;
  stage_protein_tables(sTab, tid, NT);  // visible after the chunk loop's first barrier
  const gptr<v4u> gI = (gptr<v4u>)(J.planes + (size_t)bi * TILE);
  const gptr<v4u> gJ = (gptr<v4u>)(J.planes + (size_t)bj * TILE);
  const size_t row4 = J.n_pad / 4;
  const size_t plane4 = J.plane_stride / 4;

  uint32_t acc_x[MI][MJ], acc_g[MI][MJ];
#pragma unroll
  for (int a = 0; a < MI; ++a)
#pragma unroll
    for (int b = 0; b < MJ; ++b) acc_x[a][b] = acc_g[a][b] = 0;

  // this range's words that hold sites (padded words are code 0 without the gap bit and would
  // count nothing); ranges start on chunk boundaries, the last chunk may be partial
  const unsigned w0 = split * J.words_per_split;
  const unsigned w1 = min(J.W_live, w0 + J.words_per_split);
  const unsigned c0 = w0 / CW, c1 = (w1 + CW - 1) / CW;

  v4u pI[EP], pJ[EP];
  if (c0 < c1) fetch_chunk<EP, NT, Q, CW, E>(pI, pJ, gI, gJ, tid, c0, row4, plane4);
  for (unsigned c = c0; c < c1; ++c) {
    __syncthreads();
#pragma unroll
    for (int e = 0; e < EP; ++e) {
      if (E % NT != 0 && tid + e * NT >= (unsigned)E) break;
      sI[tid + e * NT] = pI[e];
      sJ[tid + e * NT] = pJ[e];
    }
    __syncthreads();
    if (c + 1 < c1) fetch_chunk<EP, NT, Q, CW, E>(pI, pJ, gI, gJ, tid, c + 1, row4, plane4);

    const int live = (int)min((unsigned)CW, w1 - c * CW);  // uniform over the work-group
#pragma unroll 1
    for (int w = 0; w < CW; ++w) {
      if (w >= live) break;
      uint32_t ci[P][MI], cj[P][MJ];
#pragma unroll
      for (int p = 0; p < P; ++p) {
#pragma unroll
        for (int r = 0; r < MI / 4; ++r) {
          const v4u v = sI[(p * CW + w) * Q + ty * (MI / 4) + r];
          ci[p][4 * r] = v.x; ci[p][4 * r + 1] = v.y; ci[p][4 * r + 2] = v.z; ci[p][4 * r + 3] = v.w;
        }
#pragma unroll
        for (int r = 0; r < MJ / 4; ++r) {
          const v4u v = sJ[(p * CW + w) * Q + tx * (MJ / 4) + r];
          cj[p][4 * r] = v.x; cj[p][4 * r + 1] = v.y; cj[p][4 * r + 2] = v.z; cj[p][4 * r + 3] = v.w;
        }
      }
#pragma unroll
      for (int a = 0; a < MI; ++a)
#pragma unroll
        for (int b = 0; b < MJ; ++b) {
          uint32_t x = ci[1][a] ^ cj[1][b];
#pragma unroll
          for (int p = 2; p < P; ++p)  // x | (ci ^ cj)
            x = __builtin_amdgcn_bitop3_b32(x, ci[p][a], cj[p][b], 0xF6u);
          acc_x[a][b] = bcnt_add(x, acc_x[a][b]);
          acc_g[a][b] = bcnt_add(ci[0][a] & cj[0][b], acc_g[a][b]);
        }
    }
  }

  if (J.splits > 1) {
    uint32_t* slot = partial_slot<NT, MI, MJ>(J, tile, split) + (size_t)tid * MI * MJ;
    const size_t plane = (size_t)NT * MI * MJ;
#pragma unroll
    for (int a = 0; a < MI; ++a)
#pragma unroll
      for (int b = 0; b < MJ; ++b) {
        slot[a * MJ + b] = acc_x[a][b];
        slot[plane + a * MJ + b] = acc_g[a][b];
      }
    return;
  }
  if (c0 >= c1) __syncthreads();  // no chunk, so no barrier yet since the table was staged
  protein_epilogue<TILE, MI, MJ, kExact>(J, sTab, tile, bi, bj, ty, tx, tid, acc_x, acc_g);
}

// Sums the partial counts of the splits of one tile and finishes it (kernels.hip numbering).
template <int TILE, int MI, int MJ, bool kExact>
__global__ __launch_bounds__((TILE / MI) * (TILE / MJ)) void protein_finish_kernel(
    const JobDesc* __restrict__ jobs, unsigned njobs) {
  constexpr int NTX = TILE / MJ;
  constexpr int NT = (TILE / MI) * NTX;
  const unsigned item = blockIdx.x;
  unsigned lo = 0, hi = njobs - 1;
  while (lo < hi) {
    const unsigned mid = (lo + hi + 1) / 2;
    if (jobs[mid].fin_base <= item) lo = mid; else hi = mid - 1;
  }
  const JobDesc& J = jobs[lo];
  const unsigned tile = item - J.fin_base;
  if (J.splits <= 1 || tile >= J.tiles) return;  // uniform over the work-group
  __shared__ ProteinTables sTab;
  stage_protein_tables(sTab, threadIdx.x, NT);
  __syncthreads();
  unsigned bi, bj;
  patch_coords(tile, J.nblk, bi, bj);
  const unsigned tid = threadIdx.x;
  const unsigned ty = tid / NTX, tx = tid % NTX;
  uint32_t acc_x[MI][MJ], acc_g[MI][MJ];
#pragma unroll
  for (int a = 0; a < MI; ++a)
#pragma unroll
    for (int b = 0; b < MJ; ++b) acc_x[a][b] = acc_g[a][b] = 0;
  const size_t plane = (size_t)NT * MI * MJ;
  for (unsigned sp = 0; sp < J.splits; ++sp) {
    const uint32_t* slot = partial_slot<NT, MI, MJ>(J, tile, sp) + (size_t)tid * MI * MJ;
#pragma unroll
    for (int a = 0; a < MI; ++a)
#pragma unroll
      for (int b = 0; b < MJ; ++b) {
        acc_x[a][b] += slot[a * MJ + b];
        acc_g[a][b] += slot[plane + a * MJ + b];
      }
  }
  protein_epilogue<TILE, MI, MJ, kExact>(J, sTab, tile, bi, bj, ty, tx, tid, acc_x, acc_g);
}

template <int TILE, int MI, int MJ, int CW, int B>
hipError_t launch_protein_cfg(bool exact, const JobDesc* jobs, unsigned njobs, unsigned total,
                              unsigned finish_items, hipStream_t s) {
  constexpr int NT = (TILE / MI) * (TILE / MJ);
  const dim3 grid(total), block(NT);
  if (exact) protein_tile_kernel<TILE, MI, MJ, CW, B, true><<<grid, block, 0, s>>>(jobs, njobs);
  else protein_tile_kernel<TILE, MI, MJ, CW, B, false><<<grid, block, 0, s>>>(jobs, njobs);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || finish_items == 0) return e;
  const dim3 fgrid(finish_items);
  if (exact) protein_finish_kernel<TILE, MI, MJ, true><<<fgrid, block, 0, s>>>(jobs, njobs);
  else protein_finish_kernel<TILE, MI, MJ, false><<<fgrid, block, 0, s>>>(jobs, njobs);
  return hipGetLastError();
}

template <int B>
hipError_t launch_protein_b(int config, bool exact, const JobDesc* jobs, unsigned njobs,
                            unsigned total, unsigned f, hipStream_t s) {
  switch (config) {  // the configurations of kTileConfigs
    case 0: return launch_protein_cfg<64, 4, 4, 16, B>(exact, jobs, njobs, total, f, s);
    case 1: return launch_protein_cfg<64, 4, 4, 8, B>(exact, jobs, njobs, total, f, s);
    case 2: return launch_protein_cfg<128, 8, 4, 16, B>(exact, jobs, njobs, total, f, s);
    case 3: return launch_protein_cfg<32, 4, 4, 16, B>(exact, jobs, njobs, total, f, s);
    case 4: return launch_protein_cfg<32, 4, 4, 8, B>(exact, jobs, njobs, total, f, s);
    default: return hipErrorInvalidValue;
  }
}

// ---- census and packing ----------------------------------------------------------------------
// Byte values present in the characters: a 256-bit set per work-group in LDS (a bit is tested
// before the atomic, so a work-group issues about one atomic per distinct value), or-ed into
// `present` at the end.
__global__ __launch_bounds__(256) void symbol_census_kernel(const uint8_t* __restrict__ chars,
                                                            size_t pitch, unsigned n, unsigned L,
                                                            uint32_t* __restrict__ present) {
  __shared__ uint32_t seen[8];
  if (threadIdx.x < 8) seen[threadIdx.x] = 0;
  __syncthreads();
  for (unsigned r = blockIdx.y; r < n; r += gridDim.y) {
    const uint8_t* row = chars + (size_t)r * pitch;
    for (unsigned c = blockIdx.x * blockDim.x + threadIdx.x; c < L; c += gridDim.x * blockDim.x) {
      const unsigned v = row[c];
      const uint32_t bit = 1u << (v & 31);
      if (!(seen[v >> 5] & bit)) atomicOr(&seen[v >> 5], bit);
    }
  }
  __syncthreads();
  if (threadIdx.x < 8 && seen[threadIdx.x]) atomicOr(&present[threadIdx.x], seen[threadIdx.x]);
}

// One work-group packs 64 taxa x 1024 sites, staged through LDS like the DNA packer
// (kernels.hip): rows read coalesced into padded LDS rows, lane t of each wave builds taxon t's
// plane words, stores coalesced over taxa, and adds the taxon's gap sites to gaps[t].
constexpr unsigned kPTaxa = 64, kPWords = 32, kPRow = kPWords * 8 + 1;

__global__ __launch_bounds__(256) void pack_protein_kernel(
    const uint8_t* __restrict__ chars, size_t pitch, unsigned n, unsigned L, unsigned n_pad,
    unsigned W_pad, const uint8_t* __restrict__ codes, unsigned B, uint32_t* __restrict__ planes,
    uint32_t* __restrict__ gaps) {
  __shared__ uint32_t tile[kPTaxa * kPRow];
  __shared__ uint8_t table[256];
  table[threadIdx.x] = codes[threadIdx.x];
  const unsigned t0 = blockIdx.x * kPTaxa;
  const unsigned lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const size_t ps = (size_t)W_pad * n_pad;
  const unsigned t = t0 + lane;
  uint32_t gap_sites = 0;
  for (unsigned w0 = blockIdx.y * kPWords; w0 < W_pad; w0 += gridDim.y * kPWords) {
    __syncthreads();  // table written / the previous word block's tile consumed
    for (unsigned e = threadIdx.x; e < kPTaxa * kPWords * 8; e += blockDim.x) {
      const unsigned r = e / (kPWords * 8), q = e % (kPWords * 8);
      const size_t s = (size_t)w0 * 32 + 4 * q;
      uint32_t x = 0;
      if (t0 + r < n) {
        const uint8_t* row = chars + (size_t)(t0 + r) * pitch;
#pragma unroll
        for (unsigned k = 0; k < 4; ++k)
          if (s + k < L) x |= (uint32_t)row[s + k] << (8 * k);
      }
      tile[r * kPRow + q] = x;
    }
    __syncthreads();
    for (unsigned ww = wave; ww < kPWords; ww += blockDim.x / 64) {
      const unsigned w = w0 + ww;
      if (w >= W_pad) break;
      uint32_t G = 0, C[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
      for (unsigned j = 0; j < 8; ++j) {
        const uint32_t x = tile[lane * kPRow + ww * 8 + j];
#pragma unroll
        for (unsigned k = 0; k < 4; ++k) {
          const uint32_t c = table[(x >> (8 * k)) & 255u];
          const unsigned bit = 4 * j + k;
          G |= (uint32_t)(c == 0) << bit;
#pragma unroll
          for (unsigned p = 0; p < 8; ++p) C[p] |= ((c >> p) & 1u) << bit;
        }
      }
      // sites past L and taxa past n: code 0, no gap
      const unsigned long long first = (unsigned long long)w * 32;
      uint32_t keep = first >= L ? 0u : (L - first >= 32 ? ~0u : (1u << (L - first)) - 1u);
      if (t >= n) keep = 0;
      G &= keep;
      gap_sites += __popc(G);
      if (t < n_pad) {
        planes[(size_t)w * n_pad + t] = G;
#pragma unroll
        for (unsigned p = 0; p < 8; ++p)
          if (p < B) planes[(1 + p) * ps + (size_t)w * n_pad + t] = C[p] & keep;
      }
    }
  }
  if (t < n && gap_sites) atomicAdd(&gaps[t], gap_sites);
}

// Per-site code bytes of a protein alignment for the column gathers: codes[s][t], the residue
// code of taxon t at site s (0 for a gap), one thread per (site, 4 taxa), one u32 of output.
__global__ void protein_codes_kernel(const uint32_t* __restrict__ planes, unsigned n_pad,
                                     unsigned W_pad, unsigned L, unsigned B,
                                     uint32_t* __restrict__ codes) {
  const unsigned g = blockIdx.x * blockDim.x + threadIdx.x;  // group of 4 taxa
  const unsigned groups = n_pad / 4;
  if (g >= groups) return;
  const size_t ps = (size_t)W_pad * n_pad;
  for (unsigned s = blockIdx.y; s < L; s += gridDim.y) {
    const size_t o = (size_t)(s / 32) * n_pad + 4 * g;
    const unsigned sh = s % 32;
    uint32_t x = 0;
    for (unsigned j = 0; j < 4; ++j) {
      uint32_t c = 0;
      for (unsigned p = 0; p < B; ++p) c |= ((planes[(1 + p) * ps + o + j] >> sh) & 1u) << p;
      x |= c << (8 * j);
    }
    codes[(size_t)s * groups + g] = x;
  }
}

// Gap counts of every replicate to zero (they are summed by atomics in the gather).
__global__ void protein_zero_gaps_kernel(uint32_t* __restrict__ planes, size_t rep_words,
                                         size_t gaps_off, unsigned n_pad) {
  const unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n_pad) planes[(size_t)blockIdx.y * rep_words + gaps_off + t] = 0u;
}

// Replicate r's planes from the expanded column list (target site s takes source column
// col_index[r][s]; sites at or beyond lengths[r] are code 0 without the gap bit). A thread takes
// 4 taxa and a range of 32-site words: per word, 32 u32 loads (one per site, the 4 taxa's code
// bytes), a bit transpose (word 8j + p: bit p of taxon j's codes), the B code planes and the gap
// plane (code 0 at a live site) stored 16 bytes at a time, and the taxa's gap sites summed; one
// atomic add per taxon at the end.
__global__ __launch_bounds__(64) void protein_gather_kernel(
    const uint32_t* __restrict__ codes, unsigned n_pad, unsigned B,
    const unsigned* __restrict__ col_index, unsigned idx_stride,
    const unsigned* __restrict__ lengths, unsigned W_pad, uint32_t* __restrict__ planes,
    size_t rep_words) {
  const unsigned groups = n_pad / 4;
  const unsigned g = blockIdx.x * blockDim.x + threadIdx.x;
  const unsigned r = blockIdx.z;
  if (g >= groups) return;
  const unsigned len = min(lengths[r], idx_stride);
  const unsigned* idx = col_index + (size_t)r * idx_stride;
  const size_t ps = (size_t)W_pad * n_pad;
  uint32_t* pr = planes + (size_t)r * rep_words;
  uint32_t gap_sum[4] = {0, 0, 0, 0};
  for (unsigned w = blockIdx.y; w < W_pad; w += gridDim.y) {
    uint32_t a[32];
#pragma unroll
    for (unsigned k = 0; k < 32; ++k) {
      const unsigned s = w * 32 + k;
      a[k] = s < len ? codes[(size_t)idx[s] * groups + g] : 0u;
    }
    transpose32(a);
    const unsigned live = w * 32 >= len ? 0u : (len - w * 32 >= 32 ? 32u : len - w * 32);
    const uint32_t live_mask = live == 32 ? ~0u : ((1u << live) - 1u);
    const size_t o = (size_t)w * n_pad + 4 * g;
    uint32_t gap[4];
#pragma unroll
    for (unsigned j = 0; j < 4; ++j) {
      uint32_t any = 0;
#pragma unroll
      for (unsigned p = 0; p < 8; ++p) any |= a[8 * j + p];
      gap[j] = ~any & live_mask;
      gap_sum[j] += __popc(gap[j]);
    }
    *reinterpret_cast<uint4*>(pr + o) = make_uint4(gap[0], gap[1], gap[2], gap[3]);
    for (unsigned p = 0; p < B; ++p)
      *reinterpret_cast<uint4*>(pr + (1 + p) * ps + o) =
          make_uint4(a[p], a[8 + p], a[16 + p], a[24 + p]);
  }
  uint32_t* gaps = pr + (1 + B) * ps;
#pragma unroll
  for (unsigned j = 0; j < 4; ++j)
    if (gap_sum[j]) atomicAdd(&gaps[4 * g + j], gap_sum[j]);
}

}  // namespace

// ---- launchers ---------------------------------------------------------------------------------
static unsigned ygrid_p(unsigned rows) { return rows < 1 ? 1 : (rows > 65535 ? 65535 : rows); }

hipError_t launch_symbol_census(const uint8_t* chars, size_t pitch, unsigned n, unsigned L,
                                uint32_t* present, hipStream_t stream) {
  const unsigned gx = std::min(32u, (L + 255) / 256);
  const unsigned gy = ygrid_p(std::min(n, 4096u));
  symbol_census_kernel<<<dim3(gx, gy), dim3(256), 0, stream>>>(chars, pitch, n, L, present);
  return hipGetLastError();
}

hipError_t launch_pack_protein(const uint8_t* chars, size_t pitch, unsigned n, unsigned L,
                               unsigned n_pad, unsigned W_pad, const uint8_t* codes, unsigned B,
                               uint32_t* planes, uint32_t* gaps, hipStream_t stream) {
  if (B != 5 && B != 6 && B != 8) return hipErrorInvalidValue;
  const unsigned wblocks = (W_pad + kPWords - 1) / kPWords;
  pack_protein_kernel<<<dim3((n_pad + kPTaxa - 1) / kPTaxa, ygrid_p(wblocks)), dim3(256), 0,
                        stream>>>(chars, pitch, n, L, n_pad, W_pad, codes, B, planes, gaps);
  return hipGetLastError();
}

hipError_t launch_protein_tiles(int config, unsigned B, bool exact, const JobDesc* jobs,
                                unsigned njobs, unsigned total_items, unsigned finish_items,
                                hipStream_t stream) {
  if (total_items == 0) return hipSuccess;
  switch (B) {
    case 5: return launch_protein_b<5>(config, exact, jobs, njobs, total_items, finish_items, stream);
    case 6: return launch_protein_b<6>(config, exact, jobs, njobs, total_items, finish_items, stream);
    case 8: return launch_protein_b<8>(config, exact, jobs, njobs, total_items, finish_items, stream);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace rnj

namespace rnj {

hipError_t launch_protein_codes(const uint32_t* planes, unsigned n_pad, unsigned W_pad,
                                unsigned L, unsigned B, uint8_t* codes, hipStream_t stream) {
  const unsigned groups = n_pad / 4;
  const unsigned ys = L < 1 ? 1 : (L > 65535 ? 65535 : L);
  protein_codes_kernel<<<dim3((groups + 255) / 256, ys), dim3(256), 0, stream>>>(
      planes, n_pad, W_pad, L, B, reinterpret_cast<uint32_t*>(codes));
  return hipGetLastError();
}

hipError_t launch_protein_gather(const uint8_t* codes, unsigned n_pad, unsigned B,
                                 const unsigned* col_index, unsigned idx_stride,
                                 const unsigned* lengths, unsigned R, unsigned W_pad,
                                 uint32_t* planes, size_t rep_words, hipStream_t stream) {
  if (R == 0) return hipSuccess;
  const size_t gaps_off = (size_t)(1 + B) * W_pad * n_pad;
  protein_zero_gaps_kernel<<<dim3((n_pad + 255) / 256, R), dim3(256), 0, stream>>>(
      planes, rep_words, gaps_off, n_pad);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const unsigned groups = n_pad / 4;
  const unsigned ys = W_pad < 16 ? W_pad : 16;  // word ranges per taxon group: few atomics
  protein_gather_kernel<<<dim3((groups + 63) / 64, ys, R), dim3(64), 0, stream>>>(
      reinterpret_cast<const uint32_t*>(codes), n_pad, B, col_index, idx_stride, lengths, W_pad,
      planes, rep_words);
  return hipGetLastError();
}

}  // namespace rnj
