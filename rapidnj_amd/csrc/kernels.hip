// HIP kernels for gfx950: alignment packing, the pair-tile count kernel with the fused JC / K2P
// correction, the saturation patch, and bootstrap column resampling.
//
// The count kernel is integer VALU work: per 32 sites and taxon pair it issues
//   g = Vi & Vj; u = bitop3(~(Hi ^ Hj) & g); s = bitop3((Li ^ Lj) & u);
//   valid += bcnt(g); same += bcnt(u); ts += bcnt(s)
// i.e. 6 VALU instructions (v_and, 2x v_bitop3, 3x v_bcnt with accumulate): u holds the valid
// sites whose bases share H (both purines or both pyrimidines), so tv = valid - same, and s the
// transitions among them. JC needs only ts + tv, which costs 5: g, u, the identical bases
// ~(Li ^ Lj) & u, and two counts (mismatches = valid - identical).
#include "device_common.hpp"

namespace rnj {

const TileConfig kTileConfigs[] = {
    {"t64_m4x4_c16", 64, 4, 4, 16},
    {"t64_m4x4_c8", 64, 4, 4, 8},
    {"t128_m8x4_c16", 128, 8, 4, 16},
    {"t32_m4x4_c16", 32, 4, 4, 16},
    {"t32_m4x4_c8", 32, 4, 4, 8},
};
const int kNumTileConfigs = sizeof(kTileConfigs) / sizeof(kTileConfigs[0]);

namespace {

// ---- epilogue shared by the tile kernel and the split-finish kernel ------------------------
// Thread (ty, tx) of tile (bi, bj) holds the counts of its MI x MJ pairs: a = tv (mismatches for
// the JC-only loop), b = ts, v = valid. Writes distances (or counts), records saturated pairs and
// folds the wave's maximum into the job's running maximum. A tile's pairs are those with i < j,
// or for a transposed tile of a kRows job (lower) those with i > j.
template <int TILE, int MI, int MJ, bool JC6, bool kExact>
__device__ __forceinline__ void tile_epilogue(const JobDesc& J, const LogEntry* log_tab,
                                              unsigned tile, unsigned bi,
                                              unsigned bj, bool lower, unsigned ty, unsigned tx,
                                              unsigned tid,
                                              const uint32_t (&acc_a)[MI][MJ],
                                              const uint32_t (&acc_b)[MI][MJ],
                                              const uint32_t (&acc_v)[MI][MJ]) {
  constexpr int NT = (TILE / MI) * (TILE / MJ);
  static_assert(NT % 64 == 0, "whole waves per work-group");
  const unsigned n = J.n;
  float lmax = 0.0f;
  uint32_t sat = 0;  // saturated pairs of the micro-tile, bit a * MJ + b
  if (J.model == kCounts) {
#pragma unroll
    for (int a = 0; a < MI; ++a) {
      const unsigned i = bi * TILE + ty * MI + a;
#pragma unroll
      for (int b = 0; b < MJ; ++b) {
        const unsigned j = bj * TILE + tx * MJ + b;
        // JC6 launches hold no counts jobs (those need ts and tv apart)
        if (j < n && i < j) finish_counts(J, i, j, acc_b[a][b], acc_a[a][b], acc_v[a][b]);
      }
    }
    return;
  }
  const LogConsts K = log_consts(log_tab);
  // Inner tiles of an upper-triangle output (bi < bj, the whole tile below n, as almost every
  // tile of a large job is): every pair has i < j < n, so each row's MJ corrections run as
  // straight-line code (no valid >= 2^24 branch: fewer than 2^24 sites were counted) that the
  // compiler interleaves, and the row's MJ entries, consecutive in the triangle, are stored from
  // one index. The rest take the general path below.
  const bool inner = !kExact && J.layout == kTriangle && !lower && bi < bj &&
                     (bj + 1) * TILE <= n && J.W_live < (1u << 19);
  if (inner) {
    const unsigned j0 = bj * TILE + tx * MJ;
    auto rows = [&](auto model) {
#pragma unroll
      for (int a = 0; a < MI; ++a) {
        const unsigned i = bi * TILE + ty * MI + a;
        float d[MJ];
#pragma unroll
        for (int b = 0; b < MJ; ++b) {
          if constexpr (decltype(model)::value == kJC)
            d[b] = jc_value<false, true>(JC6 ? acc_a[a][b] : acc_a[a][b] + acc_b[a][b],
                                         acc_v[a][b], K);
          else
            d[b] = k2p_value<false, true>(acc_b[a][b], acc_a[a][b], acc_v[a][b], K);
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
    if (JC6 || J.model == kJC) rows(std::integral_constant<int, kJC>());
    else rows(std::integral_constant<int, kK2P>());
    wave_record(J, tile * (NT / 64) + tid / 64, bi << 16 | bj, sat, lmax);
    return;
  }
  // General path: one row of MJ corrections, then the stores each pair's position allows.
#pragma unroll
  for (int a = 0; a < MI; ++a) {
    const unsigned i = bi * TILE + ty * MI + a;
    float d[MJ];
#pragma unroll
    for (int b = 0; b < MJ; ++b) {
      if (J.model == kJC)
        d[b] = jc_value<kExact, false>(JC6 ? acc_a[a][b] : acc_a[a][b] + acc_b[a][b], acc_v[a][b],
                                       K);
      else
        d[b] = k2p_value<kExact, false>(acc_b[a][b], acc_a[a][b], acc_v[a][b], K);
    }
#pragma unroll
    for (int b = 0; b < MJ; ++b) {
      const unsigned j = bj * TILE + tx * MJ + b;
      if (j >= n || i >= n) continue;
      if (lower ? i > j : i < j) {
        if (d[b] == -1.0f) sat |= 1u << (a * MJ + b);
        lmax = fmaxf(lmax, d[b]);  // a saturated pair's -1 leaves it as it is
        store_pair(J, i, j, d[b]);
      } else if (i == j) {
        finish_diagonal(J, i);
      }
    }
  }
  wave_record(J, tile * (NT / 64) + tid / 64, bi << 16 | bj, sat, lmax);
}

// ---- the pair-tile kernel ------------------------------------------------------------------
// Word w of the staged chunk into thread (ty, tx)'s MI x MJ accumulators (6 VALU ops per pair
// for K2P, 5 for JC, 9 for the naive rules).
template <int MI, int MJ, int CW, int Q, int MODE>
__device__ __forceinline__ void count_word(const v4u* __restrict__ sI, const v4u* __restrict__ sJ,
                                           int w, unsigned ty, unsigned tx,
                                           uint32_t (&acc_a)[MI][MJ], uint32_t (&acc_b)[MI][MJ],
                                           uint32_t (&acc_v)[MI][MJ]) {
  // Operands of plane p for the i side (ty's MI taxa) and the j side (tx's MJ taxa), read in the
  // order the counts use them: V first, then (naive rules) U, then H and L.
  uint32_t vi[MI], vj[MJ], hi[MI], hj[MJ], li[MI], lj[MJ];
  uint32_t ui[MODE == kModeNaive ? MI : 1], uj[MODE == kModeNaive ? MJ : 1];
  auto read = [&](int p, uint32_t (&oi)[MI], uint32_t (&oj)[MJ]) {
#pragma unroll
    for (int r = 0; r < MI / 4; ++r) {
      const v4u q = sI[(p * CW + w) * Q + ty * (MI / 4) + r];
      oi[4 * r] = q.x; oi[4 * r + 1] = q.y; oi[4 * r + 2] = q.z; oi[4 * r + 3] = q.w;
    }
#pragma unroll
    for (int r = 0; r < MJ / 4; ++r) {
      const v4u q = sJ[(p * CW + w) * Q + tx * (MJ / 4) + r];
      oj[4 * r] = q.x; oj[4 * r + 1] = q.y; oj[4 * r + 2] = q.z; oj[4 * r + 3] = q.w;
    }
  };
  read(2, vi, vj);
  if constexpr (MODE == kModeNaive) read(3, ui, uj);
  read(0, hi, hj);
  read(1, li, lj);
  if constexpr (MODE == kModeNaive) {
    // Naive rules (simpleDistanceCalculator.cpp:58-98): valid = both counted; ts = both
    // uppercase ACGT, same H, different L; tv = i uppercase ACGT, j counted, and j not uppercase
    // ACGT or a different H. Ordered asm in two passes over the micro-tile: the valid counts need
    // only the V planes, so they run while the other planes' reads are still in flight.
#pragma unroll
    for (int a = 0; a < MI; ++a)
#pragma unroll
      for (int b = 0; b < MJ; ++b) ordered_bcnt_add(ordered_and(vi[a], vj[b]), acc_v[a][b]);
#pragma unroll
    for (int a = 0; a < MI; ++a)
#pragma unroll
      for (int b = 0; b < MJ; ++b) {
        uint32_t t;
        asm volatile("v_xor_b32 %0, %1, %2" : "=v"(t) : "v"(li[a]), "v"(lj[b]));
        const uint32_t x = ordered_bitop3<0x80>(ui[a], uj[b], t);         // ui & uj & (li ^ lj)
        const uint32_t y = ordered_bitop3<0x7D>(hi[a], hj[b], uj[b]);     // (hi ^ hj) | ~uj
        ordered_bcnt_add(ordered_bitop3<0x82>(hi[a], hj[b], x), acc_b[a][b]);  // ~(hi ^ hj) & x
        ordered_bcnt_add(ordered_bitop3<0x80>(ui[a], vj[b], y), acc_a[a][b]);  // ui & vj & y
      }
  } else {
#pragma unroll
    for (int a = 0; a < MI; ++a)
#pragma unroll
      for (int b = 0; b < MJ; ++b) {
        // One pair's instructions back to back, in this order (ordered asm): the VALU runs the
        // mix of bitwise and v_bcnt instructions about 4 % faster this way than grouped by class
        // over the micro-tile, which is how the compiler schedules the same operations.
        // g: sites valid in both; u: those whose bases share H (purine/purine or
        // pyrimidine/pyrimidine), so #tv = #g - #u (counts_from_same turns acc_a into it);
        // K2P: x = the transitions among u; JC: x = the identical bases among u, and the
        // mismatches are #g - #x.
        const uint32_t g = ordered_and(vi[a], vj[b]);
        const uint32_t u = ordered_bitop3<0x82>(hi[a], hj[b], g);
        const uint32_t x = ordered_bitop3<MODE == kModeJC6 ? 0x82 : 0x28>(li[a], lj[b], u);
        ordered_bcnt_add(g, acc_v[a][b]);
        if constexpr (MODE == kModeJC6) {
          ordered_bcnt_add(x, acc_a[a][b]);
        } else {
          ordered_bcnt_add(u, acc_a[a][b]);
          ordered_bcnt_add(x, acc_b[a][b]);
        }
      }
  }
}

// After the word loop of the K2P and JC modes acc_a holds #u (K2P: same-class sites) or #same
// (JC: identical bases); both complements within the valid sites are what the epilogue takes
// (K2P: transversions, JC: mismatches). The naive mode counts them directly.
template <int MI, int MJ, int MODE>
__device__ __forceinline__ void counts_from_same(uint32_t (&acc_a)[MI][MJ],
                                                 const uint32_t (&acc_v)[MI][MJ]) {
  if constexpr (MODE != kModeNaive) {
#pragma unroll
    for (int a = 0; a < MI; ++a)
#pragma unroll
      for (int b = 0; b < MJ; ++b) acc_a[a][b] = acc_v[a][b] - acc_a[a][b];
  }
}

// Waves per SIMD the register allocation aims for: 6 for the JC count loop of 64-pair tiles with
// 8-word chunks (its count loop needs fewer than 84 VGPRs; left alone, the compiler interleaves
// the epilogue's corrections so far that the kernel needs 107 and runs 2 % slower at 4 waves),
// otherwise no constraint.
template <int TILE, int CW, int MODE, bool kExact>
constexpr int kTileWaves = (TILE == 64 && CW == 8 && MODE == kModeJC6 && !kExact) ? 6 : 1;

// One work-group computes a TILE x TILE block of taxon pairs (bi <= bj) over one range of words
// (the whole alignment, or split `split` of J.splits); thread (ty, tx) owns the MI x MJ
// micro-tile i = bi*TILE + ty*MI + a, j = bj*TILE + tx*MJ + b. Sites are walked in chunks of CW
// words; the chunk's H/L/V words of both taxon blocks are staged in LDS (layout
// [plane][word][taxon], written linearly) while the previous chunk is counted:
//  * work-groups of several waves load them straight from global memory into LDS
//    (global_load_lds_dwordx4, each wave filling 64 consecutive uint4) into two buffers in
//    turn, so no registers hold a chunk on the way (t64_m4x4_c8: K2P 96 VGPRs, 5 waves per
//    SIMD; JC 79, 6 waves);
//  * one-wave work-groups prefetch the next chunk into registers and write it to one buffer:
//    a second buffer would cost them occupancy, which LDS bounds for them.
// Unsplit jobs finish in place; split jobs leave exact partial counts for finish_kernel.
template <int TILE, int MI, int MJ, int CW, int MODE, bool kExact>
__global__ __launch_bounds__((TILE / MI) * (TILE / MJ))
__attribute__((amdgpu_waves_per_eu(kTileWaves<TILE, CW, MODE, kExact>))) void tile_kernel(
    const JobDesc* __restrict__ jobs, unsigned njobs) {
  constexpr int NTX = TILE / MJ;
  constexpr int NT = (TILE / MI) * NTX;
  constexpr int Q = TILE / 4;                 // uint4 per plane-word row of one block
  constexpr int P = MODE == kModeNaive ? 4 : 3;  // planes
  constexpr int E = P * CW * Q;               // uint4 per block and chunk
  static_assert(MI % 4 == 0 && MJ % 4 == 0, "micro-tiles are read as uint4");
  static_assert(NT % 64 == 0, "whole waves");
  constexpr int EP = (E + NT - 1) / NT;       // uint4 per thread and block (last may be partial)
  constexpr bool kDirect = NT >= 128;         // LDS loads, two buffers

  // Two staging buffers as distinct LDS variables: the LDS-DMA writes into one are then known
  // not to alias the reads of the other, so the compiler does not wait for the next chunk's
  // loads before counting the current one.
  __shared__ v4u sI0[E], sJ0[E];
  __shared__ v4u sI1[kDirect ? E : 1], sJ1[kDirect ? E : 1];
  static_assert(E * sizeof(v4u) >= 128 * sizeof(LogEntry), "the log table fits a staging buffer");

  const unsigned item = xcd_item(blockIdx.x, gridDim.x);
  const unsigned k = find_job(jobs, njobs, item);
  const JobDesc& J = jobs[k];
  const unsigned local = item - J.item_base;
  const unsigned tile = local / J.splits, split = local % J.splits;
  unsigned bi, bj;
  const bool lower = job_tile(J, tile, bi, bj);

  const unsigned tid = threadIdx.x;
  const unsigned ty = tid / NTX, tx = tid % NTX;
  const unsigned wave0 = tid & ~63u;  // the wave's first thread
  const gptr<v4u> gI = (gptr<v4u>)(J.planes + (size_t)bi * TILE);
  const gptr<v4u> gJ = (gptr<v4u>)(J.planes + (size_t)bj * TILE);
  const size_t row4 = J.n_pad / 4;                       // uint4 per word row
  const size_t plane4 = J.plane_stride / 4;

  uint32_t acc_a[MI][MJ], acc_b[MI][MJ], acc_v[MI][MJ];  // a: tv (K2P) / mismatches (JC6); b: ts
#pragma unroll
  for (int a = 0; a < MI; ++a)
#pragma unroll
    for (int b = 0; b < MJ; ++b) acc_a[a][b] = acc_b[a][b] = acc_v[a][b] = 0;

  // this range's words that hold sites: whole chunks, then at most one partial chunk (the
  // padded words past W_live are zero and would count nothing)
  const unsigned w0 = split * J.words_per_split;
  const unsigned w1 = min(J.W_live, w0 + J.words_per_split);
  const unsigned c0 = w0 / CW, c1 = (w1 + CW - 1) / CW;  // ranges start on chunk boundaries

  auto count_chunk = [&](const v4u* bI, const v4u* bJ, unsigned c) {
    // words in fours (the later words' reads take immediate offsets from the first's address);
    // `live` is uniform over the work-group, and readfirstlane lets the compiler keep the loop
    // counter in an SGPR
    const int live = __builtin_amdgcn_readfirstlane((int)min((unsigned)CW, w1 - c * CW));
    int w = 0;
#pragma unroll 1
    for (; w + 3 < live; w += 4) {
      count_word<MI, MJ, CW, Q, MODE>(bI, bJ, w, ty, tx, acc_a, acc_b, acc_v);
      count_word<MI, MJ, CW, Q, MODE>(bI, bJ, w + 1, ty, tx, acc_a, acc_b, acc_v);
      count_word<MI, MJ, CW, Q, MODE>(bI, bJ, w + 2, ty, tx, acc_a, acc_b, acc_v);
      count_word<MI, MJ, CW, Q, MODE>(bI, bJ, w + 3, ty, tx, acc_a, acc_b, acc_v);
    }
#pragma unroll 1
    for (; w < live; ++w) count_word<MI, MJ, CW, Q, MODE>(bI, bJ, w, ty, tx, acc_a, acc_b, acc_v);
  };

  if constexpr (kDirect) {
    // The next chunk of both blocks into buffers dI / dJ: element lin = q + Q*(w + CW*p) of
    // chunk c is word row (c*CW + w) of plane p, uint4 column q; lane l of a wave writes element
    // (base + l). Chunks are staged in order from c0, so each thread keeps pointers to its
    // elements of the next chunk and advances them by one chunk (CW word rows) per stage.
    gptr<v4u> nI[EP], nJ[EP];
#pragma unroll
    for (int e = 0; e < EP; ++e) {
      const unsigned lin = tid + e * NT;
      const unsigned q = lin % Q, w = (lin / Q) % CW, p = lin / (Q * CW);
      const size_t off = p * plane4 + (size_t)(c0 * CW + w) * row4 + q;
      nI[e] = gI + off;
      nJ[e] = gJ + off;
    }
    const size_t chunk4 = (size_t)CW * row4;
    auto stage = [&](v4u* dI, v4u* dJ) {
#pragma unroll
      for (int e = 0; e < EP; ++e) {
        const unsigned lin = tid + e * NT;
        if (E % NT != 0 && lin >= (unsigned)E) break;
        __builtin_amdgcn_global_load_lds(
            (const void*)nI[e], (__attribute__((address_space(3))) void*)&dI[wave0 + e * NT],
            16, 0, 0);
        __builtin_amdgcn_global_load_lds(
            (const void*)nJ[e], (__attribute__((address_space(3))) void*)&dJ[wave0 + e * NT],
            16, 0, 0);
        nI[e] += chunk4;
        nJ[e] += chunk4;
      }
    };
    if constexpr (MODE == kModeNaive) {
      // The naive rules' loop (four planes, 8 ds_read_b128 per word) alternates the buffers by
      // chunk parity, so the compiler cannot tell them apart and waits for the next chunk's loads
      // before the first read of each chunk: measured 7 % faster at 20000 x 5000 than counting
      // while those loads land.
      if (c0 < c1) stage(sI0, sJ0);
      for (unsigned c = c0; c < c1; ++c) {
        const bool odd = ((c - c0) & 1u) != 0;
        __builtin_amdgcn_s_waitcnt(0x0f70);  // vmcnt(0): this wave's part of chunk c is in LDS
        __syncthreads();  // every wave's part too, and every wave is done with the other buffer
        if (c + 1 < c1) stage(odd ? sI0 : sI1, odd ? sJ0 : sJ1);
        count_chunk(odd ? sI1 : sI0, odd ? sJ1 : sJ0, c);
      }
    } else {
      // chunks in pairs, so that each buffer is a compile-time choice and the next chunk's
      // loads land while this one is counted
      if (c0 < c1) stage(sI0, sJ0);
      for (unsigned c = c0; c < c1; c += 2) {
        __builtin_amdgcn_s_waitcnt(0x0f70);  // vmcnt(0): this wave's part of chunk c is in LDS
        __syncthreads();  // every wave's part too, and every wave is done with buffer 1
        if (c + 1 < c1) stage(sI1, sJ1);
        count_chunk(sI0, sJ0, c);
        if (c + 1 >= c1) break;
        __builtin_amdgcn_s_waitcnt(0x0f70);
        __syncthreads();
        if (c + 2 < c1) stage(sI0, sJ0);
        count_chunk(sI1, sJ1, c + 1);
      }
    }
  } else {
    v4u pI[EP], pJ[EP];
    if (c0 < c1) fetch_chunk<EP, NT, Q, CW, E>(pI, pJ, gI, gJ, tid, c0, row4, plane4);
    for (unsigned c = c0; c < c1; ++c) {
      __syncthreads();
#pragma unroll
      for (int e = 0; e < EP; ++e) {
        if (E % NT != 0 && tid + e * NT >= (unsigned)E) break;
        sI0[tid + e * NT] = pI[e];
        sJ0[tid + e * NT] = pJ[e];
      }
      __syncthreads();
      if (c + 1 < c1) fetch_chunk<EP, NT, Q, CW, E>(pI, pJ, gI, gJ, tid, c + 1, row4, plane4);
      count_chunk(sI0, sJ0, c);
    }
  }
  counts_from_same<MI, MJ, MODE>(acc_a, acc_v);

  if (J.splits > 1) {
    uint32_t* slot = partial_slot<NT, MI, MJ>(J, tile, split) + (size_t)tid * MI * MJ;
    const size_t plane = (size_t)NT * MI * MJ;
#pragma unroll
    for (int a = 0; a < MI; ++a)
#pragma unroll
      for (int b = 0; b < MJ; ++b) {
        slot[a * MJ + b] = acc_a[a][b];
        slot[plane + a * MJ + b] = acc_b[a][b];
        slot[2 * plane + a * MJ + b] = acc_v[a][b];
      }
    return;
  }
  // The log table goes into the first staging buffer once every wave is done counting, so that
  // it costs the work-group no LDS of its own (at 2 KB it would cost the naive rules' t64_m4x4_c8
  // work-groups their fifth wave per SIMD).
  LogEntry* const sLog = reinterpret_cast<LogEntry*>(sI0);
  __syncthreads();
  stage_log_table(sLog, tid, NT);
  __syncthreads();
  tile_epilogue<TILE, MI, MJ, MODE == kModeJC6, kExact>(J, sLog, tile, bi, bj, lower, ty, tx, tid, acc_a,
                                                        acc_b, acc_v);
}

// Sums the partial counts of the splits of one tile of a split job and finishes the tile. Work
// items are numbered per job from fin_base over the fin_tiles (= tiles when split, else 0).
template <int TILE, int MI, int MJ, bool JC6, bool kExact>
__global__ __launch_bounds__((TILE / MI) * (TILE / MJ)) void finish_kernel(
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
  __shared__ LogEntry sLog[128];
  stage_log_table(sLog, threadIdx.x, NT);
  __syncthreads();
  unsigned bi, bj;
  const bool lower = job_tile(J, tile, bi, bj);
  const unsigned tid = threadIdx.x;
  const unsigned ty = tid / NTX, tx = tid % NTX;
  uint32_t acc_a[MI][MJ], acc_b[MI][MJ], acc_v[MI][MJ];
#pragma unroll
  for (int a = 0; a < MI; ++a)
#pragma unroll
    for (int b = 0; b < MJ; ++b) acc_a[a][b] = acc_b[a][b] = acc_v[a][b] = 0;
  const size_t plane = (size_t)NT * MI * MJ;
  for (unsigned sp = 0; sp < J.splits; ++sp) {
    const uint32_t* slot = partial_slot<NT, MI, MJ>(J, tile, sp) + (size_t)tid * MI * MJ;
#pragma unroll
    for (int a = 0; a < MI; ++a)
#pragma unroll
      for (int b = 0; b < MJ; ++b) {
        acc_a[a][b] += slot[a * MJ + b];
        acc_b[a][b] += slot[plane + a * MJ + b];
        acc_v[a][b] += slot[2 * plane + a * MJ + b];
      }
  }
  tile_epilogue<TILE, MI, MJ, JC6, kExact>(J, sLog, tile, bi, bj, lower, ty, tx, tid, acc_a, acc_b,
                                           acc_v);
}

template <int TILE, int MI, int MJ, int CW>
hipError_t launch_cfg(int mode, bool exact, const JobDesc* jobs, unsigned njobs, unsigned total,
                      unsigned finish_items, hipStream_t s) {
  constexpr int NT = (TILE / MI) * (TILE / MJ);
  const dim3 grid(total), block(NT);
#define RNJ_TILES(M, X) tile_kernel<TILE, MI, MJ, CW, M, X><<<grid, block, 0, s>>>(jobs, njobs)
  if (mode == kModeJC6) {
    if (exact) RNJ_TILES(kModeJC6, true); else RNJ_TILES(kModeJC6, false);
  } else if (mode == kModeNaive) {
    if (exact) RNJ_TILES(kModeNaive, true); else RNJ_TILES(kModeNaive, false);
  } else {
    if (exact) RNJ_TILES(kModeK2P, true); else RNJ_TILES(kModeK2P, false);
  }
#undef RNJ_TILES
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || finish_items == 0) return e;
  const dim3 fgrid(finish_items);
  const bool jc6 = mode == kModeJC6;
  if (jc6 && exact) finish_kernel<TILE, MI, MJ, true, true><<<fgrid, block, 0, s>>>(jobs, njobs);
  else if (jc6) finish_kernel<TILE, MI, MJ, true, false><<<fgrid, block, 0, s>>>(jobs, njobs);
  else if (exact) finish_kernel<TILE, MI, MJ, false, true><<<fgrid, block, 0, s>>>(jobs, njobs);
  else finish_kernel<TILE, MI, MJ, false, false><<<fgrid, block, 0, s>>>(jobs, njobs);
  return hipGetLastError();
}

// ---- packing ---------------------------------------------------------------------------------
// Bits of one character: H (pyrimidine), L (G or C), V (counted base). Only ACGT of either case
// are counted (dataloaderFasta.hpp:147-182 sets the gap-filter slot for exactly these); c | 0x20
// folds case and maps no other byte onto a, c, g or t.
__device__ __forceinline__ void char_bits(uint8_t c, uint32_t& h, uint32_t& l, uint32_t& v) {
  const uint8_t lc = c | 0x20;
  v = (lc == 'a') | (lc == 'c') | (lc == 'g') | (lc == 't');
  h = (lc == 'c') | (lc == 't');
  l = (lc == 'g') | (lc == 'c');
}

// One work-group packs a tile of 64 taxa x 1024 sites (32 words): the tile's bytes are read row
// by row (consecutive lanes, consecutive bytes: coalesced) into LDS rows padded by one word, then
// lane t of every wave packs taxon t's words from LDS (rows one word apart, so the 64 lanes hit
// 64 different banks) and stores them (consecutive lanes, consecutive taxa: coalesced). Sites at
// or past L and taxa at or past n pack as 0 (not counted).
constexpr unsigned kPackTaxa = 64, kPackWords = 32, kPackRow = kPackWords * 8 + 1;

// Bits of one character under the naive counter's rules: the reference's loader first turns
// every byte but ACGTU of either case into '-' (dataloaderFasta.hpp:184-191), and
// simpleDistanceCalculator::calculateTsTv (simpleDistanceCalculator.cpp:58-98) counts a site
// when neither is '-' (V), and classifies a mismatch only when the row's character is uppercase
// A, C, G or T (U plane), by exact byte: H and L as char_bits for those four.
__device__ __forceinline__ void naive_bits(uint8_t c, uint32_t& h, uint32_t& l, uint32_t& v,
                                           uint32_t& u) {
  const uint8_t lc = c | 0x20;
  v = (lc == 'a') | (lc == 'c') | (lc == 'g') | (lc == 't') | (lc == 'u');
  u = (c == 'A') | (c == 'C') | (c == 'G') | (c == 'T');
  h = (c == 'C') | (c == 'T');
  l = (c == 'G') | (c == 'C');
}

template <bool NAIVE>
__global__ __launch_bounds__(256) void pack_chars_kernel(const uint8_t* __restrict__ chars,
                                                         size_t pitch, unsigned n, unsigned L,
                                                         unsigned n_pad, unsigned W_pad,
                                                         uint32_t* __restrict__ planes) {
  __shared__ uint32_t tile[kPackTaxa * kPackRow];
  const unsigned t0 = blockIdx.x * kPackTaxa;
  const unsigned lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const size_t ps = (size_t)W_pad * n_pad;
  const unsigned t = t0 + lane;
  for (unsigned w0 = blockIdx.y * kPackWords; w0 < W_pad; w0 += gridDim.y * kPackWords) {
    for (unsigned e = threadIdx.x; e < kPackTaxa * kPackWords * 8; e += blockDim.x) {
      const unsigned r = e / (kPackWords * 8), q = e % (kPackWords * 8);
      const size_t s = (size_t)w0 * 32 + 4 * q;
      uint32_t x = 0;
      if (t0 + r < n) {
        const uint8_t* row = chars + (size_t)(t0 + r) * pitch;
#pragma unroll
        for (unsigned k = 0; k < 4; ++k)
          if (s + k < L) x |= (uint32_t)row[s + k] << (8 * k);
      }
      tile[r * kPackRow + q] = x;
    }
    __syncthreads();
    for (unsigned ww = wave; ww < kPackWords; ww += blockDim.x / 64) {
      const unsigned w = w0 + ww;
      if (w >= W_pad) break;
      uint32_t H = 0, Lo = 0, V = 0, U = 0;
#pragma unroll
      for (unsigned j = 0; j < 8; ++j) {
        const uint32_t x = tile[lane * kPackRow + ww * 8 + j];
#pragma unroll
        for (unsigned k = 0; k < 4; ++k) {
          uint32_t h, l, v, u = 0;
          if constexpr (NAIVE) naive_bits((uint8_t)(x >> (8 * k)), h, l, v, u);
          else char_bits((uint8_t)(x >> (8 * k)), h, l, v);
          H |= h << (4 * j + k);
          Lo |= l << (4 * j + k);
          V |= v << (4 * j + k);
          U |= u << (4 * j + k);
        }
      }
      if (t < n_pad) {
        planes[0 * ps + (size_t)w * n_pad + t] = H;
        planes[1 * ps + (size_t)w * n_pad + t] = Lo;
        planes[2 * ps + (size_t)w * n_pad + t] = V;
        if constexpr (NAIVE) planes[3 * ps + (size_t)w * n_pad + t] = U;
      }
    }
    __syncthreads();  // the tile is rewritten by the next word block
  }
}

// Bits 0, 2, 4, .., 30 of x gathered into bits 0..15.
__device__ __forceinline__ uint32_t even_bits(uint32_t x) {
  x &= 0x55555555u;
  x = (x | (x >> 1)) & 0x33333333u;
  x = (x | (x >> 2)) & 0x0f0f0f0fu;
  x = (x | (x >> 4)) & 0x00ff00ffu;
  x = (x | (x >> 8)) & 0x0000ffffu;
  return x;
}

// From the reference's packed layout: per taxon `stride` u32 words of 16 sites, site s at bits
// 2*(s%16) of word s/16 holding the code A=00 G=01 T=10 C=11, and a gap filter with 01 in the
// slot of every counted site (dataloaderFasta.hpp:219-251, dataloader.hpp:38-70).
__global__ void pack_legacy_kernel(const uint32_t* __restrict__ bits,
                                   const uint32_t* __restrict__ gaps, size_t stride,
                                   unsigned words16, unsigned n, unsigned n_pad, unsigned W_pad,
                                   uint32_t* __restrict__ planes) {
  const unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_pad) return;
  for (unsigned w = blockIdx.y; w < W_pad; w += gridDim.y) {
  uint32_t b0 = 0, b1 = 0, g0 = 0, g1 = 0;
  if (t < n) {
    const size_t base = (size_t)t * stride;
    if (2 * w < words16) { b0 = bits[base + 2 * w]; g0 = gaps[base + 2 * w]; }
    if (2 * w + 1 < words16) { b1 = bits[base + 2 * w + 1]; g1 = gaps[base + 2 * w + 1]; }
  }
  const size_t ps = (size_t)W_pad * n_pad;
  planes[0 * ps + (size_t)w * n_pad + t] = even_bits(b0 >> 1) | (even_bits(b1 >> 1) << 16);
  planes[1 * ps + (size_t)w * n_pad + t] = even_bits(b0) | (even_bits(b1) << 16);
  planes[2 * ps + (size_t)w * n_pad + t] = even_bits(g0) | (even_bits(g1) << 16);
  }
}

// One job description, passed by value, into device memory (a launch whose job table lives in
// memory the caller brings, with no host-to-device copy).
__global__ void store_job_kernel(JobDesc d, JobDesc* __restrict__ dst) {
  if (threadIdx.x == 0) *dst = d;
}

__global__ void copy_max_kernel(const unsigned* __restrict__ maxbits, float* __restrict__ dst) {
  if (threadIdx.x == 0) *dst = __uint_as_float(*maxbits);
}

__global__ void reset_stats_kernel(const JobDesc* __restrict__ jobs, unsigned njobs) {
  const unsigned k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= njobs) return;
  *jobs[k].maxbits = 0u;
}

// The largest distance of each job (grid y = job): the job's wave words folded per work-group
// (grid-stride over its units), then one atomicMax per work-group into the job's maxbits, which
// reset_stats_kernel zeroed. waves: waves per tile-kernel work-group.
__global__ __launch_bounds__(256) void job_max_kernel(const JobDesc* __restrict__ jobs,
                                                      unsigned waves) {
  const JobDesc& J = jobs[blockIdx.y];
  if (J.model == kCounts) return;  // uniform over the work-group
  const unsigned long long units = (unsigned long long)J.tiles * waves;
  uint32_t m = 0;
  for (unsigned long long u = (unsigned long long)blockIdx.x * 256 + threadIdx.x; u < units;
       u += (unsigned long long)gridDim.x * 256)
    m = max(m, J.sat_word[u] & ~kSatFlag);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) m = max(m, (uint32_t)__shfl_xor(m, off));
  __shared__ uint32_t wm[4];
  if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    m = max(max(wm[0], wm[1]), max(wm[2], wm[3]));
    if (m != 0) atomicMax(J.maxbits, m);
  }
}

// The reference post-process (KimuraDistance.cpp:366-395, JCdistance.cpp:201-230) replaces every
// -1 by 2 * maxDistance. Each wave takes kPatchUnits consecutive wave units of the job (grid y)
// and issues all their loads at once (lane k < kPatchUnits: unit k's word and block
// coordinates; every lane: its mask of each unit), so that one memory latency covers them. For a
// flagged unit, lane l's mask is the one lane l of the epilogue's wave stored; its bits map back
// to pairs with the unit's block coordinates and the launch's tile geometry (tile edge,
// micro-tile mi x mj).
constexpr unsigned kPatchUnits = 8;

__global__ __launch_bounds__(256) void patch_kernel(const JobDesc* __restrict__ jobs,
                                                    unsigned edge, unsigned mi, unsigned mj) {
  const JobDesc& J = jobs[blockIdx.y];
  if (J.model == kCounts) return;
  const unsigned ntx = edge / mj;
  const unsigned waves = (edge / mi) * ntx / 64;
  const unsigned long long units = (unsigned long long)J.tiles * waves;
  const unsigned long long u0 =
      ((unsigned long long)blockIdx.x * 4 + (threadIdx.x >> 6)) * kPatchUnits;
  if (u0 >= units) return;  // uniform over the wave
  const unsigned lane = threadIdx.x & 63;
  uint32_t word = 0, bij = 0;
  if (lane < kPatchUnits && u0 + lane < units) {
    word = J.sat_word[u0 + lane];
    bij = J.sat_tile[u0 + lane];
  }
  uint32_t mask[kPatchUnits];
#pragma unroll
  for (unsigned k = 0; k < kPatchUnits; ++k)
    mask[k] = u0 + k < units ? J.sat_mask[(u0 + k) * 64 + lane] : 0u;
  const float val = __uint_as_float(*J.maxbits) * 2.0f;
  const unsigned long long flagged = __ballot((word & kSatFlag) != 0u);
#pragma unroll
  for (unsigned k = 0; k < kPatchUnits; ++k) {
    const unsigned b = (unsigned)__shfl((int)bij, (int)k);
    if (((flagged >> k) & 1ull) == 0ull) continue;  // uniform over the wave
    uint32_t m = mask[k];
    const unsigned tid = (unsigned)((u0 + k) % waves) * 64 + lane;
    const unsigned i0 = (b >> 16) * edge + (tid / ntx) * mi;
    const unsigned j0 = (b & 0xffffu) * edge + (tid % ntx) * mj;
    while (m) {
      const unsigned bit = (unsigned)(__ffs(m) - 1);
      m &= m - 1;
      store_pair(J, i0 + bit / mj, j0 + bit % mj, val);
    }
  }
}

// Per-site codes of every taxon for the column gathers: nibbles, two taxa per byte, site-major
// (codes[s][t / 2], taxon t in the low nibble when even), nibble = U << 3 | V << 2 | H << 1 | L
// (U, the naive rules' fourth plane, only when naive). One thread per (site, 8 taxa): one u32 of
// output.
template <bool NAIVE>
__global__ void planes_to_codes_kernel(const uint32_t* __restrict__ planes, unsigned n_pad,
                                       unsigned W_pad, unsigned L, uint32_t* __restrict__ codes) {
  const unsigned g = blockIdx.x * blockDim.x + threadIdx.x;  // group of 8 taxa
  const unsigned groups = n_pad / 8;
  if (g >= groups) return;
  const size_t ps = (size_t)W_pad * n_pad;
  for (unsigned s = blockIdx.y; s < L; s += gridDim.y) {
    const size_t o = (size_t)(s / 32) * n_pad + 8 * g;
    const unsigned sh = s % 32;
    uint32_t x = 0;
#pragma unroll
    for (unsigned j = 0; j < 8; ++j) {
      const uint32_t h = (planes[o + j] >> sh) & 1u;
      const uint32_t l = (planes[ps + o + j] >> sh) & 1u;
      const uint32_t v = (planes[2 * ps + o + j] >> sh) & 1u;
      const uint32_t u = NAIVE ? (planes[3 * ps + o + j] >> sh) & 1u : 0u;
      x |= (u << 3 | v << 2 | h << 1 | l) << (4 * j);
    }
    codes[(size_t)s * groups + g] = x;
  }
}

// One work-group per replicate: exclusive prefix sum of the column weights, then column c is
// written weights[c] times, in column order (the expanded alignment's column list).
__global__ __launch_bounds__(1024) void bootstrap_index_kernel(
    const uint16_t* __restrict__ weights, unsigned L, unsigned* __restrict__ col_index,
    unsigned idx_stride, unsigned* __restrict__ lengths) {
  __shared__ unsigned wave_sum[16];
  __shared__ unsigned carry_s;
  const unsigned r = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const uint16_t* wr = weights + (size_t)r * L;
  unsigned* out = col_index + (size_t)r * idx_stride;
  if (tid == 0) carry_s = 0;
  __syncthreads();
  for (unsigned base = 0; base < L; base += 1024) {
    const unsigned c = base + tid;
    const unsigned w = c < L ? wr[c] : 0u;
    unsigned incl = w;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const unsigned y = __shfl_up(incl, off);
      if (lane >= (unsigned)off) incl += y;
    }
    if (lane == 63) wave_sum[wid] = incl;
    __syncthreads();
    unsigned before = carry_s;
    for (unsigned q = 0; q < wid; ++q) before += wave_sum[q];
    const unsigned excl = before + incl - w;
    for (unsigned m = 0; m < w; ++m)
      if (excl + m < idx_stride) out[excl + m] = c;
    __syncthreads();
    if (tid == 0) {
      unsigned tot = 0;
      for (unsigned q = 0; q < 16; ++q) tot += wave_sum[q];
      carry_s += tot;
    }
    __syncthreads();
  }
  if (tid == 0) lengths[r] = carry_s;
}


// Replicate r's packed planes from the expanded column list: target site s takes source column
// col_index[r][s]. Sites at or beyond lengths[r] are zero (not counted). A thread takes 8 taxa
// and one 32-site word: 32 u32 loads (one per site, the 8 taxa's nibbles), a bit transpose that
// turns them into 32 words of site bits (word 4j + p: taxon j's plane p), and 6 16-byte stores
// (8 with the naive rules' U plane, word 4j + 3).
template <bool NAIVE>
__global__ __launch_bounds__(64) void bootstrap_gather_kernel(
    const uint32_t* __restrict__ codes, unsigned n_pad, const unsigned* __restrict__ col_index,
    unsigned idx_stride, const unsigned* __restrict__ lengths, unsigned W_pad,
    uint32_t* __restrict__ planes) {
  const unsigned groups = n_pad / 8;
  const unsigned g = blockIdx.x * blockDim.x + threadIdx.x;
  const unsigned r = blockIdx.z;
  if (g >= groups) return;
  const unsigned len = min(lengths[r], idx_stride);
  const unsigned* idx = col_index + (size_t)r * idx_stride;
  const size_t ps = (size_t)W_pad * n_pad;
  uint32_t* pr = planes + (size_t)r * (NAIVE ? 4 : 3) * ps;
  for (unsigned w = blockIdx.y; w < W_pad; w += gridDim.y) {
    uint32_t a[32];
#pragma unroll
    for (unsigned k = 0; k < 32; ++k) {
      const unsigned s = w * 32 + k;
      a[k] = s < len ? codes[(size_t)idx[s] * groups + g] : 0u;
    }
    transpose32(a);
    // a[4j + p] = taxon (8g + j), plane p (0: L, 1: H, 2: V) over the word's 32 sites
    const size_t o = (size_t)w * n_pad + 8 * g;
    uint4* h = reinterpret_cast<uint4*>(pr + o);
    uint4* l = reinterpret_cast<uint4*>(pr + ps + o);
    uint4* v = reinterpret_cast<uint4*>(pr + 2 * ps + o);
    h[0] = make_uint4(a[1], a[5], a[9], a[13]);
    h[1] = make_uint4(a[17], a[21], a[25], a[29]);
    l[0] = make_uint4(a[0], a[4], a[8], a[12]);
    l[1] = make_uint4(a[16], a[20], a[24], a[28]);
    v[0] = make_uint4(a[2], a[6], a[10], a[14]);
    v[1] = make_uint4(a[18], a[22], a[26], a[30]);
    if constexpr (NAIVE) {
      uint4* u = reinterpret_cast<uint4*>(pr + 3 * ps + o);
      u[0] = make_uint4(a[3], a[7], a[11], a[15]);
      u[1] = make_uint4(a[19], a[23], a[27], a[31]);
    }
  }
}

// ---- issue-rate probe ----------------------------------------------------------------------
// 16 independent dependency chains per lane of one instruction class (OP 0: v_bitop3_b32,
// OP 1: v_bcnt_u32_b32), 32 instructions per loop iteration; the chains' ends are stored so
// that nothing is dead.
template <int OP>
__global__ __launch_bounds__(256) void issue_probe_kernel(uint32_t* __restrict__ out, int iters,
                                                          uint32_t seed) {
  constexpr int K = 16;
  uint32_t a[K];
  const uint32_t b = seed * 2654435761u + threadIdx.x, c = seed ^ 0x5bd1e995u;
#pragma unroll
  for (int k = 0; k < K; ++k) a[k] = seed + threadIdx.x * K + k;
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
      for (int k = 0; k < K; ++k) {
        if constexpr (OP == 0)
          asm volatile("v_bitop3_b32 %0, %0, %1, %2 bitop3:0x28" : "+v"(a[k]) : "v"(b), "v"(c));
        else
          asm volatile("v_bcnt_u32_b32 %0, %1, %0" : "+v"(a[k]) : "v"(b));
      }
  }
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < K; ++k) s += a[k];
  out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

// ---- instruction-mix probe ------------------------------------------------------------------
// The count loop's per-pair instruction sequence, executed back to back on 2 x 2 pairs per
// thread whose operands stay in registers (16 pairs per loop iteration): K2P and JC as count_word
// issues them; the naive rules (v_and, v_xor, 4 v_bitop3, 3 v_bcnt) and protein with B code
// planes (v_xor, B - 1 v_bitop3, v_and, 2 v_bcnt) in their source order. Volatile asm keeps every
// instruction in the loop although its operands do not change. About 30 VGPRs, so 8 waves per
// SIMD: the rate at which the VALU executes that sequence with all the latency hiding the SIMD
// can have, and no LDS reads, staging, barriers or epilogue.
template <int MIX, int B>
__global__ __launch_bounds__(256) void mix_probe_kernel(uint32_t* __restrict__ out, int iters,
                                                        uint32_t seed) {
  constexpr int P = MIX == kMixProtein ? 1 + B : 4;  // operand planes
  uint32_t ci[P][2], cj[P][2];
#pragma unroll
  for (int p = 0; p < P; ++p)
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      ci[p][k] = (seed + threadIdx.x) * (2654435761u + 2u * p + k);
      cj[p][k] = (seed ^ threadIdx.x) * (40503u + 7u * p + 3u * k);
    }
  uint32_t acc[3][2][2] = {};
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int rep = 0; rep < 4; ++rep)
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
          if constexpr (MIX == kMixK2P || MIX == kMixJC) {
            const uint32_t g = ordered_and(ci[2][a], cj[2][b]);
            const uint32_t u = ordered_bitop3<0x82>(ci[0][a], cj[0][b], g);
            const uint32_t x = ordered_bitop3<MIX == kMixJC ? 0x82 : 0x28>(ci[1][a], cj[1][b], u);
            ordered_bcnt_add(g, acc[0][a][b]);
            if constexpr (MIX == kMixJC) {
              ordered_bcnt_add(x, acc[1][a][b]);
            } else {
              ordered_bcnt_add(u, acc[1][a][b]);
              ordered_bcnt_add(x, acc[2][a][b]);
            }
          } else if constexpr (MIX == kMixNaive) {
            uint32_t t;
            const uint32_t g = ordered_and(ci[2][a], cj[2][b]);
            ordered_bcnt_add(g, acc[0][a][b]);
            asm volatile("v_xor_b32 %0, %1, %2" : "=v"(t) : "v"(ci[1][a]), "v"(cj[1][b]));
            const uint32_t x = ordered_bitop3<0x80>(ci[3][a], cj[3][b], t);
            const uint32_t y = ordered_bitop3<0x7D>(ci[0][a], cj[0][b], cj[3][b]);
            ordered_bcnt_add(ordered_bitop3<0x82>(ci[0][a], cj[0][b], x), acc[1][a][b]);
            ordered_bcnt_add(ordered_bitop3<0x80>(ci[3][a], cj[2][b], y), acc[2][a][b]);
          } else {
            uint32_t x;
            asm volatile("v_xor_b32 %0, %1, %2" : "=v"(x) : "v"(ci[1][a]), "v"(cj[1][b]));
#pragma unroll
            for (int p = 2; p < P; ++p) x = ordered_bitop3<0xF6>(x, ci[p][a], cj[p][b]);
            ordered_bcnt_add(x, acc[0][a][b]);
            ordered_bcnt_add(ordered_and(ci[0][a], cj[0][b]), acc[1][a][b]);
          }
        }
  }
  uint32_t s = 0;
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) s += acc[c][a][b];
  out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

}  // namespace

// VALU instructions per pair of one mix: K2P 6, JC 5, naive 9, protein B + 3.
int mix_ops_per_pair(int mix, int planes) {
  switch (mix) {
    case kMixK2P: return 6;
    case kMixJC: return 5;
    case kMixNaive: return 9;
    default: return planes + 3;
  }
}

hipError_t launch_mix_probe(int mix, int planes, uint32_t* out, unsigned blocks, int iters,
                            hipStream_t s) {
  const dim3 g(blocks), b(256);
  switch (mix) {
    case kMixK2P: mix_probe_kernel<kMixK2P, 0><<<g, b, 0, s>>>(out, iters, 1u); break;
    case kMixJC: mix_probe_kernel<kMixJC, 0><<<g, b, 0, s>>>(out, iters, 1u); break;
    case kMixNaive: mix_probe_kernel<kMixNaive, 0><<<g, b, 0, s>>>(out, iters, 1u); break;
    case kMixProtein:
      if (planes == 5) mix_probe_kernel<kMixProtein, 5><<<g, b, 0, s>>>(out, iters, 1u);
      else if (planes == 6) mix_probe_kernel<kMixProtein, 6><<<g, b, 0, s>>>(out, iters, 1u);
      else if (planes == 8) mix_probe_kernel<kMixProtein, 8><<<g, b, 0, s>>>(out, iters, 1u);
      else return hipErrorInvalidValue;
      break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_issue_probe(int op, uint32_t* out, unsigned blocks, int iters, hipStream_t s) {
  if (op == 0) issue_probe_kernel<0><<<dim3(blocks), dim3(256), 0, s>>>(out, iters, 1u);
  else issue_probe_kernel<1><<<dim3(blocks), dim3(256), 0, s>>>(out, iters, 1u);
  return hipGetLastError();
}

// ---- launchers ---------------------------------------------------------------------------------
// grid.y is limited to 65535; the packing kernels loop over the rest.
static unsigned ygrid(unsigned rows) { return rows < 1 ? 1 : (rows > 65535 ? 65535 : rows); }

hipError_t launch_pack_chars(const uint8_t* chars, size_t pitch, unsigned n, unsigned L,
                             unsigned n_pad, unsigned W_pad, uint32_t* planes,
                             hipStream_t stream, bool naive) {
  const unsigned wblocks = (W_pad + kPackWords - 1) / kPackWords;
  const dim3 grid((n_pad + kPackTaxa - 1) / kPackTaxa, ygrid(wblocks)), block(256);
  if (naive)
    pack_chars_kernel<true><<<grid, block, 0, stream>>>(chars, pitch, n, L, n_pad, W_pad, planes);
  else
    pack_chars_kernel<false><<<grid, block, 0, stream>>>(chars, pitch, n, L, n_pad, W_pad, planes);
  return hipGetLastError();
}

hipError_t launch_pack_legacy(const uint32_t* bits, const uint32_t* gaps, size_t stride,
                              unsigned words16, unsigned n, unsigned n_pad, unsigned W_pad,
                              uint32_t* planes, hipStream_t stream) {
  pack_legacy_kernel<<<dim3(n_pad / 256 + (n_pad % 256 != 0), ygrid(W_pad)), dim3(256), 0, stream>>>(
      bits, gaps, stride, words16, n, n_pad, W_pad, planes);
  return hipGetLastError();
}

hipError_t launch_store_job(const JobDesc& d, JobDesc* dst, hipStream_t stream) {
  store_job_kernel<<<dim3(1), dim3(64), 0, stream>>>(d, dst);
  return hipGetLastError();
}

hipError_t launch_copy_max(const unsigned* maxbits, float* dst, hipStream_t stream) {
  copy_max_kernel<<<dim3(1), dim3(64), 0, stream>>>(maxbits, dst);
  return hipGetLastError();
}

hipError_t launch_reset_stats(const JobDesc* jobs, unsigned njobs, hipStream_t stream) {
  reset_stats_kernel<<<dim3((njobs + 255) / 256), dim3(256), 0, stream>>>(jobs, njobs);
  return hipGetLastError();
}

hipError_t launch_tiles(int config, int mode, bool exact, const JobDesc* jobs, unsigned njobs,
                        unsigned total_items, unsigned finish_items, hipStream_t stream) {
  if (total_items == 0) return hipSuccess;
  const unsigned f = finish_items;
  switch (config) {
    case 0: return launch_cfg<64, 4, 4, 16>(mode, exact, jobs, njobs, total_items, f, stream);
    case 1: return launch_cfg<64, 4, 4, 8>(mode, exact, jobs, njobs, total_items, f, stream);
    case 2: return launch_cfg<128, 8, 4, 16>(mode, exact, jobs, njobs, total_items, f, stream);
    case 3: return launch_cfg<32, 4, 4, 16>(mode, exact, jobs, njobs, total_items, f, stream);
    case 4: return launch_cfg<32, 4, 4, 8>(mode, exact, jobs, njobs, total_items, f, stream);
    default: return hipErrorInvalidValue;
  }
}

unsigned tile_threads(int config) {
  const TileConfig& c = kTileConfigs[config];
  return (c.tile / c.micro_i) * (c.tile / c.micro_j);
}

// Row r of the lower triangle with its diagonal (r + 1 values) from the square matrix, rows back
// to back at r (r + 1) / 2: the rows rapidNJMem keeps (distMatrixReader.cpp:26-35).
__global__ void lower_pack_kernel(const float* __restrict__ sq, size_t ld, unsigned n,
                                  float* __restrict__ out) {
  for (unsigned r = blockIdx.x; r < n; r += gridDim.x) {
    const float* row = sq + (size_t)r * ld;
    float* dst = out + (size_t)r * (r + 1) / 2;
    for (unsigned j = threadIdx.x; j <= r; j += blockDim.x) dst[j] = row[j];
  }
}

hipError_t launch_lower_pack(const float* sq, size_t ld, unsigned n, float* out,
                             hipStream_t stream) {
  if (n == 0) return hipSuccess;
  lower_pack_kernel<<<dim3(n < 65536 ? n : 65536), dim3(256), 0, stream>>>(sq, ld, n, out);
  return hipGetLastError();
}

unsigned long long sat_units(unsigned long long tiles, int config) {
  return tiles * (tile_threads(config) / 64);
}

hipError_t launch_saturation(const JobDesc* jobs, unsigned njobs, unsigned long long max_units,
                             int config, bool patch, hipStream_t stream) {
  if (njobs == 0 || max_units == 0) return hipSuccess;
  const TileConfig& c = kTileConfigs[config];
  // about 4 words per thread for the maximum; kPatchUnits units per wave for the patch
  const unsigned mx = (unsigned)std::min<unsigned long long>(2048, (max_units + 1023) / 1024);
  job_max_kernel<<<dim3(mx, njobs), dim3(256), 0, stream>>>(jobs, tile_threads(config) / 64);
  if (!patch) return hipGetLastError();
  const unsigned long long per_block = 4ull * kPatchUnits;
  if ((max_units + per_block - 1) / per_block >= (1ull << 31)) return hipErrorInvalidValue;
  const unsigned px = (unsigned)((max_units + per_block - 1) / per_block);
  patch_kernel<<<dim3(px, njobs), dim3(256), 0, stream>>>(jobs, c.tile, c.micro_i, c.micro_j);
  return hipGetLastError();
}

hipError_t launch_planes_to_codes(const uint32_t* planes, unsigned n_pad, unsigned W_pad,
                                  unsigned L, uint32_t* codes, hipStream_t stream, bool naive) {
  const unsigned groups = n_pad / 8;
  const dim3 grid((groups + 255) / 256, ygrid(L)), block(256);
  if (naive) planes_to_codes_kernel<true><<<grid, block, 0, stream>>>(planes, n_pad, W_pad, L, codes);
  else planes_to_codes_kernel<false><<<grid, block, 0, stream>>>(planes, n_pad, W_pad, L, codes);
  return hipGetLastError();
}

hipError_t launch_bootstrap_index(const uint16_t* weights, unsigned R, unsigned L,
                                  unsigned* col_index, unsigned idx_stride, unsigned* lengths,
                                  hipStream_t stream) {
  bootstrap_index_kernel<<<dim3(R), dim3(1024), 0, stream>>>(weights, L, col_index, idx_stride,
                                                             lengths);
  return hipGetLastError();
}

hipError_t launch_bootstrap_gather(const uint32_t* codes, unsigned n_pad, const unsigned* col_index,
                                   unsigned idx_stride, const unsigned* lengths, unsigned R,
                                   unsigned W_pad, uint32_t* planes, hipStream_t stream,
                                   bool naive) {
  const unsigned groups = n_pad / 8;
  const dim3 grid((groups + 63) / 64, ygrid(W_pad), R), block(64);
  if (naive)
    bootstrap_gather_kernel<true><<<grid, block, 0, stream>>>(codes, n_pad, col_index, idx_stride,
                                                              lengths, W_pad, planes);
  else
    bootstrap_gather_kernel<false><<<grid, block, 0, stream>>>(codes, n_pad, col_index,
                                                               idx_stride, lengths, W_pad, planes);
  return hipGetLastError();
}

}  // namespace rnj
