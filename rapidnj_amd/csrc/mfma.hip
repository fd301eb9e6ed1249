// Matrix-core path: the pair counts as Gram products of {-1, 0, +1} site embeddings.
//
// For a site where taxon t holds a counted base (V = 1) with code bits H (pyrimidine) and L,
// four components are +-1 or 0:  e0 = V,  e1 = V(-1)^H,  e2 = V(-1)^L,  e3 = V(-1)^(H^L).
// Summed over the sites of a pair, with n00/n01/n10/n11 the valid sites whose (H, L) bits
// differ in (no, no)/(no, yes)/(yes, no)/(yes, yes):
//   A = sum e0 e0 = valid = n00+n01+n10+n11     B = sum e1 e1 = n00+n01-n10-n11
//   C = sum e2 e2 = n00-n01+n10-n11             D = sum e3 e3 = n00-n01-n10+n11
// so tv = n10 + n11 = (A - B) / 2 and ts = n01 = (A + B - C - D) / 4, exactly the reference's
// transversions (high bits differ) and transitions (only the low bit differs,
// bitDistanceGap.cpp:138-161). Four multiply-adds per site-pair.
//
// The products and sums are exact: the operands are FP4 (e2m1) +1 = 0x2, -1 = 0xA, 0 = 0x0,
// accumulated in FP32, whose integers are exact below 2^24 (alignments of fewer than 2^24
// sites). v_mfma_f32_32x32x64_f8f6f4 (unscaled, both operands FP4) multiplies a 32 x 64 by a
// 64 x 32 block: lane l holds 32 nibbles (4 dwords) of row/column l % 32 for K-half l / 32, and
// A and B use the same nibble-to-K convention, so feeding the same fragment layout on both
// sides forms a Gram block whatever the hardware's internal K order. Result C[m][n] sits in
// lane n + 32*((m/4)%2), register 4*(m/8) + m%4 (checked on MI355X with a one-wave probe).
#include "device_common.hpp"

namespace rnj {
namespace {

typedef int v8i __attribute__((ext_vector_type(8)));
typedef float v16f __attribute__((ext_vector_type(16)));

// Bits 0..7 of x to bits 0, 4, .., 28.
__device__ __forceinline__ uint32_t spread8(uint32_t x) {
  x &= 0xffu;
  x = (x | (x << 12)) & 0x000F000Fu;
  x = (x | (x << 6)) & 0x03030303u;
  x = (x | (x << 3)) & 0x11111111u;
  return x;
}

// FP4 nibbles of 8 sites: +1 where counted and not negated, -1 where counted and negated.
__device__ __forceinline__ uint32_t fp4_8(uint32_t v8, uint32_t neg8) {
  return (spread8(v8) << 1) | (spread8(v8 & neg8) << 3);
}

// Fragment (64-site block b, component c, 32-taxon tile tt) of the embedding: 256 u32 at
// ((b * 4 + c) * ntile + tt) * 256; lane (t % 32) + 32 h holds dwords [4 lane, 4 lane + 4),
// nibble k of dword d being site 64 b + 32 h + 8 d + k. One thread per (taxon, block, half).
__global__ void pack_fp4_kernel(const uint32_t* __restrict__ planes, unsigned n_pad,
                                unsigned W_pad, unsigned nblocks, uint32_t* __restrict__ emb) {
  const unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_pad) return;
  const size_t ps = (size_t)W_pad * n_pad;
  const unsigned ntile = n_pad / 32;
  for (unsigned b = blockIdx.y; b < nblocks; b += gridDim.y) {
    for (unsigned h = 0; h < 2; ++h) {
      const size_t o = (size_t)(2 * b + h) * n_pad + t;
      const uint32_t H = planes[o], Lo = planes[ps + o], V = planes[2 * ps + o];
      const uint32_t neg[4] = {0u, H, Lo, H ^ Lo};
#pragma unroll
      for (unsigned c = 0; c < 4; ++c) {
        uint4 q;
        q.x = fp4_8(V, neg[c]);
        q.y = fp4_8(V >> 8, neg[c] >> 8);
        q.z = fp4_8(V >> 16, neg[c] >> 16);
        q.w = fp4_8(V >> 24, neg[c] >> 24);
        uint4* dst = reinterpret_cast<uint4*>(emb + ((size_t)(b * 4 + c) * ntile + t / 32) * 256) +
                     (t % 32) + 32 * h;
        *dst = q;
      }
    }
  }
}

// Work-group barrier that orders LDS accesses only: __syncthreads() also releases global
// memory, which makes every wave wait for its in-flight prefetch loads at each barrier.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

__device__ __forceinline__ v16f mfma_fp4(const v4u& a, const v4u& b, const v16f& c) {
  const v8i av = {(int)a.x, (int)a.y, (int)a.z, (int)a.w, 0, 0, 0, 0};
  const v8i bv = {(int)b.x, (int)b.y, (int)b.z, (int)b.w, 0, 0, 0, 0};
  // scale operands 0: the compiler emits the unscaled v_mfma_f32_32x32x64_f8f6f4
  return __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(av, bv, c, 4, 4, 0, 0, 0, 0);
}

// Counts of one wave's 64 x 64 pairs (in its accumulators) to the job's output, one 32 x 32 MFMA
// block at a time: the block's three accumulators go to the wave's 12 KB of LDS as integers
// (cnt[o][r][lane]), then a loop that need not be unrolled takes four registers r = 4g .. 4g+3 at
// a time, which hold four consecutive rows i of column j = lane % 32. The four corrections of such
// a group are straight-line code that the compiler interleaves (one wave per SIMD leaves nothing
// else to hide their latency); the stores and the rare sentinels follow. Pairs with j >= n or
// i >= j are skipped; i == j zeroes the diagonal of a square output.
template <int kModel, bool kExact>
__device__ __forceinline__ void mfma_epilogue(const JobDesc& J, const v16f (&acc)[2][2][3],
                                              uint32_t* cnt, unsigned bi, unsigned bj,
                                              unsigned wr, unsigned wc, unsigned lane,
                                              unsigned tid) {
  const unsigned n = J.n;
  float lmax = 0.0f;
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y) {
#pragma unroll
      for (int o = 0; o < 3; ++o)
#pragma unroll
        for (int r = 0; r < 16; ++r) cnt[(o * 16 + r) * 64 + lane] = (uint32_t)(int)acc[x][y][o][r];
      const unsigned j = bj * 128 + (2 * wc + y) * 32 + (lane % 32);
#pragma unroll 1
      for (int g = 0; g < 4; ++g) {
        const unsigned i0 = bi * 128 + (2 * wr + x) * 32 + 8 * g + 4 * (lane / 32);
        uint32_t ts[4], tv[4], va[4];
        float d[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int r = 4 * g + u;
          const int A = (int)cnt[(0 * 16 + r) * 64 + lane];
          const int B = (int)cnt[(1 * 16 + r) * 64 + lane];
          const int CD = (int)cnt[(2 * 16 + r) * 64 + lane];
          tv[u] = (uint32_t)((A - B) / 2);
          ts[u] = (uint32_t)((A + B - CD) / 4);
          va[u] = (uint32_t)A;
          if (kModel == kK2P) d[u] = k2p_value<kExact, true>(ts[u], tv[u], va[u]);
          if (kModel == kJC) d[u] = jc_value<kExact, true>(ts[u] + tv[u], va[u]);
        }
        if (j < n) {
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const unsigned i = i0 + u;
            if (i < j) {
              if (kModel == kCounts) {
                const unsigned long long q = tri_index(i, j, n);
                J.out_tv[q] = tv[u];
                J.out_ts[q] = ts[u];
                J.out_valid[q] = va[u];
                continue;
              }
              if (d[u] == -1.0f) push_sentinel(J, i, j);
              else lmax = fmaxf(lmax, d[u]);
              if (J.layout == kSquare) {
                J.out[(size_t)i * J.ld + j] = d[u];
                J.out[(size_t)j * J.ld + i] = d[u];
              } else {
                J.out[tri_index(i, j, n)] = d[u];
              }
            } else if (i == j) {
              finish_diagonal(J, i);
            }
          }
        }
      }
    }
  if (kModel != kCounts) fold_max(J, lmax, tid);
}

// One work-group (4 waves) computes a 128 x 128 block of taxon pairs (bi <= bj); wave w takes the
// 64 x 64 quarter (row half w / 2, column half w % 2) as 2 x 2 MFMA blocks. Per 64-site block
// the work-group stages the 16 A and 16 B fragments (4 components x 4 taxon tiles, 32 KB) in LDS
// (two stages) and each wave issues 16 MFMAs into three accumulators per MFMA block: A, B and
// C + D. Global loads run PF blocks ahead of the MFMAs in registers (8 x 16 bytes per thread per
// block), enough to cover L2 / HBM latency at one work-group per CU; tiles follow patch order so
// that the work-groups resident on one XCD share their row and column blocks in L2.
template <bool kExact, int PF>
__global__ __launch_bounds__(256) void mfma_kernel(const JobDesc* __restrict__ jobs,
                                                   unsigned njobs) {
  constexpr int kFrag = 64;                    // v4u per fragment (one per lane)
  constexpr int kStage = 2 * 4 * 4 * kFrag;    // [A/B][component][tile][lane]
  constexpr int kPer = kStage / 256;           // v4u per thread per block
  static_assert(PF % 2 == 0 && kWordPad % (2 * PF) == 0, "steps must tile the padded words");
  __shared__ v4u sfrag[2 * kStage];

  const unsigned item = xcd_item(blockIdx.x, gridDim.x);
  const JobDesc& J = jobs[find_job(jobs, njobs, item)];
  unsigned bi, bj;
  patch_coords(item - J.item_base, J.nblk, bi, bj);
  const unsigned tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const unsigned wr = wave >> 1, wc = wave & 1;
  const unsigned ntile = J.n_pad / 32;
  const unsigned nb64 = J.W / 2;  // > 0: the word count is a positive multiple of 16
  const gptr<v4u> emb = (gptr<v4u>)J.emb;

  // Thread tid moves element e = tid + 256 k of a stage: A/B = k / 4, component k % 4, taxon
  // tile `wave` of the block, v4u `lane`; its source is fragment ((b*4 + c)*ntile + 4*blk + wave).
  const size_t block_step = (size_t)4 * ntile * kFrag;
  const size_t base_a = ((size_t)4 * bi + wave) * kFrag + lane;
  const size_t base_b = ((size_t)4 * bj + wave) * kFrag + lane;
  auto load = [&](unsigned b, v4u (&dst)[kPer]) {
    const size_t o = b * block_step;
#pragma unroll
    for (int k = 0; k < kPer; ++k)
      dst[k] = emb[o + (size_t)(k & 3) * ntile * kFrag + (k < 4 ? base_a : base_b)];
  };

  v16f acc[2][2][3];
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y)
#pragma unroll
      for (int o = 0; o < 3; ++o) acc[x][y][o] = v16f{};

  // The word count is a multiple of 16, so nb64 is a multiple of 8 and of PF: every pass of the
  // outer loop runs PF whole steps. Loads past the end re-read the last block (never used), which
  // keeps the loop free of branches so the compiler's in-order load counting stays exact.
  //
  // Step b reads LDS stage b & 1 (written during step b - 1) and, between its MFMAs, writes step
  // b + 1's fragments from registers into the other stage; one LDS-only barrier ends the step.
  // pre[q] holds the fragments of the steps congruent to q mod PF, loaded PF - 1 steps ahead.
  auto last = [nb64](unsigned b) { return b < nb64 ? b : nb64 - 1; };
  v4u pre[PF][kPer];
#pragma unroll
  for (int p = 0; p < PF; ++p) {
    load(last(p), pre[p]);
    __builtin_amdgcn_sched_barrier(0);  // issue in slot order, as the loop consumes them
  }
#pragma unroll
  for (int k = 0; k < kPer; ++k) sfrag[tid + 256 * k] = pre[0][k];
  load(last(PF), pre[0]);
  __builtin_amdgcn_sched_barrier(0);
  lds_barrier();

  for (unsigned b0 = 0; b0 < nb64; b0 += PF) {
#pragma unroll
    for (int p = 0; p < PF; ++p) {
      const unsigned b = b0 + p;
      const v4u* stage = sfrag + (p & 1) * kStage;
      v4u* next = sfrag + ((p + 1) & 1) * kStage;
      const int q = (p + 1) % PF;
      v4u op[2][4];
      auto read_ops = [&](int c, v4u (&o)[4]) {
        o[0] = stage[((0 * 4 + c) * 4 + 2 * wr + 0) * kFrag + lane];
        o[1] = stage[((0 * 4 + c) * 4 + 2 * wr + 1) * kFrag + lane];
        o[2] = stage[((1 * 4 + c) * 4 + 2 * wc + 0) * kFrag + lane];
        o[3] = stage[((1 * 4 + c) * 4 + 2 * wc + 1) * kFrag + lane];
      };
      read_ops(0, op[0]);
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        if (c < 3) read_ops(c + 1, op[(c + 1) & 1]);
        next[tid + 256 * (2 * c)] = pre[q][2 * c];
        next[tid + 256 * (2 * c + 1)] = pre[q][2 * c + 1];
        const v4u(&o4)[4] = op[c & 1];
        const int o = c < 2 ? c : 2;
        acc[0][0][o] = mfma_fp4(o4[0], o4[2], acc[0][0][o]);
        acc[0][1][o] = mfma_fp4(o4[0], o4[3], acc[0][1][o]);
        acc[1][0][o] = mfma_fp4(o4[1], o4[2], acc[1][0][o]);
        acc[1][1][o] = mfma_fp4(o4[1], o4[3], acc[1][1][o]);
      }
      load(last(b + 1 + PF), pre[q]);
      __builtin_amdgcn_sched_barrier(0);
      lds_barrier();
    }
  }

  // Every wave has passed the last step's barrier, so the stages are free: each wave takes 12 KB.
  uint32_t* cnt = reinterpret_cast<uint32_t*>(sfrag) + wave * (3 * 16 * 64);
  if (J.model == kCounts)
    mfma_epilogue<kCounts, kExact>(J, acc, cnt, bi, bj, wr, wc, lane, tid);
  else if (J.model == kJC)
    mfma_epilogue<kJC, kExact>(J, acc, cnt, bi, bj, wr, wc, lane, tid);
  else
    mfma_epilogue<kK2P, kExact>(J, acc, cnt, bi, bj, wr, wc, lane, tid);
}

}  // namespace

hipError_t launch_pack_fp4(const uint32_t* planes, unsigned n_pad, unsigned W_pad,
                           uint32_t* emb, hipStream_t stream) {
  const unsigned nblocks = W_pad / 2;
  const unsigned gy = nblocks < 1 ? 1 : (nblocks > 65535 ? 65535 : nblocks);
  pack_fp4_kernel<<<dim3((n_pad + 255) / 256, gy), dim3(256), 0, stream>>>(planes, n_pad, W_pad,
                                                                           nblocks, emb);
  return hipGetLastError();
}

template <int PF>
static void launch_depth(bool exact, const JobDesc* jobs, unsigned njobs, unsigned total,
                         hipStream_t stream) {
  if (exact)
    mfma_kernel<true, PF><<<dim3(total), dim3(256), 0, stream>>>(jobs, njobs);
  else
    mfma_kernel<false, PF><<<dim3(total), dim3(256), 0, stream>>>(jobs, njobs);
}

hipError_t launch_mfma(unsigned depth, bool exact, const JobDesc* jobs, unsigned njobs,
                       unsigned total, hipStream_t stream) {
  if (total == 0) return hipSuccess;
  switch (depth) {
    case 2: launch_depth<2>(exact, jobs, njobs, total, stream); break;
    case 4: launch_depth<4>(exact, jobs, njobs, total, stream); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace rnj
