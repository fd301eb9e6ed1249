// Neighbour joining on the device: rapidNJ::run (rapidNJ.cpp:45-88) with the matrix in HBM.
//
// rapidNJ keeps each row's partners sorted by distance (std::sort on the distance alone) and
// scans a row's live pairs in that order until the first whose float
// distance - (sep[row] + maxRowSeparation[row]) is not below the running minimum; it joins the
// first pair, in its visiting order, that lowers the minimum of the criterion
// (d - sep[partner]) - sep[row] last. Here every live pair of every row that can hold a record
// is evaluated each iteration (a pass over those rows of the live matrix, which HBM and the
// 256 MB last-level cache serve at several TB/s; see "Row skipping" below), with the
// reference's float arithmetic and visiting order:
//   * a pair (a, b) lives in the row of the cluster with the larger id; the matrix is kept in
//     "slots" ordered by cluster id (new clusters take the next slot), so row slot a's partners
//     are exactly the live slots b < a;
//   * rows are visited by matrix index, cyclically from the row rapidNJ cached in the previous
//     search, and a row's pairs by increasing distance;
//   * a row with redundant copies offers the copy join at -2 * sep first.
// Two things the full evaluation cannot see are left to the host, which the step halts for
// (NjState::error = kNjHalted) and which then supplies the search result (NjState::forced):
//   * exact ties: the winning row has several pairs of equal criterion and equal distance, and
//     rapidNJ takes the one std::sort put first in the row's list;
//   * the scan's break: its bound is rounded differently from the criterion, so a row can stop
//     one pair before a criterion one unit in the last place below the running minimum. The scan
//     gives each row a bound sigma (below) such that the row's own result stands whenever the
//     running minimum that reaches it is above sigma; a record row reached at or below it
//     halts the step.
// Per iteration: nj_scan_kernel (16 row slots per work-group, one per wave: each row's best
// pair; it also writes the previous join's column) and nj_step_kernel (one work-group: the
// global winner and rapidNJ's row caches, the merge with its redundancy bookkeeping, the update
// of distances and separations in id order, the new slot's row, logged for the host). The new
// cluster's separation sum (a sequential float sum in id order) is finished by the next scan.
// Every few iterations nj_compact_* rebuilds the slot matrix from the live slots. The host
// enqueues iterations in windows without waiting: the iteration count (n - 2) and the slot
// bounds are known ahead.
//
// Row skipping. Most rows cannot be records, and a scan skips a row it can prove is not one.
// Each full scan of row a keeps m_a = min over its live pairs of float(d - sep[b]) and the
// cumulative drift D at that scan, where each step adds its largest rise of any live
// separation. Until row a is scanned again its pairs keep their d, and each partner's
// separation has risen by at most D(now) - D(then) (partners only go), so every criterion of
// the row is at least L_a = m_a - (D(now) - D(then)) - sep[a], less a rounding margin (1e-5
// relative, far above the few float roundings involved). The step gives the next search's start
// row (rapidNJ's caches name it) an upper bound U of its answer: the criterion, in the scan's
// arithmetic, of the pair that was that row's best at its last scan, if the partner is still a
// live slot below it (and its copy join). The start row is visited first, so the running
// minimum that reaches any other row is at most U, and a row with L_a >= U cannot be a record
// (records are strictly below). Such a row is not scanned: it publishes L_a as its key (never a
// record for the step's search) and flags itself (row_tie = 2), so that the host, if it
// settles that search, scans it over its sorted list like any row whose answer does not stand.
// The new cluster's separation sum (a sequential float sum, about 10 cycles per term) is then
// what bounds a scan launch's length, so this saves about a fifth of the scan time.

#include "nj.hpp"

#include <float.h>

namespace rnj {
namespace {

constexpr int kScanThreads = 1024;  // 16 waves, one row each
constexpr int kScanRows = kScanThreads / 64;
constexpr int kSepChunk = 4096;     // separations staged in LDS per step of a row scan
constexpr int kStepThreads = 1024;
constexpr int kStepWaves = kStepThreads / 64;
constexpr int kKeyCache = 32768;    // row keys the step's search keeps in LDS
constexpr uint32_t kNone = 0xffffffffu;

// Floats as order-preserving unsigned keys (after -0 -> +0, so that equal values compare equal
// as the reference's float comparisons do).
__device__ __forceinline__ uint32_t fkey(float x) {
  const uint32_t u = __float_as_uint(x + 0.0f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

constexpr float kInf = __builtin_inff();
constexpr unsigned long long kNoKey = ~0ull;

// A row's best pair: criterion v, distance d, partner index p, and the smallest criterion among
// the row's other pairs (v2; v2 == v when another pair shares the best criterion). Among pairs
// of equal criterion any may stand: such a row is settled by the host whenever it wins, and its
// break bound (below) is conservative for either distance. Merged by (v, d, p), with NaN v (gone
// partners) never better.
struct Cand {
  float v, d;
  int p;
  float v2;
};

// Merge o into c (symmetric, so that both lanes of a shuffle agree).
__device__ __forceinline__ void cand_merge(Cand& c, const Cand& o) {
  if (o.v < c.v) {
    const float v2 = fminf(o.v2, c.v);
    c = o;
    c.v2 = v2;
  } else if (c.v < o.v) {
    c.v2 = fminf(c.v2, o.v);
  } else {
    c.v2 = fminf(fminf(c.v2, o.v2), c.v);
    if (o.d < c.d || (o.d == c.d && o.p < c.p)) {
      c.d = o.d;
      c.p = o.p;
    }
  }
}

__device__ __forceinline__ Cand wave_merge(Cand c) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const Cand o{__shfl_xor(c.v, off), __shfl_xor(c.d, off), __shfl_xor(c.p, off),
                 __shfl_xor(c.v2, off)};
    cand_merge(c, o);
  }
  return c;
}

__device__ __forceinline__ float wave_fmax(float x) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) x = fmaxf(x, __shfl_xor(x, off));
  return x;
}

__device__ __forceinline__ float wave_fmin(float x) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) x = fminf(x, __shfl_xor(x, off));
  return x;
}

// A fully scanned row's skip bound: the smallest float(d - sep[b]) of its live pairs, with the
// drift it was taken at.
__device__ __forceinline__ void keep_bound(const NjDevice& D, int a, float mn) {
  D.smin[a] = mn;
  D.sref[a] = D.st->drift;
  D.shave[a] = 1;
}

// One lane's running state over its share of a row: best criterion, its distance and slot, the
// smallest criterion of the lane's other pairs, and the largest separation seen
// (maxRowSeparation's running max, from FLT_MIN).
struct Lane {
  float bv, bd, v2, mx, mn;  // mn: the smallest float(d - sep[b]) (the row's skip bound)
  int bb;
};

__device__ __forceinline__ Lane lane_start() { return Lane{kInf, kInf, kInf, FLT_MIN, kInf, -1}; }

// Candidate (d, separation) of partner slot b into this lane's state: the criterion, then five
// operations (NaN from gone slots leaves every field unchanged).
__device__ __forceinline__ void consider(float d, float s_b, float s_a, int b, Lane& l) {
  const float t = __fsub_rn(d, s_b);
  const float v = __fsub_rn(t, s_a);
  l.mn = fminf(l.mn, t);
  l.mx = fmaxf(l.mx, s_b);
  const bool lt = v < l.bv;
  l.v2 = fminf(l.v2, lt ? l.bv : v);
  l.bv = lt ? v : l.bv;
  l.bd = lt ? d : l.bd;
  l.bb = lt ? b : l.bb;
}

// A lane's state as a candidate for the wave merge.
__device__ __forceinline__ Cand lane_cand(const NjDevice& D, const Lane& l) {
  if (l.bb < 0) return Cand{kInf, kInf, 0x7fffffff, l.v2};
  return Cand{l.bv, l.bd, D.slot_index[l.bb], l.v2};
}

// The row's answer as rapidNJ's findRowMin gives it: the redundant-copy join first when the row
// has copies (searchRedundantEntries), then the best pair unless it is not strictly better.
// Rows whose answer does not beat FLT_MAX (the search's start value) get no key.
// The break bound: rapidNJ reaches the best pair e (distance d, criterion v) unless the running
// minimum G it meets there is at most t = float(d - float(sep[a] + maxsep)). When t <= v that
// cannot happen (G > v at e, or e would not improve it). Otherwise G at e is the minimum that
// reaches the row, lowered first by the copy join and by pairs before e, whose criteria are at
// least v2: with v2 and the copy join above t, the row stands exactly when the minimum reaching
// it is above sigma = t; else sigma = +inf (the host searches whenever the row is a record).
__device__ __forceinline__ void write_row(const NjDevice& D, int a, Cand best, float s_a,
                                          float maxsep) {
  const int idx = D.slot_index[a];
  const bool has_pair = best.p != 0x7fffffff && best.v < kInf;
  const bool copies = D.red_count[idx] > 1;
  const float cv = copies ? -2.0f * s_a : kInf;
  const bool pair = has_pair && best.v < cv;
  float v = kInf;
  int p = -1;
  if (pair) {
    v = best.v;
    p = best.p;
  } else if (copies) {
    v = cv;
    p = D.red_head[idx];
  }
  float sigma = -kInf;
  if (pair) {
    const float t = __fsub_rn(best.d, __fadd_rn(s_a, maxsep));
    if (t > best.v) sigma = (best.v2 <= t || cv <= t) ? kInf : t;
  }
  D.row_key[a] = v < FLT_MAX ? ((unsigned long long)fkey(v) << 32) | (uint32_t)idx : kNoKey;
  D.key_by_index[idx] = v < FLT_MAX ? fkey(v) : kNone;
  D.row_p[a] = p;
  D.row_sigma[a] = fkey(sigma);
  D.row_tie[a] = pair && best.v2 <= best.v ? 1 : 0;
  D.row_maxsep[a] = maxsep;
}

// Finish the previous join's new cluster (slot used - 1), on one work-group: the reference's
// sequential float sum of its separation sum in id order (terms of gone or merged slots are +0,
// which leaves a sum that is never -0 unchanged), its separation, and its row's best pair.
__device__ void finish_new_cluster(const NjDevice& D, const float* __restrict__ S) {
  NjState& st = *D.st;
  const int ns = st.used - 1;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  __shared__ Cand wbest[kScanThreads / 64];
  __shared__ float wmax[kScanThreads / 64];
  __shared__ float wmin[kScanThreads / 64];
  constexpr int kChunk = 4096;
  __shared__ float terms[kChunk];
  float acc = 0.0f;
  for (int base = 0; base < ns; base += kChunk) {
    const int cnt = min(kChunk, ns - base);
    for (int k = tid; k < kChunk; k += kScanThreads) terms[k] = k < cnt ? D.term[base + k] : 0.0f;
    __syncthreads();
    if (wave == 0) {
      const float4* t4 = reinterpret_cast<const float4*>(terms);
      for (int k = 0; k < (cnt + 3) / 4; ++k) {
        const float4 x = t4[k];
        acc = __fadd_rn(acc, x.x);
        acc = __fadd_rn(acc, x.y);
        acc = __fadd_rn(acc, x.z);
        acc = __fadd_rn(acc, x.w);
      }
    }
    __syncthreads();
  }
  if (tid == 0) {
    D.sepsum[ns] = acc;
    D.sep[ns] = acc / (float)(st.cluster_count - 2);
  }
  __syncthreads();
  const float s_ns = D.sep[ns];
  // the new row: partners are every live slot below it
  Lane l = lane_start();
  for (int b = tid; b < ns; b += kScanThreads) consider(D.newcol[b], D.sep[b], s_ns, b, l);
  Cand best = wave_merge(lane_cand(D, l));
  const float mx = wave_fmax(l.mx);
  const float mn = wave_fmin(l.mn);
  if (lane == 0) {
    wbest[wave] = best;
    wmax[wave] = mx;
    wmin[wave] = mn;
  }
  __syncthreads();
  if (tid == 0) {
    Cand b = wbest[0];
    float m = wmax[0], lo = wmin[0];
    for (int w = 1; w < kScanThreads / 64; ++w) {
      cand_merge(b, wbest[w]);
      m = fmaxf(m, wmax[w]);
      lo = fminf(lo, wmin[w]);
    }
    write_row(D, ns, b, s_ns, m);
    keep_bound(D, ns, lo);
  }
}

// Launch A: work-group 0 finishes the previous join's new cluster when asked; every other
// work-group scans 16 consecutive row slots, one per wave, over partner slots b < row, the
// highest slots in the first work-groups. The
// separations of the partners are staged in LDS once per work-group and shared by its rows.
// With `finish`, each wave also writes its row's entry in the new slot's column (S[a][ns] =
// newcol[a], a < ns): the step wrote the new row, and no scan reads that column (row a's
// partners are the slots below a), so the column is complete before the next step reads it.
__global__ __launch_bounds__(kScanThreads) void nj_scan_kernel(NjDevice D, float* __restrict__ S,
                                                               int finish) {
  const NjState& st = *D.st;
  if (st.error) return;
  int blk = blockIdx.x;
  if (finish) {
    if (blk == 0) {
      finish_new_cluster(D, S);
      return;
    }
    --blk;
  }
  __shared__ float ssep[kSepChunk];
  __shared__ int s_reach;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int used = st.used;
  // longest rows first: row a has a partners, so the groups go out in decreasing slot order and
  // the launch ends on its shortest rows
  const int groups = (int)gridDim.x - (finish ? 1 : 0);
  const int a0 = (groups - 1 - blk) * kScanRows;
  if (a0 >= used) return;  // uniform over the work-group
  const int a = a0 + wave;
  if (finish && lane == 0 && a < used - 1) S[(size_t)a * D.ld + (used - 1)] = D.newcol[a];
  const bool mine = a < used && !(finish && a == used - 1) && D.alive[a];
  const float s_a = mine ? D.sep[a] : 0.0f;
  // skip the row when its bound shows it cannot be a record (see the top of this file)
  bool scan = mine;
  const float U = st.start_bound;
  if (mine && U < FLT_MAX && D.shave[a] && a != D.slot_of_index[st.start_row]) {
    const int idx = D.slot_index[a];
    if (D.red_count[idx] <= 1) {
      const float m = D.smin[a];
      const float dr = (float)(st.drift - D.sref[a]);
      float lb = __fsub_rn(__fsub_rn(m, dr), s_a);
      lb = __fsub_rn(lb, 1e-5f * (1.0f + fabsf(m) + dr + fabsf(s_a)));
      if (lb >= U) {
        scan = false;
        if (lane == 0) {
          D.row_key[a] = ((unsigned long long)fkey(lb) << 32) | (uint32_t)idx;
          D.key_by_index[idx] = fkey(lb);
          D.row_sigma[a] = fkey(kInf);
          D.row_tie[a] = 2;
        }
      }
    }
  }
  if (tid == 0) s_reach = 0;
  __syncthreads();
  if (scan && lane == 0) atomicMax(&s_reach, a);
  __syncthreads();
  const int reach = s_reach;  // longest row prefix this group scans (0: none)
  const float* row = S + (size_t)a * D.ld;
  Lane l = lane_start();
  for (int c0 = 0; c0 < reach; c0 += kSepChunk) {
    const int cnt = min(kSepChunk, reach - c0);
    for (int k = 4 * tid; k < cnt; k += 4 * kScanThreads)
      *reinterpret_cast<float4*>(ssep + k) = *reinterpret_cast<const float4*>(D.sep + c0 + k);
    __syncthreads();
    if (scan) {
      // four 4-wide loads in flight per lane
      const int end = min(a, c0 + cnt);
      for (int b0 = c0 + 4 * lane; b0 < end; b0 += 1024) {
        float4 dv[4];
#pragma unroll
        for (int u = 0; u < 4; ++u)
          if (b0 + 256 * u < end) dv[u] = *reinterpret_cast<const float4*>(row + b0 + 256 * u);
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int bu = b0 + 256 * u;
          if (bu >= end) break;
          const float4 sv = *reinterpret_cast<const float4*>(ssep + (bu - c0));
          consider(dv[u].x, sv.x, s_a, bu, l);
          if (bu + 1 < end) consider(dv[u].y, sv.y, s_a, bu + 1, l);
          if (bu + 2 < end) consider(dv[u].z, sv.z, s_a, bu + 2, l);
          if (bu + 3 < end) consider(dv[u].w, sv.w, s_a, bu + 3, l);
        }
      }
    }
    __syncthreads();
  }
  if (!scan) return;
  const Cand best = wave_merge(lane_cand(D, l));
  // maxRowSeparation: the largest separation of the live clusters of smaller id (rapidNJ.cpp:
  // 177-214), and for the first search that of every smaller index (:114-119)
  const float mx = wave_fmax(l.mx);
  const float mn = wave_fmin(l.mn);
  if (lane == 0) {
    write_row(D, a, best, s_a, st.step == 0 ? D.init_maxsep[a] : mx);
    keep_bound(D, a, mn);
  }
}

// Launch A's finishing work alone (before a compaction, which copies whole rows and columns):
// the new cluster's separation and row, and its column.
__global__ __launch_bounds__(kScanThreads) void nj_finish_kernel(NjDevice D, float* __restrict__ S) {
  if (D.st->error) return;
  const int ns = D.st->used - 1;
  for (int b = threadIdx.x; b < ns; b += kScanThreads) S[(size_t)b * D.ld + ns] = D.newcol[b];
  finish_new_cluster(D, S);
}

// ---- the one-work-group step ------------------------------------------------------------------
// Exclusive prefix (in thread order) of x under min (identity kNone) or + (identity 0) over
// kStepThreads threads; *total gets the whole block's (`sh` holds kStepWaves words).
template <bool kMin>
__device__ __forceinline__ uint32_t block_exclusive(uint32_t x, uint32_t* sh, uint32_t* total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t inc = x;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t o = (uint32_t)__shfl_up((int)inc, off);
    if (lane >= off) inc = kMin ? min(inc, o) : inc + o;
  }
  if (lane == 63) sh[wave] = inc;
  __syncthreads();
  uint32_t before = kMin ? kNone : 0u, all = kMin ? kNone : 0u;
#pragma unroll
  for (int w = 0; w < kStepWaves; ++w) {
    const uint32_t y = sh[w];
    if (w < wave) before = kMin ? min(before, y) : before + y;
    all = kMin ? min(all, y) : all + y;
  }
  __syncthreads();
  const uint32_t excl_lane = (uint32_t)__shfl_up((int)inc, 1);
  const uint32_t mine = lane == 0 ? (kMin ? kNone : 0u) : excl_lane;
  *total = all;
  return kMin ? min(before, mine) : before + mine;
}

// findMin over the rows' answers (rapidNJ.cpp:290-321): rows are visited by matrix index,
// cyclically from `start`; a record is a row whose key is below every earlier one (and below
// FLT_MAX's). The keys come by matrix index (key_by_index: kNone for rows without an answer,
// gone rows and hidden copies), read coalesced into LDS, then each thread takes a run of
// consecutive visiting positions: a block-wide exclusive minimum gives each run the minimum
// before it. Gives the last four records' slots in visiting order (s_rec[0..*s_nrec-1]) and sets
// *s_event when a record meets a running minimum at or below its break bound.
__device__ void search_records(const NjDevice& D, int n, int start, uint32_t* kc, int* s_rec,
                               int* s_nrec, int* s_event, uint32_t* shu) {
  const int tid = threadIdx.x;
  auto key_at = [&](int c) {  // visiting position c: index start + c (mod n)
    int idx = start + c;
    if (idx >= n) idx -= n;
    return D.key_by_index[idx];
  };
  const int cached = min(n, kKeyCache);
  for (int c0 = 0; c0 < cached; c0 += 4 * kStepThreads) {  // four loads in flight per thread
    uint32_t k[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int c = c0 + u * kStepThreads + tid;
      k[u] = c < cached ? key_at(c) : kNone;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int c = c0 + u * kStepThreads + tid;
      if (c < cached) kc[c] = k[u];
    }
  }
  __syncthreads();
  const int per = (n + kStepThreads - 1) / kStepThreads;
  const int c0 = min(n, tid * per), c1 = min(n, c0 + per);
  uint32_t seg = kNone;
  for (int c = c0; c < c1; ++c) seg = min(seg, c < kKeyCache ? kc[c] : key_at(c));
  uint32_t unused;
  uint32_t G = block_exclusive<true>(seg, shu, &unused);
  G = min(G, fkey(FLT_MAX));  // the search starts at FLT_MAX (rapidNJ.cpp:291)
  int cnt = 0, r0 = -1, r1 = -1, r2 = -1, r3 = -1;  // this thread's last records, newest first
  bool unsafe = false;
  for (int c = c0; c < c1; ++c) {
    const uint32_t key = c < kKeyCache ? kc[c] : key_at(c);
    if (key < G) {
      int idx = start + c;
      if (idx >= n) idx -= n;
      const int slot = D.slot_of_index[idx];
      unsafe |= G <= D.row_sigma[slot];
      G = key;
      r3 = r2;
      r2 = r1;
      r1 = r0;
      r0 = slot;
      ++cnt;
    }
  }
  uint32_t total;
  const int rank0 = (int)block_exclusive<false>((uint32_t)cnt, shu, &total);
  const int R = (int)total, first = R > 4 ? R - 4 : 0;
  const int mine[4] = {r0, r1, r2, r3};
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int rank = rank0 + cnt - 1 - j;  // the j-th newest of this thread's records
    if (j < cnt && rank >= first) s_rec[rank - first] = mine[j];
  }
  if (tid == 0) {
    *s_nrec = R < 4 ? R : 4;
    *s_event = 0;
  }
  __syncthreads();
  if (unsafe) *s_event = 1;
  __syncthreads();
}

__device__ __forceinline__ int red_pop(const NjDevice& D, int i) {
  const int j = D.red_head[i];
  D.red_head[i] = D.red_next[j];
  return j;
}

// Launch B, one work-group: findMin's result and caches (rapidNJ.cpp:290-345) -- searched here,
// or the host's when it set NjState::forced --, mergeMinNodes (:240-281), clusterCount--,
// updateData (:166-238) except the new separation sum, which the next launch A finishes. The new
// row goes to the log at log_off (by slot: distance and partner index, -1 where none).
__global__ __launch_bounds__(kStepThreads) void nj_step_kernel(NjDevice D, float* __restrict__ S,
                                                               size_t log_off) {
  NjState& st = *D.st;
  const int tid = threadIdx.x;
  __shared__ uint32_t key_cache[kKeyCache];
  __shared__ uint32_t shu[kStepWaves];
  __shared__ int s_rec[4], s_nrec, s_event;
  __shared__ int s_rows[2], s_skip[2], s_fail;
  __shared__ float s_mutual;
  __shared__ unsigned s_rise;  // the largest separation rise (non-negative float bits)
  if (st.error) return;
  const int n = st.n, used = st.used;
  const bool forced = st.forced != 0;
  if (!forced) {
    const int start = D.index_to_id[st.cache1] != -1 ? st.cache1 : st.cache2;
    search_records(D, n, start, key_cache, s_rec, &s_nrec, &s_event, shu);
  }
  if (tid == 0) {
    // the winner must name live rows with slots, and the new cluster needs a free slot
    int min1 = 0, min2 = 0, r1 = 0, r2 = 0, s1 = 0, s2 = 0;
    int c1 = st.cache1, c2 = st.cache2;
    bool bad = used >= (int)D.ld;
    bool halt = false;
    if (forced) {
      min1 = st.f_min1;
      min2 = st.f_min2;
      c1 = st.f_cache1;
      c2 = st.f_cache2;
      bad = bad || min1 < 0 || min1 >= n;
    } else {
      const int nrec = s_nrec;
      bad = bad || nrec == 0;
      if (!bad) {
        const int win = s_rec[nrec - 1];
        min1 = (int)(uint32_t)D.row_key[win];
        min2 = D.row_p[win];
        halt = s_event != 0 || D.row_tie[win] != 0;
        // caches: replay the improvements in visiting order (with 4 or more records only the
        // last three matter, and they are distinct rows)
        int m = st.min1;
        for (int t = 0; t < nrec; ++t) {
          const int row = (int)(uint32_t)D.row_key[s_rec[t]];
          if (row != m) {
            c2 = c1;
            c1 = m;
            m = row;
          }
        }
      }
    }
    if (!bad && !halt) {
      bad = min2 < 0 || min2 >= n;
      if (!bad) {
        r1 = D.rtm[min1];
        r2 = D.rtm[min2];
        bad = r1 < 0 || r1 >= n || r2 < 0 || r2 >= n;
        if (!bad) {
          s1 = D.slot_of_index[r1];
          s2 = D.slot_of_index[r2];
          bad = s1 < 0 || s2 < 0;
        }
      }
    }
    s_fail = bad || halt;
    if (bad) {
      st.error = kNjFailed;
    } else if (halt) {
      st.error = kNjHalted;
    } else {
      st.forced = 0;
      st.min1 = min1;
      st.min2 = min2;
      st.cache1 = c1;
      st.cache2 = c2;
      // mergeMinNodes
      const float mutual = S[(size_t)s1 * D.ld + s2];
      const double dist = mutual;
      const double sep1 = D.sep[s1], sep2 = D.sep[s2];
      double dist1 = (0.5 * dist) + (0.5 * (sep1 - sep2));
      double dist2 = (0.5 * dist) + (0.5 * (sep2 - sep1));
      if (st.negative_branches) {
        if (dist1 < 0.0) {
          dist2 += dist1;
          dist1 = 0;
        }
        if (dist2 < 0.0) {
          dist1 += dist2;
          if (dist1 < 0) dist1 = 0;
          dist2 = 0;
        }
      }
      int new_row = r1, obs_row = r2;
      if (D.red_count[r1] != 1) {
        new_row = red_pop(D, r1);
        D.red_count[r1]--;
        D.index_to_id[new_row] = new_row;
      }
      const int m2 = D.rtm[min2];
      if (D.red_count[m2] != 1) {
        obs_row = red_pop(D, m2);
        D.red_count[m2]--;
        D.index_to_id[obs_row] = obs_row;
      }
      const int k = st.step;
      D.left[k] = D.index_to_id[new_row];
      D.right[k] = D.index_to_id[obs_row];
      D.llen[k] = dist1;
      D.rlen[k] = dist2;
      D.newidx[k] = new_row;
      st.new_row = new_row;
      st.obs_row = obs_row;
      st.cluster_count--;
      s_rows[0] = s1;
      s_rows[1] = s2;
      s_mutual = mutual;
      s_skip[0] = new_row;
      s_skip[1] = obs_row;
      s_rise = 0u;
    }
  }
  __syncthreads();
  if (s_fail) return;

  // updateData over every live cluster in id order (= live slots in order); the merged rows'
  // slots die (separation NaN)
  const float mutual = s_mutual;
  const float* row1 = S + (size_t)s_rows[0] * D.ld;
  const float* row2 = S + (size_t)s_rows[1] * D.ld;
  const float denom = (float)(st.cluster_count - 2);
  const int ns = used;
  float* new_row_p = S + (size_t)ns * D.ld;
  float* log_v = D.log_val + log_off;
  int* log_i = D.log_idx + log_off;
  // Four slots per thread and pass, their loads issued together (every index is in bounds for
  // dead slots too: slots below `used`, and slot_index always holds a matrix index).
  const int skip0 = s_skip[0], skip1 = s_skip[1];
  const bool neg = st.negative_branches != 0;
  float rise = 0.0f;
  for (int b0 = tid; b0 < used; b0 += 4 * kStepThreads) {
    int al[4], si[4], rcnt[4];
    float v1[4], v2[4], ssum[4], sold[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int b = b0 + u * kStepThreads;
      al[u] = 0;
      si[u] = 0;
      v1[u] = v2[u] = ssum[u] = sold[u] = 0.0f;
      if (b < used) {
        al[u] = D.alive[b];
        si[u] = D.slot_index[b];
        v1[u] = row1[b];
        v2[u] = row2[b];
        ssum[u] = D.sepsum[b];
        sold[u] = D.sep[b];
      }
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) rcnt[u] = b0 + u * kStepThreads < used ? D.red_count[si[u]] : 0;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int b = b0 + u * kStepThreads;
      if (b >= used) break;
      float dist = 0.0f, term = 0.0f;
      int partner = -1;
      if (al[u]) {
        const int i = si[u];
        if (i != skip0 && i != skip1) {
          const float val1 = v1[u], val2 = v2[u];
          const float t = __fsub_rn(__fadd_rn(val1, val2), mutual);
          dist = (float)((double)t / 2.0);
          if (neg && dist < 0) dist = 0;
          term = __fmul_rn(dist, (float)rcnt[u]);
          const float ss = __fadd_rn(ssum[u], __fsub_rn(__fsub_rn(dist, val1), val2));
          const float sn = ss / denom;
          D.sepsum[b] = ss;
          D.sep[b] = sn;
          rise = fmaxf(rise, __fsub_rn(sn, sold[u]));
          partner = i;
        } else {
          D.alive[b] = 0;
          D.sep[b] = __int_as_float(0x7fc00000);
          D.row_key[b] = kNoKey;
        }
      }
      D.newcol[b] = dist;
      D.term[b] = term;
      new_row_p[b] = dist;  // the column S[b][ns] is written by the next launch A
      log_v[b] = dist;
      log_i[b] = partner;
    }
  }
  rise = wave_fmax(rise);
  if ((tid & 63) == 0) atomicMax(&s_rise, __float_as_uint(rise));  // rise >= 0: ordered as bits
  __syncthreads();  // also makes the new separations visible to thread 0 below
  if (tid == 0) {
    new_row_p[ns] = 0.0f;
    const int new_row = st.new_row, obs_row = st.obs_row;
    D.slot_of_index[obs_row] = -1;
    D.key_by_index[obs_row] = kNone;
    D.id_to_index[D.index_to_id[new_row]] = -1;
    D.id_to_index[D.index_to_id[obs_row]] = -1;
    D.index_to_id[obs_row] = -1;
    D.rtm[new_row] = new_row;
    D.rtm[obs_row] = obs_row;
    const int next_id = st.current_id++;
    D.index_to_id[new_row] = next_id;
    D.id_to_index[next_id] = new_row;
    D.alive[ns] = 1;
    D.slot_index[ns] = new_row;
    D.slot_of_index[new_row] = ns;
    D.sep[ns] = __int_as_float(0x7fc00000);  // until launch A finishes it
    D.row_key[ns] = kNoKey;
    D.shave[ns] = 0;
    st.used = ns + 1;
    st.step++;
    // row skipping: the drift, and the next search's start row (rapidNJ.cpp:292-297) with the
    // criterion of the best pair of its last scan, if that partner is still a live slot below
    // it (one of its pairs, so at least its answer), and its copy join
    st.drift += (double)__uint_as_float(s_rise);
    const int start = D.index_to_id[st.cache1] != -1 ? st.cache1 : st.cache2;
    float U = FLT_MAX;
    int srow = -1;
    if (start >= 0 && start < n && D.index_to_id[start] != -1) {
      const int sa = D.slot_of_index[start];
      if (sa >= 0) {
        const int p = D.row_p[sa];
        const int sb = p >= 0 && p < n ? D.slot_of_index[p] : -1;
        if (sb >= 0 && sb < sa && D.alive[sb])
          U = __fsub_rn(__fsub_rn(S[(size_t)sa * D.ld + sb], D.sep[sb]), D.sep[sa]);
        if (D.red_count[start] > 1) U = fminf(U, -2.0f * D.sep[sa]);
        srow = U < FLT_MAX ? start : -1;
      }
    }
    st.start_bound = srow >= 0 ? U : FLT_MAX;
    st.start_row = srow;
  }
}

// ---- compaction ------------------------------------------------------------------------------
// Inclusive scan of part[0..1023] in place (one work-group of 1024).
__device__ __forceinline__ void scan1024(int* part) {
  const int tid = threadIdx.x;
  for (int off = 1; off < 1024; off <<= 1) {
    const int x = tid >= off ? part[tid - off] : 0;
    __syncthreads();
    part[tid] += x;
    __syncthreads();
  }
}

// Positions of the live slots (one work-group, in slot order) into D.map, count into st.alive.
__global__ __launch_bounds__(1024) void nj_compact_map_kernel(NjDevice D) {
  NjState& st = *D.st;
  __shared__ int part[1024];
  if (st.error) return;
  const int used = st.used, tid = threadIdx.x;
  const int per = (used + 1023) / 1024;
  const int b0 = min(used, tid * per), b1 = min(used, b0 + per);
  int cnt = 0;
  for (int b = b0; b < b1; ++b) cnt += D.alive[b] ? 1 : 0;
  part[tid] = cnt;
  __syncthreads();
  scan1024(part);
  int pos = part[tid] - cnt;
  for (int b = b0; b < b1; ++b)
    if (D.alive[b]) D.map[pos++] = b;
  if (tid == 1023) st.alive = part[1023];
}

// New slot matrix: S2[a][b] = S[map[a]][map[b]] for a, b < alive (grid: rows x column blocks).
__global__ __launch_bounds__(256) void nj_compact_copy_kernel(NjDevice D, const float* __restrict__ S,
                                                              float* __restrict__ S2) {
  if (D.st->error) return;
  const int alive = D.st->alive;
  for (int a = blockIdx.y; a < alive; a += gridDim.y) {
    const float* src = S + (size_t)D.map[a] * D.ld;
    float* dst = S2 + (size_t)a * D.ld;
    for (int b = blockIdx.x * 256 + threadIdx.x; b < alive; b += gridDim.x * 256)
      dst[b] = src[D.map[b]];
  }
}

// Per-slot arrays follow the new numbering (one work-group; in place, ascending: map[a] >= a).
__global__ __launch_bounds__(1024) void nj_compact_slots_kernel(NjDevice D) {
  NjState& st = *D.st;
  if (st.error) return;
  const int alive = st.alive, used = st.used;
  for (int base = 0; base < alive; base += 1024) {
    const int a = base + threadIdx.x;
    // with row skipping, a row's bound and its last answer (or skip key) must follow it: the
    // next scan may skip it, and the host reads them by slot when it settles a search
    int idx = 0, hv = 0, rp = -1, tie = 0;
    float ss = 0, s = 0, m = 0, mx = 0;
    double r = 0;
    unsigned long long key = kNoKey;
    unsigned sig = 0;
    if (a < alive) {
      const int b = D.map[a];
      idx = D.slot_index[b];
      ss = D.sepsum[b];
      s = D.sep[b];
      m = D.smin[b];
      r = D.sref[b];
      hv = D.shave[b];
      rp = D.row_p[b];
      key = D.row_key[b];
      sig = D.row_sigma[b];
      tie = D.row_tie[b];
      mx = D.row_maxsep[b];
    }
    __syncthreads();
    if (a < alive) {
      D.slot_index[a] = idx;
      D.sepsum[a] = ss;
      D.sep[a] = s;
      D.slot_of_index[idx] = a;
      D.smin[a] = m;  // a row's pairs are the same after the copy, so its bound holds
      D.sref[a] = r;
      D.shave[a] = hv;
      D.row_p[a] = rp;
      D.row_key[a] = key;
      D.row_sigma[a] = sig;
      D.row_tie[a] = tie;
      D.row_maxsep[a] = mx;
    }
    __syncthreads();
  }
  for (int b = threadIdx.x; b < used; b += 1024) D.alive[b] = b < alive ? 1 : 0;
  __syncthreads();
  if (threadIdx.x == 0) st.used = alive;
}

// ---- set-up ----------------------------------------------------------------------------------
// Row sums in index order, as rapidNJ::initialize accumulates them (float, j = 0..n-1).
__global__ void nj_row_sums_kernel(const float* __restrict__ M, size_t ldm, int n,
                                   float* __restrict__ sums) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float* row = M + (size_t)i * ldm;
  float s = 0.0f;
  for (int j = 0; j < n; ++j) s = __fadd_rn(s, row[j]);
  sums[i] = s;
}

// Pairs (i, j), j < i, with M[i][j] == 0 (rapidNJ.cpp:118-129): rows in parallel, entries of a
// row across the work-group.
__global__ __launch_bounds__(256) void nj_zero_pairs_kernel(const float* __restrict__ M, size_t ldm,
                                                            int n, int* __restrict__ pairs, int cap,
                                                            int* __restrict__ count) {
  for (int i = blockIdx.x; i < n; i += gridDim.x) {
    const float* ri = M + (size_t)i * ldm;
    for (int j = threadIdx.x; j < i; j += 256) {
      if (ri[j] != 0.0f) continue;
      const int at = atomicAdd(count, 1);
      if (at < cap) {
        pairs[2 * at] = i;
        pairs[2 * at + 1] = j;
      }
    }
  }
}

// handleRedundancy's test (rapidNJ.cpp:155-164): one work-group per listed pair.
__global__ __launch_bounds__(256) void nj_same_rows_kernel(const float* __restrict__ M, size_t ldm,
                                                           int n, const int* __restrict__ pairs,
                                                           int npairs, int* __restrict__ same) {
  __shared__ int diff;
  for (int k = blockIdx.x; k < npairs; k += gridDim.x) {
    if (threadIdx.x == 0) diff = 0;
    __syncthreads();
    const float* ri = M + (size_t)pairs[2 * k] * ldm;
    const float* rj = M + (size_t)pairs[2 * k + 1] * ldm;
    for (int c = threadIdx.x; c < n; c += 256)
      if (ri[c] != rj[c]) diff = 1;
    __syncthreads();
    if (threadIdx.x == 0) same[k] = diff ? 0 : 1;
    __syncthreads();
  }
}

// Initial slot matrix: S[a][b] = M[vis[a]][vis[b]].
__global__ __launch_bounds__(256) void nj_gather_kernel(const float* __restrict__ M, size_t ldm,
                                                        const int* __restrict__ vis, int v,
                                                        float* __restrict__ S, size_t ld) {
  for (int a = blockIdx.y; a < v; a += gridDim.y) {
    const float* src = M + (size_t)vis[a] * ldm;
    float* dst = S + (size_t)a * ld;
    for (int b = blockIdx.x * 256 + threadIdx.x; b < v; b += gridDim.x * 256) dst[b] = src[vis[b]];
  }
}

}  // namespace

// ---- launchers -----------------------------------------------------------------------------------
static unsigned gy(unsigned rows) { return rows < 1 ? 1 : (rows > 65535 ? 65535 : rows); }

hipError_t nj_launch_row_sums(const float* M, size_t ldm, int n, float* sums, hipStream_t s) {
  nj_row_sums_kernel<<<dim3((n + 255) / 256), dim3(256), 0, s>>>(M, ldm, n, sums);
  return hipGetLastError();
}

hipError_t nj_launch_zero_pairs(const float* M, size_t ldm, int n, int* pairs, int cap, int* count,
                                hipStream_t s) {
  nj_zero_pairs_kernel<<<dim3(n < 16384 ? n : 16384), dim3(256), 0, s>>>(M, ldm, n, pairs, cap,
                                                                          count);
  return hipGetLastError();
}

hipError_t nj_launch_same_rows(const float* M, size_t ldm, int n, const int* pairs, int npairs,
                               int* same, hipStream_t s) {
  if (npairs <= 0) return hipSuccess;
  nj_same_rows_kernel<<<dim3(npairs < 16384 ? npairs : 16384), dim3(256), 0, s>>>(M, ldm, n, pairs,
                                                                                  npairs, same);
  return hipGetLastError();
}

hipError_t nj_launch_gather(const float* M, size_t ldm, const int* vis, int v, float* S, size_t ld,
                            hipStream_t s) {
  if (v <= 0) return hipSuccess;
  nj_gather_kernel<<<dim3((v + 255) / 256 < 64 ? (v + 255) / 256 : 64, gy(v)), dim3(256), 0, s>>>(
      M, ldm, vis, v, S, ld);
  return hipGetLastError();
}

hipError_t nj_launch_iteration(const NjDevice& D, float* S, int used_bound, bool finish_previous,
                               size_t log_off, hipStream_t s) {
  const int rows_blocks = (used_bound + kScanRows - 1) / kScanRows;
  const int grid = rows_blocks + (finish_previous ? 1 : 0);
  if (grid > 0) {
    nj_scan_kernel<<<dim3(grid), dim3(kScanThreads), 0, s>>>(D, S, finish_previous ? 1 : 0);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  nj_step_kernel<<<dim3(1), dim3(kStepThreads), 0, s>>>(D, S, log_off);
  return hipGetLastError();
}

hipError_t nj_launch_step(const NjDevice& D, float* S, size_t log_off, hipStream_t s) {
  nj_step_kernel<<<dim3(1), dim3(kStepThreads), 0, s>>>(D, S, log_off);
  return hipGetLastError();
}

hipError_t nj_launch_finish(const NjDevice& D, float* S, hipStream_t s) {
  nj_finish_kernel<<<dim3(1), dim3(kScanThreads), 0, s>>>(D, S);
  return hipGetLastError();
}

hipError_t nj_launch_compact(const NjDevice& D, const float* S, float* S2, int alive_bound,
                             hipStream_t s) {
  nj_compact_map_kernel<<<dim3(1), dim3(1024), 0, s>>>(D);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const int bx = (alive_bound + 255) / 256 < 64 ? (alive_bound + 255) / 256 : 64;
  nj_compact_copy_kernel<<<dim3(bx < 1 ? 1 : bx, gy(alive_bound)), dim3(256), 0, s>>>(D, S, S2);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  nj_compact_slots_kernel<<<dim3(1), dim3(1024), 0, s>>>(D);
  return hipGetLastError();
}

}  // namespace rnj
