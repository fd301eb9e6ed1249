// Neighbour joining on the device: rapidNJ::run (rapidNJ.cpp:45-88) with the matrix in HBM.
//
// rapidNJ keeps each row's partners sorted by distance and stops scanning a row at a bound; the
// pair it joins is the first one, in its visiting order, reaching the minimum of the criterion
// (d - sep[partner]) - sep[row] over all live pairs. Here every live pair is evaluated each
// iteration (the scan is a full pass over the live matrix, which HBM and the 256 MB last-level
// cache serve at several TB/s), with the reference's float arithmetic and its visiting order:
//   * a pair (a, b) lives in the row of the cluster with the larger id; the matrix is kept in
//     "slots" ordered by cluster id (new clusters take the next slot), so row slot a's partners
//     are exactly the live slots b < a;
//   * rows are visited by matrix index, cyclically from the row rapidNJ cached in the previous
//     search, and a row's pairs by increasing distance, then increasing partner index (for
//     exact ties in both distance and criterion the reference's order is std::sort's);
//   * a row with redundant copies offers the copy join at -2 * sep first.
// Per iteration: nj_scan_kernel (16 row slots per work-group, one per wave: each row's best
// pair; it also writes the previous join's column) and nj_step_kernel (one work-group: the
// global winner and rapidNJ's row caches, the merge with its redundancy bookkeeping, the update
// of distances and separations in id order, the new slot's row). The new cluster's separation
// sum (a sequential float sum in id order) is finished by the next scan. Every few iterations
// nj_compact_* rebuilds the slot matrix from the live slots. The host enqueues all iterations
// without waiting: the iteration count (n - 2) and the slot bounds are known ahead.
#include "nj.hpp"

#include <float.h>

namespace rnj {
namespace {

constexpr int kScanThreads = 1024;  // 16 waves, one row each
constexpr int kScanRows = kScanThreads / 64;
constexpr int kSepChunk = 4096;     // separations staged in LDS per step of a row scan
constexpr int kStepThreads = 1024;
constexpr int kKeyCache = 20000;    // row keys the step keeps in LDS between its record passes

// Floats as order-preserving unsigned keys (after -0 -> +0, so that equal values compare equal
// as the reference's float comparisons do).
__device__ __forceinline__ uint32_t fkey(float x) {
  const uint32_t u = __float_as_uint(x + 0.0f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

// A row's candidate: criterion v, distance d, partner index p; ordered by (v, d, p) with NaN v
// (gone partners) never better.
struct Cand {
  float v, d;
  int p;
};

__device__ __forceinline__ bool cand_less(const Cand& x, const Cand& y) {
  if (x.v < y.v) return true;
  if (!(x.v == y.v)) return false;
  if (x.d < y.d) return true;
  if (!(x.d == y.d)) return false;
  return x.p < y.p;
}

__device__ __forceinline__ Cand wave_min(Cand c) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    Cand o{__shfl_xor(c.v, off), __shfl_xor(c.d, off), __shfl_xor(c.p, off)};
    if (cand_less(o, c)) c = o;
  }
  return c;
}

constexpr float kInf = __builtin_inff();
constexpr unsigned long long kNoKey = ~0ull;

// The row's answer as rapidNJ's findRowMin gives it: the redundant-copy join first when the row
// has copies (searchRedundantEntries), then the best pair unless it is not strictly better.
// Rows whose answer does not beat FLT_MAX (the search's start value) get no key.
__device__ __forceinline__ void write_row(const NjDevice& D, int a, const Cand& best, float s_a) {
  const int idx = D.slot_index[a];
  float v = best.p >= 0 ? best.v : kInf;
  int p = best.p;
  if (D.red_count[idx] > 1) {
    const float rv = -2.0f * s_a;
    if (!(best.p >= 0 && best.v < rv)) {
      v = rv;
      p = D.red_head[idx];
    }
  }
  D.row_key[a] = v < FLT_MAX ? ((unsigned long long)fkey(v) << 32) | (uint32_t)idx : kNoKey;
  D.row_p[a] = p;
}

// Candidate (d, separation) of partner b into this lane's best.
__device__ __forceinline__ void consider(const NjDevice& D, float d_raw, float s_b, float s_a,
                                         int b, Cand& best, int& best_b) {
  const float v = __fsub_rn(__fsub_rn(d_raw, s_b), s_a) + 0.0f;
  const float d = d_raw + 0.0f;
  if (v < best.v || (v == best.v && d < best.d)) {
    best.v = v;
    best.d = d;
    best_b = b;
  } else if (v == best.v && d == best.d && best_b >= 0 && D.slot_index[b] < D.slot_index[best_b]) {
    best_b = b;  // exact tie: the lower partner index first
  }
}

// Finish the previous join's new cluster (slot used - 1), on one work-group: the reference's
// sequential float sum of its separation sum in id order (terms of gone or merged slots are +0,
// which leaves a sum that is never -0 unchanged), its separation, and its row's best pair.
__device__ void finish_new_cluster(const NjDevice& D, const float* __restrict__ S) {
  NjState& st = *D.st;
  const int ns = st.used - 1;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  constexpr int kChunk = 4096;
  __shared__ float terms[kChunk];
  __shared__ Cand wbest[kScanThreads / 64];
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
  Cand best{kInf, kInf, -1};
  int best_b = -1;
  for (int b = tid; b < ns; b += kScanThreads) consider(D, D.newcol[b], D.sep[b], s_ns, b, best, best_b);
  best.p = best_b >= 0 ? D.slot_index[best_b] : 0x7fffffff;
  if (best_b < 0) best.v = kInf;
  best = wave_min(best);
  if (lane == 0) wbest[wave] = best;
  __syncthreads();
  if (tid == 0) {
    Cand b = wbest[0];
    for (int w = 1; w < kScanThreads / 64; ++w)
      if (cand_less(wbest[w], b)) b = wbest[w];
    if (!(b.v < kInf)) b.p = -1;
    write_row(D, ns, b, s_ns);
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
  const int reach = min(a0 + kScanRows, used) - 1;  // longest row prefix in this group
  const float s_a = mine ? D.sep[a] : 0.0f;
  const float* row = S + (size_t)a * D.ld;
  Cand best{kInf, kInf, -1};
  int best_b = -1;
  for (int c0 = 0; c0 < reach; c0 += kSepChunk) {
    const int cnt = min(kSepChunk, reach - c0);
    for (int k = 4 * tid; k < cnt; k += 4 * kScanThreads)
      *reinterpret_cast<float4*>(ssep + k) = *reinterpret_cast<const float4*>(D.sep + c0 + k);
    __syncthreads();
    if (mine) {
      // four 4-wide loads in flight per lane; a lane's candidates are the same as one at a time
      // (ties are settled by partner index, not by visiting order)
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
          consider(D, dv[u].x, sv.x, s_a, bu, best, best_b);
          if (bu + 1 < end) consider(D, dv[u].y, sv.y, s_a, bu + 1, best, best_b);
          if (bu + 2 < end) consider(D, dv[u].z, sv.z, s_a, bu + 2, best, best_b);
          if (bu + 3 < end) consider(D, dv[u].w, sv.w, s_a, bu + 3, best, best_b);
        }
      }
    }
    __syncthreads();
  }
  if (!mine) return;
  best.p = best_b >= 0 ? D.slot_index[best_b] : 0x7fffffff;
  if (best_b < 0) best.v = kInf;
  best = wave_min(best);
  if (lane == 0) {
    if (!(best.v < kInf)) best.p = -1;
    write_row(D, a, best, s_a);
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
struct RowKey {  // (v, cyc) lexicographic; slot -1 = none
  uint32_t v;
  int cyc, slot;
};

__device__ __forceinline__ bool rk_less(const RowKey& x, const RowKey& y) {
  if (x.slot < 0) return false;
  if (y.slot < 0) return true;
  if (x.v != y.v) return x.v < y.v;
  return x.cyc < y.cyc;
}

// Fold the key of slot b into k when it has one and its cyclic position is below `below`.
__device__ __forceinline__ void fold_key(unsigned long long key, int b, int start, int n, int below,
                                         RowKey& k) {
  if (key == kNoKey) return;
  int cyc = (int)(uint32_t)key - start;
  if (cyc < 0) cyc += n;
  if (cyc >= below) return;
  const RowKey c{(uint32_t)(key >> 32), cyc, b};
  if (rk_less(c, k)) k = c;
}

// Wave-wide minimum (every lane gets it).
__device__ __forceinline__ RowKey wave_min_key(RowKey k) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    RowKey o{(uint32_t)__shfl_xor((int)k.v, off), __shfl_xor(k.cyc, off), __shfl_xor(k.slot, off)};
    if (rk_less(o, k)) k = o;
  }
  return k;
}

// Block-wide minimum: shuffles within waves, then one pass over the 16 wave results.
__device__ RowKey block_min(RowKey k, RowKey* sh) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  k = wave_min_key(k);
  if (lane == 0) sh[wave] = k;
  __syncthreads();
  RowKey r = sh[0];
  for (int w = 1; w < kStepThreads / 64; ++w)
    if (rk_less(sh[w], r)) r = sh[w];
  __syncthreads();
  return r;
}

// The first row (in visiting order) reaching the minimum among rows with a key (live, answer
// below FLT_MAX) whose cyclic position is below `below`. The first pass (fill) copies the keys of
// the first kKeyCache slots into `cache`; later passes read them from there (block_min's
// barriers order the copy before them).
__device__ RowKey best_row_before(const NjDevice& D, int used, int start, int n, int below,
                                  RowKey* sh, unsigned long long* cache, bool fill) {
  RowKey k{0xffffffffu, 0x7fffffff, -1};
  // four keys per thread and pass, loaded together
  for (int b0 = threadIdx.x; b0 < used; b0 += 4 * kStepThreads) {
    unsigned long long key[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int b = b0 + u * kStepThreads;
      key[u] = kNoKey;
      if (b < used) key[u] = (fill || b >= kKeyCache) ? D.row_key[b] : cache[b];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int b = b0 + u * kStepThreads;
      if (fill && b < used && b < kKeyCache) cache[b] = key[u];
      fold_key(key[u], b, start, n, below, k);
    }
  }
  return block_min(k, sh);
}

__device__ __forceinline__ int red_pop(const NjDevice& D, int i) {
  const int j = D.red_head[i];
  D.red_head[i] = D.red_next[j];
  return j;
}

// Launch B, one work-group: findMin's result and caches (rapidNJ.cpp:290-345), mergeMinNodes
// (:240-281), clusterCount--, updateData (:166-238) except the new separation sum, which the
// next launch A finishes.
__global__ __launch_bounds__(kStepThreads) void nj_step_kernel(NjDevice D, float* __restrict__ S) {
  NjState& st = *D.st;
  const int tid = threadIdx.x;
  __shared__ RowKey sh[kStepThreads / 64];
  __shared__ unsigned long long key_cache[kKeyCache];
  __shared__ int s_rows[2], s_skip[2], s_fail;
  __shared__ float s_mutual;
  if (st.error) return;
  const int n = st.n, used = st.used;
  const int start = D.index_to_id[st.cache1] != -1 ? st.cache1 : st.cache2;
  // records: rows where the running minimum improved, last first
  // (unrolled with constant indices, so that rec stays in registers; nrec is uniform)
  RowKey rec[4];
  int nrec = 0;
  int below = n + 1;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    if (nrec == r) {
      rec[r] = best_row_before(D, used, start, n, below, sh, key_cache, r == 0);
      if (rec[r].slot >= 0) {
        below = rec[r].cyc;
        nrec = r + 1;
      }
    }
  }
  if (tid == 0) {
    // the winner must name live rows with slots, and the new cluster needs a free slot
    bool bad = nrec == 0 || used >= (int)D.ld;
    int min1 = 0, min2 = 0, r1 = 0, r2 = 0, s1 = 0, s2 = 0;
    if (!bad) {
      min1 = (int)(uint32_t)D.row_key[rec[0].slot];
      min2 = D.row_p[rec[0].slot];
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
    s_fail = bad;
    if (bad) {
      st.error = 1;
    } else {
      // caches: replay the improvements in visiting order (records are listed last first; with
      // 4 or more records only the last three matter, and they are distinct rows)
      int m = st.min1, c1 = st.cache1, c2 = st.cache2;
#pragma unroll
      for (int t = 3; t >= 0; --t) {
        if (t >= nrec) continue;
        const int row = (int)(uint32_t)D.row_key[rec[t].slot];
        if (row != m) {
          c2 = c1;
          c1 = m;
          m = row;
        }
      }
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
      st.new_row = new_row;
      st.obs_row = obs_row;
      st.cluster_count--;
      s_rows[0] = s1;
      s_rows[1] = s2;
      s_mutual = mutual;
      s_skip[0] = new_row;
      s_skip[1] = obs_row;
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
  // Four slots per thread and pass, their loads issued together (every index is in bounds for
  // dead slots too: slots below `used`, and slot_index always holds a matrix index).
  const int skip0 = s_skip[0], skip1 = s_skip[1];
  const bool neg = st.negative_branches != 0;
  for (int b0 = tid; b0 < used; b0 += 4 * kStepThreads) {
    int al[4], si[4], rcnt[4];
    float v1[4], v2[4], ssum[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int b = b0 + u * kStepThreads;
      al[u] = 0;
      si[u] = 0;
      v1[u] = v2[u] = ssum[u] = 0.0f;
      if (b < used) {
        al[u] = D.alive[b];
        si[u] = D.slot_index[b];
        v1[u] = row1[b];
        v2[u] = row2[b];
        ssum[u] = D.sepsum[b];
      }
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) rcnt[u] = b0 + u * kStepThreads < used ? D.red_count[si[u]] : 0;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int b = b0 + u * kStepThreads;
      if (b >= used) break;
      float dist = 0.0f, term = 0.0f;
      if (al[u]) {
        const int i = si[u];
        if (i != skip0 && i != skip1) {
          const float val1 = v1[u], val2 = v2[u];
          const float t = __fsub_rn(__fadd_rn(val1, val2), mutual);
          dist = (float)((double)t / 2.0);
          if (neg && dist < 0) dist = 0;
          term = __fmul_rn(dist, (float)rcnt[u]);
          const float ss = __fadd_rn(ssum[u], __fsub_rn(__fsub_rn(dist, val1), val2));
          D.sepsum[b] = ss;
          D.sep[b] = ss / denom;
        } else {
          D.alive[b] = 0;
          D.sep[b] = __int_as_float(0x7fc00000);
          D.row_key[b] = kNoKey;
        }
      }
      D.newcol[b] = dist;
      D.term[b] = term;
      new_row_p[b] = dist;  // the column S[b][ns] is written by the next launch A
    }
  }
  if (tid == 0) {
    new_row_p[ns] = 0.0f;
    const int new_row = st.new_row, obs_row = st.obs_row;
    D.slot_of_index[obs_row] = -1;
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
    st.used = ns + 1;
    st.step++;
  }
}

// ---- compaction ------------------------------------------------------------------------------
// Positions of the live slots (one work-group, in slot order) into D.map, count into st.alive.
__global__ __launch_bounds__(1024) void nj_compact_map_kernel(NjDevice D) {
  NjState& st = *D.st;
  __shared__ int part[1024];
  const int used = st.used, tid = threadIdx.x;
  const int per = (used + 1023) / 1024;
  const int b0 = tid * per, b1 = min(used, b0 + per);
  int cnt = 0;
  for (int b = b0; b < b1; ++b) cnt += D.alive[b] ? 1 : 0;
  part[tid] = cnt;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {  // inclusive scan
    const int x = tid >= off ? part[tid - off] : 0;
    __syncthreads();
    part[tid] += x;
    __syncthreads();
  }
  int pos = part[tid] - cnt;
  for (int b = b0; b < b1; ++b)
    if (D.alive[b]) D.map[pos++] = b;
  if (tid == 1023) st.alive = part[1023];
}

// New slot matrix: S2[a][b] = S[map[a]][map[b]] for a, b < alive (grid: rows x column blocks).
__global__ __launch_bounds__(256) void nj_compact_copy_kernel(NjDevice D, const float* __restrict__ S,
                                                              float* __restrict__ S2) {
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
  const int alive = st.alive, used = st.used;
  for (int base = 0; base < alive; base += 1024) {
    const int a = base + threadIdx.x;
    int idx = 0;
    float ss = 0, s = 0;
    if (a < alive) {
      const int b = D.map[a];
      idx = D.slot_index[b];
      ss = D.sepsum[b];
      s = D.sep[b];
    }
    __syncthreads();
    if (a < alive) {
      D.slot_index[a] = idx;
      D.sepsum[a] = ss;
      D.sep[a] = s;
      D.slot_of_index[idx] = a;
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
                               hipStream_t s) {
  const int rows_blocks = (used_bound + kScanRows - 1) / kScanRows;
  const int grid = rows_blocks + (finish_previous ? 1 : 0);
  if (grid > 0) {
    nj_scan_kernel<<<dim3(grid), dim3(kScanThreads), 0, s>>>(D, S, finish_previous ? 1 : 0);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  nj_step_kernel<<<dim3(1), dim3(kStepThreads), 0, s>>>(D, S);
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
