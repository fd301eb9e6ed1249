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
// Per iteration: nj_scan_kernel (one work-group per row slot: the row's best pair) and
// nj_step_kernel (one work-group: the global winner and rapidNJ's row caches, the merge with its
// redundancy bookkeeping, the update of distances and separations in id order including the
// sequential float sum of the new separation sum, the new slot's row and column). Every few
// iterations nj_compact_* rebuilds the slot matrix from the live slots. The host enqueues all
// iterations without waiting: the iteration count (n - 2) and the slot bounds are known ahead.
#include "nj.hpp"

#include <float.h>

namespace rnj {
namespace {

constexpr int kScanThreads = 256;
constexpr int kStepThreads = 1024;

// Floats as order-preserving unsigned keys (after -0 -> +0, so that equal values compare equal
// as the reference's float comparisons do).
__device__ __forceinline__ uint32_t fkey(float x) {
  const uint32_t u = __float_as_uint(x + 0.0f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

struct Cand {  // lexicographic (v, d, p)
  uint32_t v, d;
  int p;
};

__device__ __forceinline__ bool cand_less(const Cand& x, const Cand& y) {
  if (x.v != y.v) return x.v < y.v;
  if (x.d != y.d) return x.d < y.d;
  return x.p < y.p;
}

// Best pair of every live row slot: the row's first candidate reaching its minimum, in the
// reference's order within a row. row_v = +inf marks a row without candidates.
__global__ __launch_bounds__(kScanThreads) void nj_scan_kernel(NjDevice D, const float* __restrict__ S) {
  const NjState& st = *D.st;
  const int a = blockIdx.x;
  if (a >= st.used || st.error) return;
  const int tid = threadIdx.x;
  if (!D.alive[a]) {
    if (tid == 0) D.row_v[a] = __int_as_float(0x7f800000);
    return;
  }
  const float* row = S + (size_t)a * D.ld;
  const float s_a = D.sep[a];
  Cand best{0xffffffffu, 0xffffffffu, 0x7fffffff};
  float best_v = __int_as_float(0x7f800000);
  for (int b = tid; b < a; b += kScanThreads) {
    if (!D.alive[b]) continue;
    const float d = row[b];
    const float v = __fsub_rn(__fsub_rn(d, D.sep[b]), s_a);
    const Cand c{fkey(v), fkey(d), D.slot_index[b]};
    if (cand_less(c, best)) {
      best = c;
      best_v = v;
    }
  }
  __shared__ Cand sc[kScanThreads];
  __shared__ float sv[kScanThreads];
  sc[tid] = best;
  sv[tid] = best_v;
  __syncthreads();
  for (int off = kScanThreads / 2; off > 0; off >>= 1) {
    if (tid < off && cand_less(sc[tid + off], sc[tid])) {
      sc[tid] = sc[tid + off];
      sv[tid] = sv[tid + off];
    }
    __syncthreads();
  }
  if (tid == 0) {
    const int idx = D.slot_index[a];
    float v = sv[0];
    int p = sc[0].p;
    int red = 0;
    if (D.red_count[idx] > 1) {  // searchRedundantEntries comes first in the row
      const float rv = -2.0f * s_a;
      if (!(sc[0].v != 0xffffffffu && v < rv)) {
        v = rv;
        p = D.red_head[idx];
        red = 1;
      }
    }
    D.row_v[a] = sc[0].v == 0xffffffffu && !red ? __int_as_float(0x7f800000) : v;
    D.row_p[a] = p;
  }
}

// ---- block reductions for the one-work-group step ----------------------------------------------
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

__device__ RowKey block_min(RowKey k, RowKey* sh) {
  const int tid = threadIdx.x;
  sh[tid] = k;
  __syncthreads();
  for (int off = kStepThreads / 2; off > 0; off >>= 1) {
    if (tid < off && rk_less(sh[tid + off], sh[tid])) sh[tid] = sh[tid + off];
    __syncthreads();
  }
  const RowKey r = sh[0];
  __syncthreads();
  return r;
}

// The first row (in visiting order) reaching the minimum among live rows with candidates whose
// cyclic position is below `below`.
__device__ RowKey best_row_before(const NjDevice& D, const NjState& st, int start, int below,
                                  RowKey* sh) {
  RowKey k{0xffffffffu, 0x7fffffff, -1};
  for (int b = threadIdx.x; b < st.used; b += kStepThreads) {
    if (!D.alive[b]) continue;
    const float v = D.row_v[b];
    if (!(v < FLT_MAX)) continue;  // no candidate, or none that beats the initial FLT_MAX
    const int idx = D.slot_index[b];
    const int cyc = (idx - start + st.n) % st.n;
    if (cyc >= below) continue;
    const RowKey c{fkey(v), cyc, b};
    if (rk_less(c, k)) k = c;
  }
  return block_min(k, sh);
}

__device__ __forceinline__ int red_pop(const NjDevice& D, int i) {
  const int j = D.red_head[i];
  D.red_head[i] = D.red_next[j];
  return j;
}

// One iteration after the scan: findMin's result and caches (rapidNJ.cpp:290-345),
// mergeMinNodes (:240-281), clusterCount--, updateData (:166-238).
__global__ __launch_bounds__(kStepThreads) void nj_step_kernel(NjDevice D, float* __restrict__ S) {
  NjState& st = *D.st;
  const int tid = threadIdx.x;
  __shared__ RowKey sh[kStepThreads];
  __shared__ int s_new_slot, s_rows[2], s_skip[2];
  __shared__ float s_mutual, s_newsum;
  __shared__ int s_fail;
  if (st.error) return;

  const int start = D.index_to_id[st.cache1] != -1 ? st.cache1 : st.cache2;
  // records: rows where the running minimum improved, last first
  RowKey rec[4];
  int nrec = 0;
  int below = st.n + 1;
  for (; nrec < 4; ++nrec) {
    rec[nrec] = best_row_before(D, st, start, below, sh);
    if (rec[nrec].slot < 0) break;
    below = rec[nrec].cyc;
  }
  if (tid == 0) {
    // the winner must name live rows with slots, and the new cluster needs a free slot
    bool bad = nrec == 0 || st.used >= (int)D.ld;
    if (!bad) {
      const int w1 = D.slot_index[rec[0].slot], w2 = D.row_p[rec[0].slot];
      bad = w2 < 0 || w2 >= st.n;
      if (!bad) {
        const int q1 = D.rtm[w1], q2 = D.rtm[w2];
        bad = q1 < 0 || q1 >= st.n || q2 < 0 || q2 >= st.n || D.slot_of_index[q1] < 0 ||
              D.slot_of_index[q2] < 0;
      }
    }
    s_fail = bad;
    if (bad) st.error = 1;
  }
  __syncthreads();
  if (s_fail) return;
  if (tid == 0) {
    // caches: replay the improvements in visiting order (records are listed last first)
    int m = st.min1, c1 = st.cache1, c2 = st.cache2;
    for (int t = nrec - 1; t >= 0; --t) {
      const int row = D.slot_index[rec[t].slot];
      if (row != m) {
        c2 = c1;
        c1 = m;
        m = row;
      }
    }
    // with 4 or more records only the last three matter, and they are distinct rows
    const int wslot = rec[0].slot;
    const int min1 = D.slot_index[wslot];
    const int min2 = D.row_p[wslot];
    st.min1 = min1;
    st.min2 = min2;
    st.cache1 = c1;
    st.cache2 = c2;
    (void)m;
    // mergeMinNodes
    const int r1 = D.rtm[min1], r2 = D.rtm[min2];
    const int s1 = D.slot_of_index[r1], s2 = D.slot_of_index[r2];
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
    s_new_slot = st.used;
  }
  __syncthreads();

  // updateData: every live cluster in id order (= live slots in order)
  const int used = st.used;
  const float mutual = s_mutual;
  const float* row1 = S + (size_t)s_rows[0] * D.ld;
  const float* row2 = S + (size_t)s_rows[1] * D.ld;
  const float denom = (float)(st.cluster_count - 2);
  for (int b = tid; b < used; b += kStepThreads) {
    float dist = 0.0f, term = 0.0f;
    int use = 0;
    if (D.alive[b]) {
      const int i = D.slot_index[b];
      if (i != s_skip[0] && i != s_skip[1]) {
        const float val1 = row1[b], val2 = row2[b];
        const float t = __fsub_rn(__fadd_rn(val1, val2), mutual);
        dist = (float)((double)t / 2.0);
        if (st.negative_branches && dist < 0) dist = 0;
        term = __fmul_rn(dist, (float)D.red_count[i]);
        D.sepsum[b] = __fadd_rn(D.sepsum[b], __fsub_rn(__fsub_rn(dist, val1), val2));
        D.sep[b] = D.sepsum[b] / denom;
        use = 1;
      }
    }
    D.newcol[b] = dist;
    D.term[b] = use ? term : __int_as_float(0x7fc00000);  // NaN marks "not added"
  }
  __syncthreads();
  // newSeparationsum: the reference's sequential float sum, in id order, on one wave
  if (tid < 64) {
    float acc = 0.0f;
    for (int base = 0; base < used; base += 64) {
      const int b = base + tid;
      const float t = b < used ? D.term[b] : __int_as_float(0x7fc00000);
      const unsigned long long present = __ballot(t == t);
      const int ti = __float_as_int(t);
#pragma unroll
      for (int l = 0; l < 64; ++l) {
        const float x = __int_as_float(__builtin_amdgcn_readlane(ti, l));
        if ((present >> l) & 1ull) acc = __fadd_rn(acc, x);
      }
    }
    if (tid == 0) s_newsum = acc;
  }
  __syncthreads();
  // the new cluster: slot `used`, its row and column; the merged rows' slots die
  const int ns = s_new_slot;
  float* new_row_p = S + (size_t)ns * D.ld;
  for (int b = tid; b < used; b += kStepThreads) {
    const float x = D.alive[b] ? D.newcol[b] : 0.0f;
    new_row_p[b] = x;
    S[(size_t)b * D.ld + ns] = x;
  }
  if (tid == 0) {
    new_row_p[ns] = 0.0f;
    const int new_row = st.new_row, obs_row = st.obs_row;
    const int old_new_slot = D.slot_of_index[new_row];
    if (old_new_slot >= 0) D.alive[old_new_slot] = 0;
    const int obs_slot = D.slot_of_index[obs_row];
    if (obs_slot >= 0) D.alive[obs_slot] = 0;
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
    D.sepsum[ns] = s_newsum;
    D.sep[ns] = s_newsum / denom;
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

hipError_t nj_launch_iteration(const NjDevice& D, float* S, int used_bound, hipStream_t s) {
  nj_scan_kernel<<<dim3(used_bound), dim3(kScanThreads), 0, s>>>(D, S);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  nj_step_kernel<<<dim3(1), dim3(kStepThreads), 0, s>>>(D, S);
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
