// Host driver of the device neighbour joining (nj.hip): set-up as rapidNJ::initialize
// (rapidNJ.cpp:91-138), the iteration loop of rapidNJ::run (:45-50) enqueued without waits, the
// last join (:52-85), and the tree with polytree's Newick text (polytree.cpp:402-458).
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "nj.hpp"
#include "runtime.hpp"

struct rnj_tree {
  unsigned n = 0;
  std::vector<int> left, right;  // polytree::addInternalNode: the two ids joined at step k
  std::vector<double> llen, rlen;
  int last[2] = {0, 0};          // polytree::set_serialization_indices
  float last_dist = 0.0f;
  std::vector<std::string> names;
};

namespace rnj {
namespace {

template <class T>
struct DevBuf {
  T* p = nullptr;
  hipError_t alloc(size_t count) { return hipMalloc((void**)&p, std::max<size_t>(count, 1) * sizeof(T)); }
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
};

#define NJ_HIP(call)                                  \
  do {                                                \
    hipError_t _e = (call);                           \
    if (_e != hipSuccess) return hip_fail(_e, #call); \
  } while (0)

int run_nj(Context* ctx, const float* M, size_t ldm, unsigned n, int negative_branches,
           hipStream_t s, rnj_tree* t) {
  // slots; a multiple of 64 so that every row starts 256-byte aligned and 4-wide loads stay in
  // their row
  const size_t cap = ((size_t)n + std::max<size_t>(64, n / 2) + 63) / 64 * 64;
  // device state
  DevBuf<NjState> st;
  DevBuf<int> index_to_id, id_to_index, red_count, rtm, red_head, red_next, slot_of_index;
  DevBuf<int> slot_index, alive, row_p, map, left, right;
  DevBuf<float> sepsum, sep, newcol, term, sums, S0, S1;
  DevBuf<unsigned long long> row_key;
  DevBuf<double> llen, rlen;
  NJ_HIP(st.alloc(1));
  NJ_HIP(index_to_id.alloc(n));
  NJ_HIP(id_to_index.alloc(2 * (size_t)n));
  NJ_HIP(red_count.alloc(n));
  NJ_HIP(rtm.alloc(n));
  NJ_HIP(red_head.alloc(n));
  NJ_HIP(red_next.alloc(n));
  NJ_HIP(slot_of_index.alloc(n));
  NJ_HIP(slot_index.alloc(cap));
  NJ_HIP(alive.alloc(cap));
  NJ_HIP(row_p.alloc(cap));
  NJ_HIP(map.alloc(cap));
  NJ_HIP(left.alloc(n));
  NJ_HIP(right.alloc(n));
  NJ_HIP(sepsum.alloc(cap));
  NJ_HIP(sep.alloc(cap));
  NJ_HIP(row_key.alloc(cap));
  NJ_HIP(newcol.alloc(cap));
  NJ_HIP(term.alloc(cap));
  NJ_HIP(sums.alloc(n));
  NJ_HIP(llen.alloc(n));
  NJ_HIP(rlen.alloc(n));
  NJ_HIP(S0.alloc(cap * cap));
  NJ_HIP(S1.alloc(cap * cap));
  // declared after the buffers, so destroyed first: no buffer is freed under a running kernel
  struct StreamWait {
    hipStream_t s;
    ~StreamWait() { (void)hipStreamSynchronize(s); }
  } wait_on_exit{s};

  // rapidNJ::initialize: row sums, redundant rows
  NJ_HIP(nj_launch_row_sums(M, ldm, (int)n, sums.p, s));
  std::vector<float> h_sums(n);
  NJ_HIP(hipMemcpyAsync(h_sums.data(), sums.p, n * sizeof(float), hipMemcpyDeviceToHost, s));
  int npairs = 0;
  std::vector<int> pairs, same;
  {
    DevBuf<int> count;
    NJ_HIP(count.alloc(1));
    int cap_pairs = (int)std::min<size_t>(4 * (size_t)n + 1024, 1u << 26);
    for (int pass = 0; pass < 2; ++pass) {
      DevBuf<int> d_pairs;
      NJ_HIP(d_pairs.alloc(2 * (size_t)cap_pairs));
      NJ_HIP(hipMemsetAsync(count.p, 0, sizeof(int), s));
      NJ_HIP(nj_launch_zero_pairs(M, ldm, (int)n, d_pairs.p, cap_pairs, count.p, s));
      NJ_HIP(hipMemcpyAsync(&npairs, count.p, sizeof(int), hipMemcpyDeviceToHost, s));
      NJ_HIP(hipStreamSynchronize(s));
      if (npairs > cap_pairs) {
        if (pass == 1) return fail(RNJ_E_INVALID, "too many zero distances for the redundancy check");
        cap_pairs = npairs;
        continue;
      }
      if (npairs > 0) {
        DevBuf<int> d_same;
        NJ_HIP(d_same.alloc(npairs));
        NJ_HIP(nj_launch_same_rows(M, ldm, (int)n, d_pairs.p, npairs, d_same.p, s));
        pairs.resize(2 * (size_t)npairs);
        same.resize(npairs);
        NJ_HIP(hipMemcpyAsync(pairs.data(), d_pairs.p, pairs.size() * sizeof(int),
                              hipMemcpyDeviceToHost, s));
        NJ_HIP(hipMemcpyAsync(same.data(), d_same.p, same.size() * sizeof(int),
                              hipMemcpyDeviceToHost, s));
        NJ_HIP(hipStreamSynchronize(s));
      }
      break;
    }
  }
  // handleRedundancy in the reference's order: i from n-1 down (rows not already hidden), j up
  std::vector<int> h_index_to_id(n), h_rtm(n), h_red_count(n, 1), h_red_head(n, -1),
      h_red_next(n, -1);
  for (unsigned i = 0; i < n; ++i) {
    h_index_to_id[i] = (int)i;
    h_rtm[i] = (int)i;
  }
  {
    std::vector<std::vector<int>> ident(n);  // identical rows j < i of row i
    for (int k = 0; k < npairs; ++k)
      if (same[k]) ident[pairs[2 * k]].push_back(pairs[2 * k + 1]);
    for (int i = (int)n - 1; i >= 0; --i) {
      if (h_index_to_id[i] == -1) continue;
      std::sort(ident[i].begin(), ident[i].end());
      for (int j : ident[i]) {
        h_index_to_id[j] = -1;
        h_rtm[j] = i;
        h_red_next[j] = h_red_head[i];
        h_red_head[i] = j;
        h_red_count[i]++;
      }
    }
  }
  // slots: the visible rows in index (= id) order
  std::vector<int> vis, h_slot_of_index(n, -1), h_id_to_index(2 * (size_t)n, -1);
  for (unsigned i = 0; i < n; ++i) {
    h_id_to_index[i] = (int)i;
    if (h_index_to_id[i] != -1) {
      h_slot_of_index[i] = (int)vis.size();
      vis.push_back((int)i);
    }
  }
  const int v = (int)vis.size();
  std::vector<float> h_sepsum(cap, 0.0f), h_sep(cap, 0.0f);
  std::vector<int> h_alive(cap, 0), h_slot_index(cap, 0);
  for (int a = 0; a < v; ++a) {
    h_sepsum[a] = h_sums[vis[a]];
    h_sep[a] = h_sums[vis[a]] / (float)((int)n - 2);  // separation = sum / (clusterCount - 2)
    h_alive[a] = 1;
    h_slot_index[a] = vis[a];
  }
  NjState h_st{};
  h_st.n = (int)n;
  h_st.used = v;
  h_st.cluster_count = (int)n;
  h_st.current_id = (int)n;
  h_st.negative_branches = negative_branches ? 1 : 0;
  auto up = [&](void* dst, const void* src, size_t bytes) {
    return hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s);
  };
  DevBuf<int> d_vis;
  NJ_HIP(d_vis.alloc(std::max(v, 1)));
  NJ_HIP(up(d_vis.p, vis.data(), v * sizeof(int)));
  NJ_HIP(up(st.p, &h_st, sizeof h_st));
  NJ_HIP(up(index_to_id.p, h_index_to_id.data(), n * sizeof(int)));
  NJ_HIP(up(id_to_index.p, h_id_to_index.data(), 2 * (size_t)n * sizeof(int)));
  NJ_HIP(up(red_count.p, h_red_count.data(), n * sizeof(int)));
  NJ_HIP(up(rtm.p, h_rtm.data(), n * sizeof(int)));
  NJ_HIP(up(red_head.p, h_red_head.data(), n * sizeof(int)));
  NJ_HIP(up(red_next.p, h_red_next.data(), n * sizeof(int)));
  NJ_HIP(up(slot_of_index.p, h_slot_of_index.data(), n * sizeof(int)));
  NJ_HIP(up(slot_index.p, h_slot_index.data(), cap * sizeof(int)));
  NJ_HIP(up(alive.p, h_alive.data(), cap * sizeof(int)));
  NJ_HIP(up(sepsum.p, h_sepsum.data(), cap * sizeof(float)));
  NJ_HIP(up(sep.p, h_sep.data(), cap * sizeof(float)));
  NJ_HIP(nj_launch_gather(M, ldm, d_vis.p, v, S0.p, cap, s));

  NjDevice D{};
  D.st = st.p;
  D.ld = cap;
  D.index_to_id = index_to_id.p;
  D.id_to_index = id_to_index.p;
  D.red_count = red_count.p;
  D.rtm = rtm.p;
  D.red_head = red_head.p;
  D.red_next = red_next.p;
  D.slot_of_index = slot_of_index.p;
  D.slot_index = slot_index.p;
  D.alive = alive.p;
  D.sepsum = sepsum.p;
  D.sep = sep.p;
  D.row_key = row_key.p;
  D.row_p = row_p.p;
  D.newcol = newcol.p;
  D.term = term.p;
  D.map = map.p;
  D.left = left.p;
  D.right = right.p;
  D.llen = llen.p;
  D.rlen = rlen.p;

  // rapidNJ::run: n - 2 joins. used_bound bounds the slots in use (exact until the first
  // compaction, then the live count's bound clusterCount plus the slots added since). The slot
  // matrix is compacted when that bound passes 1.125 x clusterCount (or the capacity).
  float* S = S0.p;
  float* S2 = S1.p;
  size_t used_bound = (size_t)v;
  bool pending = false;  // the last join's new cluster still needs its separation and row
  for (unsigned it = 0; it + 2 < n; ++it) {
    const size_t cc = n - it;  // clusterCount before this join
    if (used_bound + 1 > cap || (used_bound > 256 && used_bound * 8 > cc * 9)) {
      if (pending) NJ_HIP(nj_launch_finish(D, S, s));
      pending = false;
      NJ_HIP(nj_launch_compact(D, S, S2, (int)std::min(cc, used_bound), s));
      std::swap(S, S2);
      used_bound = std::min(cc, used_bound);
    }
    NJ_HIP(nj_launch_iteration(D, S, (int)used_bound, pending, s));
    pending = true;
    ++used_bound;
  }
  // results
  std::vector<int> hl(n), hr(n);
  std::vector<double> hll(n), hrl(n);
  NJ_HIP(hipMemcpyAsync(&h_st, st.p, sizeof h_st, hipMemcpyDeviceToHost, s));
  NJ_HIP(hipMemcpyAsync(hl.data(), left.p, n * sizeof(int), hipMemcpyDeviceToHost, s));
  NJ_HIP(hipMemcpyAsync(hr.data(), right.p, n * sizeof(int), hipMemcpyDeviceToHost, s));
  NJ_HIP(hipMemcpyAsync(hll.data(), llen.p, n * sizeof(double), hipMemcpyDeviceToHost, s));
  NJ_HIP(hipMemcpyAsync(hrl.data(), rlen.p, n * sizeof(double), hipMemcpyDeviceToHost, s));
  NJ_HIP(hipMemcpyAsync(h_index_to_id.data(), index_to_id.p, n * sizeof(int), hipMemcpyDeviceToHost, s));
  NJ_HIP(hipMemcpyAsync(h_red_count.data(), red_count.p, n * sizeof(int), hipMemcpyDeviceToHost, s));
  NJ_HIP(hipMemcpyAsync(h_red_head.data(), red_head.p, n * sizeof(int), hipMemcpyDeviceToHost, s));
  NJ_HIP(hipMemcpyAsync(h_slot_of_index.data(), slot_of_index.p, n * sizeof(int),
                        hipMemcpyDeviceToHost, s));
  NJ_HIP(hipStreamSynchronize(s));
  if (h_st.error || h_st.step != (int)n - 2)
    return fail(RNJ_E_INVALID, "neighbour joining stopped early (no pair found)");
  // the last join (rapidNJ.cpp:52-85)
  int index1 = -1, index2 = -1;
  bool copy_pair = false;  // the last two are a row and its redundant copy
  for (unsigned i = 0; i < n; ++i) {
    if (h_index_to_id[i] == -1) continue;
    if (index1 == -1) {
      if (h_red_count[i] == 2) {
        index1 = (int)i;
        index2 = h_red_head[i];
        h_index_to_id[index2] = index2;
        copy_pair = true;
        break;
      }
      index1 = (int)i;
    } else {
      index2 = (int)i;
      break;
    }
  }
  if (index1 == -1 || index2 == -1) return fail(RNJ_E_INVALID, "an error occurred while constructing the tree");
  // matrix[index1][index2]. A row with a copy left is an input row never rewritten, and the
  // copy's column was never written either, so the input holds the value; otherwise both are
  // live slots, and the pair lives in the row of the larger slot (the step that created that
  // slot wrote its row; its column is only written by a later scan).
  float last_dist = 0.0f;
  const int sl1 = h_slot_of_index[index1], sl2 = h_slot_of_index[index2];
  if (copy_pair) {
    NJ_HIP(hipMemcpy(&last_dist, M + (size_t)index1 * ldm + index2, sizeof(float),
                     hipMemcpyDeviceToHost));
  } else {
    if (sl1 < 0 || sl2 < 0) return fail(RNJ_E_INVALID, "last clusters without slots");
    const size_t hi = (size_t)std::max(sl1, sl2), lo = (size_t)std::min(sl1, sl2);
    NJ_HIP(hipMemcpy(&last_dist, S + hi * cap + lo, sizeof(float), hipMemcpyDeviceToHost));
  }
  t->n = n;
  t->left.assign(hl.begin(), hl.begin() + (n - 2));
  t->right.assign(hr.begin(), hr.begin() + (n - 2));
  t->llen.assign(hll.begin(), hll.begin() + (n - 2));
  t->rlen.assign(hrl.begin(), hrl.begin() + (n - 2));
  t->last[0] = h_index_to_id[index1];
  t->last[1] = h_index_to_id[index2];
  t->last_dist = last_dist;
  (void)ctx;
  return RNJ_OK;
}

// polytree::serialize_tree / serialize_node without bootstrap counts, iterative (deep trees).
void write_newick(const rnj_tree& t, const std::vector<std::string>& names, std::string& out) {
  const unsigned n = t.n;
  std::vector<double> dist(2 * (size_t)n, 0.0);
  for (unsigned k = 0; k + 2 < n; ++k) {
    dist[t.left[k]] = t.llen[k];
    dist[t.right[k]] = t.rlen[k];
  }
  char buf[64];
  auto len = [&](double x) {
    const int k = std::snprintf(buf, sizeof buf, "%.5g", x);
    out.append(buf, (size_t)k);
  };
  auto leaf = [&](int id, double d) {
    out += '\'';
    out += names[id];
    out += "':";
    len(d);
  };
  // work items: 0 = node(l, r) of internal id x; 1 = close internal child x ("):" + length);
  // 2 = emit ","; 3 = child x (leaf or "(" node ")")
  struct Item {
    int kind, x;
  };
  std::vector<Item> stack;
  auto push_node = [&](int x) {  // node(left[x-n], right[x-n]) = child(l) "," child(r)
    stack.push_back({3, t.right[x - n]});
    stack.push_back({2, 0});
    stack.push_back({3, t.left[x - n]});
  };
  auto run = [&]() {
    while (!stack.empty()) {
      const Item it = stack.back();
      stack.pop_back();
      if (it.kind == 2) {
        out += ',';
      } else if (it.kind == 1) {
        out += "):";
        len(dist[it.x]);
      } else if (it.kind == 3) {
        if ((unsigned)it.x < n) {
          leaf(it.x, dist[it.x]);
        } else {
          out += '(';
          stack.push_back({1, it.x});
          push_node(it.x);
        }
      }
    }
  };
  const int sl = t.last[0], sr = t.last[1];
  out += '(';
  if ((unsigned)sl < n) {
    leaf(sl, t.last_dist);
  } else {
    if ((unsigned)sr >= n) out += '(';
    push_node(sl);
    run();
    if ((unsigned)sr >= n) {
      out += "):";
      len(t.last_dist);
    }
  }
  out += ',';
  if ((unsigned)sr < n) {
    leaf(sr, t.last_dist);
  } else {
    push_node(sr);
    run();
  }
  out += ");\n";
}

}  // namespace
}  // namespace rnj

using namespace rnj;

extern "C" {

int rnj_nj_device(rnj_context* ctx, const float* matrix, unsigned n, size_t ld,
                  int negative_branches, void* stream, rnj_tree** out) {
  if (!ctx || !matrix || !out) return fail(RNJ_E_INVALID, "null argument");
  if (n < 3) return fail(RNJ_E_INVALID, "need at least 3 taxa");
  if (ld < n) return fail(RNJ_E_INVALID, "ld smaller than n");
  if (n > (1u << 24)) return fail(RNJ_E_INVALID, "too many taxa");
  (void)hipGetLastError();
  RNJ_HIP(hipSetDevice(ctx->device));
  hipStream_t s = (hipStream_t)stream;
  RNJ_HIP(hipStreamSynchronize(s));
  auto* t = new rnj_tree();
  const int rc = run_nj(ctx, matrix, ld, n, negative_branches, s, t);
  if (rc != RNJ_OK) {
    delete t;
    return rc;
  }
  *out = t;
  return RNJ_OK;
}

int rnj_nj_rows(rnj_context* ctx, const float* const* rows, unsigned n, int negative_branches,
                rnj_tree** out) {
  if (!ctx || !rows || !out) return fail(RNJ_E_INVALID, "null argument");
  if (n < 3) return fail(RNJ_E_INVALID, "need at least 3 taxa");
  for (unsigned i = 0; i < n; ++i)
    if (!rows[i]) return fail(RNJ_E_INVALID, "null matrix row");
  (void)hipGetLastError();
  RNJ_HIP(hipSetDevice(ctx->device));
  float* d = nullptr;
  RNJ_HIP(hipMalloc((void**)&d, (size_t)n * n * sizeof(float)));
  hipError_t e = hipSuccess;
  {
    // rows packed into the two pinned buffers in turn, each uploaded while the other fills
    std::lock_guard<std::mutex> lock(ctx->pin_mu);
    e = ensure_pinned(ctx);
    const size_t row_bytes = (size_t)n * sizeof(float);
    const size_t per = std::max<size_t>(1, kPinBytes / row_bytes);
    int k = 0;
    for (unsigned r0 = 0; e == hipSuccess && r0 < n; r0 += (unsigned)per, k ^= 1) {
      const unsigned r1 = (unsigned)std::min<size_t>(n, r0 + per);
      e = hipEventSynchronize(ctx->pin_ev[k]);
      if (e != hipSuccess) break;
      if (row_bytes <= kPinBytes) {
        for (unsigned r = r0; r < r1; ++r)
          std::memcpy(ctx->pin[k] + (size_t)(r - r0) * row_bytes, rows[r], row_bytes);
        e = hipMemcpyAsync(d + (size_t)r0 * n, ctx->pin[k], (size_t)(r1 - r0) * row_bytes,
                           hipMemcpyHostToDevice, ctx->stream);
      } else {
        e = hipMemcpyAsync(d + (size_t)r0 * n, rows[r0], row_bytes, hipMemcpyHostToDevice,
                           ctx->stream);
      }
      if (e == hipSuccess) e = hipEventRecord(ctx->pin_ev[k], ctx->stream);
    }
  }
  rnj_tree* t = nullptr;
  int rc = e == hipSuccess ? rnj_nj_device(ctx, d, n, n, negative_branches, ctx->stream, &t)
                           : hip_fail(e, "rnj_nj_rows upload");
  (void)hipStreamSynchronize(ctx->stream);
  (void)hipFree(d);
  if (rc != RNJ_OK) return rc;
  *out = t;
  return RNJ_OK;
}

int rnj_nj_alignment(rnj_context* ctx, const rnj_alignment* aln, int model, int negative_branches,
                     rnj_tree** out) {
  if (!ctx || !aln || !out) return fail(RNJ_E_INVALID, "null argument");
  if (aln->ctx != ctx) return fail(RNJ_E_INVALID, "alignment belongs to another context");
  const unsigned n = aln->n;
  if (n < 3) return fail(RNJ_E_INVALID, "need at least 3 taxa");
  (void)hipGetLastError();
  RNJ_HIP(hipSetDevice(ctx->device));
  float* d = nullptr;
  RNJ_HIP(hipMalloc((void**)&d, (size_t)n * n * sizeof(float)));
  // the post-processed square matrix the reference hands to rapidNJ, computed with the exact
  // FP64 logs (KimuraDistance / JCdistance + postProcessDistanceMatrix)
  int rc = rnj_distances_device(ctx, aln, model, RNJ_SQUARE, d, n, RNJ_FLAG_EXACT, nullptr,
                                ctx->stream);
  rnj_tree* t = nullptr;
  if (rc == RNJ_OK) rc = rnj_nj_device(ctx, d, n, n, negative_branches, ctx->stream, &t);
  (void)hipStreamSynchronize(ctx->stream);
  (void)hipFree(d);
  if (rc != RNJ_OK) return rc;
  t->names = aln->names;
  t->names.resize(std::min<size_t>(t->names.size(), n));
  *out = t;
  return RNJ_OK;
}

unsigned rnj_tree_size(const rnj_tree* t) { return t ? t->n : 0u; }

int rnj_tree_joins(const rnj_tree* t, int* left, int* right, double* left_len, double* right_len,
                   int* last, float* last_dist) {
  if (!t) return fail(RNJ_E_INVALID, "null tree");
  const size_t k = t->left.size();
  if (left) std::copy(t->left.begin(), t->left.end(), left);
  if (right) std::copy(t->right.begin(), t->right.end(), right);
  if (left_len) std::copy(t->llen.begin(), t->llen.end(), left_len);
  if (right_len) std::copy(t->rlen.begin(), t->rlen.end(), right_len);
  if (last) {
    last[0] = t->last[0];
    last[1] = t->last[1];
  }
  if (last_dist) *last_dist = t->last_dist;
  (void)k;
  return RNJ_OK;
}

int rnj_tree_newick(const rnj_tree* t, const char* const* names, char* out, size_t cap,
                    size_t* length) {
  if (!t) return fail(RNJ_E_INVALID, "null tree");
  std::vector<std::string> nm(t->n);
  for (unsigned i = 0; i < t->n; ++i) {
    if (names && names[i]) nm[i] = names[i];
    else if (i < t->names.size()) nm[i] = t->names[i];
    else nm[i] = "t" + std::to_string(i);
  }
  std::string text;
  write_newick(*t, nm, text);
  if (length) *length = text.size();
  if (out && cap > 0) {
    const size_t k = std::min(cap - 1, text.size());
    std::memcpy(out, text.data(), k);
    out[k] = 0;
  }
  return RNJ_OK;
}

int rnj_tree_free(rnj_tree* t) {
  delete t;
  return RNJ_OK;
}

}  // extern "C"
