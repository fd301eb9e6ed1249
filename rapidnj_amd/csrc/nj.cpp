// Host driver of the device neighbour joining (nj.hip): set-up as rapidNJ::initialize
// (rapidNJ.cpp:91-138), the iteration loop of rapidNJ::run (:45-50) enqueued in windows without
// waits, the searches the device leaves to the host (exact ties and the row-scan break, settled
// with rapidNJ's own sorted row lists), the last join (:52-85), and the tree with polytree's
// Newick text (polytree.cpp:402-458).
#include <algorithm>
#include <cfloat>
#include <cstdio>
#include <cstring>
#include <deque>
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
  unsigned host_searches = 0;    // iterations whose search the host settled
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
  DevBuf<int> slot_index, alive, row_p, map, left, right, row_tie, newidx, log_idx, shave;
  DevBuf<float> sepsum, sep, newcol, term, sums, S0, S1, row_maxsep, init_maxsep, log_val, smin;
  DevBuf<unsigned> row_sigma, key_by_index;
  DevBuf<unsigned long long> row_key;
  DevBuf<double> llen, rlen, sref;
  // the iteration plan: slot bound, compaction before it, slot-matrix buffer, log offset
  const int iters = (int)n - 2;
  std::vector<int> plan_ub(iters), plan_buf(iters);
  std::vector<char> plan_compact(iters);
  std::vector<size_t> plan_log(iters);
  size_t log_total = 0;
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
  NJ_HIP(row_sigma.alloc(cap));
  NJ_HIP(row_tie.alloc(cap));
  NJ_HIP(row_maxsep.alloc(cap));
  NJ_HIP(smin.alloc(cap));
  NJ_HIP(sref.alloc(cap));
  NJ_HIP(shave.alloc(cap));
  NJ_HIP(init_maxsep.alloc(cap));
  NJ_HIP(key_by_index.alloc(n));
  NJ_HIP(newidx.alloc(n));
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
  // maxRowSeparations of the first search: the largest separation of every smaller index,
  // copies included (rapidNJ.cpp:112-119)
  std::vector<float> h_init_maxsep(cap, FLT_MIN);
  {
    std::vector<float> by_index(n);
    float mx = FLT_MIN;
    for (unsigned i = 0; i < n; ++i) {
      by_index[i] = mx;
      const float sp = h_sums[i] / (float)((int)n - 2);
      if (sp > mx) mx = sp;
    }
    for (int a = 0; a < v; ++a) h_init_maxsep[a] = by_index[vis[a]];
  }
  const std::vector<int> init_visible = h_index_to_id;  // -1: a copy hidden at set-up
  NjState h_st{};
  h_st.n = (int)n;
  h_st.used = v;
  h_st.cluster_count = (int)n;
  h_st.current_id = (int)n;
  h_st.negative_branches = negative_branches ? 1 : 0;
  h_st.start_bound = FLT_MAX;  // no bound before the first step
  h_st.start_row = -1;

  // rapidNJ::run: n - 2 joins. The plan's slot bound bounds the slots in use (exact until the
  // first compaction, then the live count's bound clusterCount plus the slots added since). The
  // slot matrix is compacted when that bound passes 1.125 x clusterCount (or the capacity).
  {
    size_t b = (size_t)v;
    int buf = 0;
    for (int it = 0; it < iters; ++it) {
      const size_t cc = n - (size_t)it;  // clusterCount before this join
      if (b + 1 > cap || (b > 256 && b * 8 > cc * 9)) {
        plan_compact[it] = 1;
        b = std::min(cc, b);
        buf ^= 1;
      }
      plan_ub[it] = (int)b;
      plan_buf[it] = buf;
      plan_log[it] = log_total;
      log_total += b;
      ++b;
    }
  }
  NJ_HIP(log_val.alloc(log_total));
  NJ_HIP(log_idx.alloc(log_total));
  NJ_HIP(hipMemsetAsync(log_idx.p, 0xff, std::max<size_t>(log_total, 1) * sizeof(int), s));

  auto up = [&](void* dst, const void* src, size_t bytes) {
    return hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s);
  };
  auto down = [&](void* dst, const void* src, size_t bytes) {
    return hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s);
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
  NJ_HIP(up(init_maxsep.p, h_init_maxsep.data(), cap * sizeof(float)));
  NJ_HIP(hipMemsetAsync(key_by_index.p, 0xff, n * sizeof(unsigned), s));  // no answers yet
  NJ_HIP(hipMemsetAsync(row_p.p, 0xff, cap * sizeof(int), s));
  NJ_HIP(hipMemsetAsync(shave.p, 0, cap * sizeof(int), s));
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
  D.row_sigma = row_sigma.p;
  D.row_tie = row_tie.p;
  D.row_maxsep = row_maxsep.p;
  D.smin = smin.p;
  D.sref = sref.p;
  D.shave = shave.p;
  D.init_maxsep = init_maxsep.p;
  D.newcol = newcol.p;
  D.term = term.p;
  D.map = map.p;
  D.key_by_index = key_by_index.p;
  D.newidx = newidx.p;
  D.log_val = log_val.p;
  D.log_idx = log_idx.p;
  D.left = left.p;
  D.right = right.p;
  D.llen = llen.p;
  D.rlen = rlen.p;
  float* const Sbuf[2] = {S0.p, S1.p};

  // Iteration `it` as planned (or only its step, after the host settled its search). The
  // previous join's new cluster is finished by this iteration's scan, or before a compaction.
  auto enqueue = [&](int it, bool step_only) -> hipError_t {
    float* S = Sbuf[plan_buf[it]];
    if (step_only) return nj_launch_step(D, S, plan_log[it], s);
    if (plan_compact[it]) {
      float* Sp = Sbuf[plan_buf[it] ^ 1];
      if (it > 0) {
        const hipError_t e = nj_launch_finish(D, Sp, s);
        if (e != hipSuccess) return e;
      }
      const hipError_t e = nj_launch_compact(D, Sp, S, plan_ub[it], s);
      if (e != hipSuccess) return e;
    }
    return nj_launch_iteration(D, S, plan_ub[it], it > 0 && !plan_compact[it], plan_log[it], s);
  };

  // The search of a halted iteration, as rapidNJ's findMin makes it (rapidNJ.cpp:290-361): rows
  // in visiting order; a row that cannot improve the running minimum G (its answer is not below
  // G) is passed over; a row whose answer stands (G above its break bound, no tie at its best) is
  // a record with that answer; any other row is scanned as rapidNJ scans it, over its sorted
  // list (the list as the row was built, std::sort on the distance alone: rapidNJ.cpp:141-153,
  // 219-234), with the break (:340). The result goes back as NjState::forced.
  int t_halt = 0;
  auto settle = [&]() -> int {
    NjState h{};
    NJ_HIP(down(&h, st.p, sizeof h));
    NJ_HIP(hipStreamSynchronize(s));
    if (h.error != kNjHalted) return fail(RNJ_E_INVALID, "neighbour joining stopped early");
    const int step = h.step, used = h.used;
    if (step < 0 || step >= iters || used <= 0 || (size_t)used > cap)
      return fail(RNJ_E_INVALID, "neighbour joining state out of range");
    std::vector<int> i2id(n), id2i(2 * (size_t)n), rc(n), rhead(n), rt(n), soi(n), nidx(step + 1);
    std::vector<unsigned long long> key(used);
    std::vector<int> rp(used), tie(used);
    std::vector<unsigned> sig(used);
    std::vector<float> sp(used), mxs(used);
    NJ_HIP(down(i2id.data(), index_to_id.p, n * sizeof(int)));
    NJ_HIP(down(id2i.data(), id_to_index.p, 2 * (size_t)n * sizeof(int)));
    NJ_HIP(down(rc.data(), red_count.p, n * sizeof(int)));
    NJ_HIP(down(rhead.data(), red_head.p, n * sizeof(int)));
    NJ_HIP(down(rt.data(), rtm.p, n * sizeof(int)));
    NJ_HIP(down(soi.data(), slot_of_index.p, n * sizeof(int)));
    if (step > 0) NJ_HIP(down(nidx.data(), newidx.p, step * sizeof(int)));
    NJ_HIP(down(key.data(), row_key.p, used * sizeof(unsigned long long)));
    NJ_HIP(down(rp.data(), row_p.p, used * sizeof(int)));
    NJ_HIP(down(tie.data(), row_tie.p, used * sizeof(int)));
    NJ_HIP(down(sig.data(), row_sigma.p, used * sizeof(unsigned)));
    NJ_HIP(down(sp.data(), sep.p, used * sizeof(float)));
    NJ_HIP(down(mxs.data(), row_maxsep.p, used * sizeof(float)));
    NJ_HIP(hipStreamSynchronize(s));
    auto slot_ok = [&](int slot) { return slot >= 0 && slot < used; };
    auto sep_of = [&](int idx) -> float {  // separations[redundantToMasterMap[idx]]
      const int slot = soi[rt[idx]];
      return slot_ok(slot) ? sp[slot] : 0.0f;
    };
    auto fkey_host = [](float x) -> unsigned {
      unsigned u;
      const float y = x + 0.0f;
      std::memcpy(&u, &y, 4);
      return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
    };
    auto unkey = [](unsigned u) -> float {
      const unsigned b = (u & 0x80000000u) ? (u & 0x7fffffffu) : ~u;
      float x;
      std::memcpy(&x, &b, 4);
      return x;
    };
    struct Pair {  // cluster_pair: compared on the distance alone (cluster_pair.h:8-10)
      int id;
      float distance;
      bool operator<(const Pair& o) const { return distance < o.distance; }
    };
    // the row list of cluster `cid` as it was built, then sorted
    auto sorted_list = [&](int cid, std::vector<Pair>& out) -> int {
      out.clear();
      if (cid < (int)n) {  // an initial row: visible indices below it (rapidNJ.cpp:141-153)
        std::vector<float> row(std::max(cid, 1));
        if (cid > 0) {
          NJ_HIP(down(row.data(), M + (size_t)cid * ldm, cid * sizeof(float)));
          NJ_HIP(hipStreamSynchronize(s));
        }
        for (int j = 0; j < cid; ++j)
          if (init_visible[j] != -1) out.push_back(Pair{j, row[j]});
      } else {  // made at step k: every other live index, in index order (:219-234)
        const int k = cid - (int)n;
        if (k < 0 || k >= step) return fail(RNJ_E_INVALID, "row list of an unknown cluster");
        const size_t len = (size_t)plan_ub[k];
        std::vector<float> val(len);
        std::vector<int> idx(len);
        NJ_HIP(down(val.data(), log_val.p + plan_log[k], len * sizeof(float)));
        NJ_HIP(down(idx.data(), log_idx.p + plan_log[k], len * sizeof(int)));
        NJ_HIP(hipStreamSynchronize(s));
        // the cluster at each index then: the last one made there before step k, else the
        // index's own
        std::vector<int> at(n);
        for (unsigned i = 0; i < n; ++i) at[i] = (int)i;
        for (int j = 0; j < k; ++j)
          if (nidx[j] >= 0 && nidx[j] < (int)n) at[nidx[j]] = (int)n + j;
        std::vector<std::pair<int, float>> by_index;
        for (size_t b = 0; b < len; ++b)
          if (idx[b] >= 0 && idx[b] < (int)n) by_index.emplace_back(idx[b], val[b]);
        std::sort(by_index.begin(), by_index.end(),
                  [](const std::pair<int, float>& x, const std::pair<int, float>& y) {
                    return x.first < y.first;
                  });
        for (const auto& e : by_index) out.push_back(Pair{at[e.first], e.second});
      }
      std::sort(out.begin(), out.end());
      return RNJ_OK;
    };
    // maxRowSeparation of a row the scan skipped: the largest separation of the live slots
    // below it, from FLT_MIN as the scan takes it (gone slots hold NaN, which no comparison
    // takes), as a prefix maximum over the slots
    std::vector<float> below(used + 1);
    below[0] = FLT_MIN;
    for (int b = 0; b < used; ++b) below[b + 1] = sp[b] > below[b] ? sp[b] : below[b];
    float G = FLT_MAX;
    int partner = -1;
    std::vector<int> records;
    std::vector<Pair> list;
    const int start = i2id[h.cache1] != -1 ? h.cache1 : h.cache2;
    for (unsigned c = 0; c < n; ++c) {
      const int r = (int)((start + c) % n);
      if (i2id[r] == -1) continue;
      const int a = soi[r];
      if (!slot_ok(a) || key[a] == ~0ull) continue;
      const float vr = unkey((unsigned)(key[a] >> 32));
      if (!(vr < G)) continue;  // for a row the scan skipped, vr is a lower bound of its answer
      if (fkey_host(G) > sig[a] && !tie[a]) {  // the row's own answer stands
        G = vr;
        partner = rp[a];
        records.push_back(r);
        continue;
      }
      // findRowMin (rapidNJ.cpp:324-361) with searchRedundantEntries (:413-430)
      bool improved = false;
      const float row_sep = sp[a];
      const float row_max_sep = row_sep + (tie[a] == 2 ? below[a] : mxs[a]);
      if (rc[r] > 1) {
        const float cv = -2.0f * row_sep;
        if (cv < G) {
          G = cv;
          partner = rhead[r];
          improved = true;
        }
      }
      const int lrc = sorted_list(i2id[r], list);
      if (lrc != RNJ_OK) return lrc;
      for (const Pair& e : list) {
        if (e.id < 0 || e.id >= 2 * (int)n) continue;
        const int o = id2i[e.id];
        if (o == -1) continue;  // a dead pair
        const float bound = e.distance - row_max_sep;
        if (bound >= G) break;
        float val = e.distance - sep_of(o);
        val = val - row_sep;
        if (val < G) {
          G = val;
          partner = o;
          improved = true;
        }
      }
      if (improved) records.push_back(r);
    }
    if (records.empty()) return fail(RNJ_E_INVALID, "neighbour joining stopped early (no pair found)");
    int m = h.min1, c1 = h.cache1, c2 = h.cache2;
    for (int r : records) {
      if (r != m) {
        c2 = c1;
        c1 = m;
        m = r;
      }
    }
    h.f_min1 = m;
    h.f_min2 = partner;
    h.f_cache1 = c1;
    h.f_cache2 = c2;
    h.forced = 1;
    h.error = 0;
    NJ_HIP(up(st.p, &h, sizeof h));
    t_halt = step;
    return RNJ_OK;
  };

  // Windows of iterations, each followed by a copy of the state into pinned memory; the next
  // window is enqueued before the previous one is checked, so the device does not wait. A halt
  // turns the kernels after it into no-ops (they test NjState::error first); the host then
  // settles the search and resumes with that iteration's step.
  struct Pinned {
    NjState* p = nullptr;
    ~Pinned() {
      if (p) (void)hipHostFree(p);
    }
  } pinned;
  NJ_HIP(hipHostMalloc((void**)&pinned.p, 2 * sizeof(NjState), hipHostMallocDefault));
  struct Events {
    hipEvent_t e[2] = {nullptr, nullptr};
    ~Events() {
      for (auto x : e)
        if (x) (void)hipEventDestroy(x);
    }
  } ev;
  NJ_HIP(hipEventCreateWithFlags(&ev.e[0], hipEventDisableTiming));
  NJ_HIP(hipEventCreateWithFlags(&ev.e[1], hipEventDisableTiming));
  int it = 0, window = 32, slot = 0;
  bool step_only = false;
  std::deque<int> inflight;
  while (true) {
    if (it < iters && inflight.size() < 2) {
      const int to = std::min(iters, it + window);
      for (int k = it; k < to; ++k) NJ_HIP(enqueue(k, step_only && k == it));
      step_only = false;
      NJ_HIP(down(&pinned.p[slot], st.p, sizeof(NjState)));
      NJ_HIP(hipEventRecord(ev.e[slot], s));
      inflight.push_back(slot);
      slot ^= 1;
      it = to;
      window = std::min(2 * window, 1024);
      continue;
    }
    if (inflight.empty()) break;
    const int w = inflight.front();
    inflight.pop_front();
    NJ_HIP(hipEventSynchronize(ev.e[w]));
    const int err = pinned.p[w].error;
    if (err == kNjFailed) break;  // reported below
    if (err == kNjHalted) {
      NJ_HIP(hipStreamSynchronize(s));
      inflight.clear();
      const int rc = settle();
      if (rc != RNJ_OK) return rc;
      ++t->host_searches;
      it = t_halt;
      step_only = true;
      window = 32;
    }
  }
  float* S = Sbuf[plan_buf[iters - 1]];
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

unsigned rnj_tree_host_searches(const rnj_tree* t) { return t ? t->host_searches : 0u; }

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
