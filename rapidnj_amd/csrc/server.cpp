// Served distance-matrix jobs: a request loop that admits jobs by HBM budget and batches them.
//
// Submitted jobs wait in a FIFO queue. A worker thread forms a batch from the queue head as long
// as the jobs' device memory (characters, packed planes, triangle, saturated-pair list) fits the
// remaining budget and a stream slot is free, then enqueues on that slot's stream: upload, pack,
// ONE multi-job tile launch for the whole batch, saturation patch and download of every job's
// triangle into pinned memory. Up to `slots` batches are on the device at once, so jobs that
// arrive while others run are admitted as soon as a slot and memory free up, and their uploads,
// counting and downloads overlap with the running batches. Each job's result is bit-identical to
// running it alone: the counts are exact integers and every job keeps its own maximum and list
// of saturated pairs.
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <thread>

#include "runtime.hpp"

using namespace rnj;
using Clock = std::chrono::steady_clock;

namespace {

enum class State { Queued, Admitted, Done };

struct ServedJob {
  uint64_t id = 0;
  unsigned n = 0, L = 0;
  int model = 0;
  std::vector<char> chars;
  std::vector<float> result;
  State state = State::Queued;
  int status = RNJ_OK;
  std::string error;
  Clock::time_point t_submit, t_admit, t_done;
  unsigned batch_id = 0, batch_jobs = 0;
  size_t need = 0;
};

struct Slot {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  float* pinned = nullptr;  // downloads of the batch's triangles
  size_t pinned_floats = 0;
  bool busy = false;
  unsigned batch_id = 0;
  std::vector<ServedJob*> jobs;
  std::vector<size_t> offsets;  // job k's triangle at pinned + offsets[k]
  size_t bytes = 0;
};

size_t pairs_of(unsigned n) { return (size_t)n * (n - 1) / 2; }

size_t job_bytes(unsigned n, unsigned L) {
  const size_t n_pad = round_up(n, kTaxaPad);
  return (size_t)n * L + (size_t)3 * words_for_sites(L) * n_pad * 4 + pairs_of(n) * 4 +
         sentinel_capacity(n) * 8;
}

}  // namespace

struct rnj_server {
  Context* ctx = nullptr;
  rnj_server_config cfg{};
  std::vector<Slot> slots;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<ServedJob*> queue;
  std::map<uint64_t, ServedJob*> jobs;
  size_t used = 0;
  uint64_t next_id = 1;
  unsigned next_batch = 1;
  bool stopping = false;
  std::thread worker;

  void run();
  void retire(Slot& s);
  void launch(Slot& s, std::vector<ServedJob*>& batch);
  void fail_batch(Slot& s, int status, const std::string& msg);
};

void rnj_server::fail_batch(Slot& s, int status, const std::string& msg) {
  for (ServedJob* j : s.jobs) {
    j->status = status;
    j->error = msg;
    j->state = State::Done;
    j->t_done = Clock::now();
  }
}

// Enqueue one batch on slot s (called with the lock held; only enqueues work).
void rnj_server::launch(Slot& s, std::vector<ServedJob*>& batch) {
  s.jobs = batch;
  s.batch_id = next_batch++;
  s.bytes = 0;
  s.offsets.assign(batch.size(), 0);
  size_t floats = 0;
  for (size_t k = 0; k < batch.size(); ++k) {
    s.offsets[k] = floats;
    floats += pairs_of(batch[k]->n);
    s.bytes += batch[k]->need;
  }
  const Clock::time_point now = Clock::now();
  for (ServedJob* j : batch) {
    j->state = State::Admitted;
    j->t_admit = now;
    j->batch_id = s.batch_id;
    j->batch_jobs = (unsigned)batch.size();
  }
  used += s.bytes;
  s.busy = true;
  hipError_t e = hipSuccess;
  if (floats > s.pinned_floats) {
    if (s.pinned) (void)hipHostFree(s.pinned);
    s.pinned = nullptr;
    s.pinned_floats = 0;
    e = hipHostMalloc((void**)&s.pinned, floats * sizeof(float), hipHostMallocDefault);
    if (e == hipSuccess) s.pinned_floats = floats;
  }
  std::vector<void*> allocs;
  std::vector<JobSpec> specs;
  std::vector<float*> tris;
  for (size_t k = 0; e == hipSuccess && k < batch.size(); ++k) {
    ServedJob* j = batch[k];
    const unsigned n_pad = round_up(j->n, kTaxaPad), W = words_for_sites(j->L);
    uint8_t* dchars = nullptr;
    uint32_t* planes = nullptr;
    float* tri = nullptr;
    e = hipMallocAsync((void**)&dchars, (size_t)j->n * j->L, s.stream);
    if (e == hipSuccess) allocs.push_back(dchars);
    if (e == hipSuccess) e = hipMallocAsync((void**)&planes, (size_t)3 * W * n_pad * 4, s.stream);
    if (e == hipSuccess) allocs.push_back(planes);
    if (e == hipSuccess) e = hipMallocAsync((void**)&tri, pairs_of(j->n) * 4, s.stream);
    if (e == hipSuccess) allocs.push_back(tri);
    if (e == hipSuccess)
      e = hipMemcpyAsync(dchars, j->chars.data(), (size_t)j->n * j->L, hipMemcpyHostToDevice,
                         s.stream);
    if (e == hipSuccess) e = launch_pack_chars(dchars, j->L, j->n, j->L, n_pad, W, planes, s.stream);
    JobSpec js{};
    js.planes = planes;
    js.n = j->n;
    js.n_pad = n_pad;
    js.W_pad = W;
    js.W = W;
    js.model = j->model;
    js.layout = kTriangle;
    js.out = tri;
    specs.push_back(js);
    tris.push_back(tri);
  }
  int rc = RNJ_OK;
  if (e == hipSuccess) rc = launch_jobs(ctx, specs, false, s.stream);
  for (size_t k = 0; e == hipSuccess && rc == RNJ_OK && k < batch.size(); ++k)
    e = hipMemcpyAsync(s.pinned + s.offsets[k], tris[k], pairs_of(batch[k]->n) * 4,
                       hipMemcpyDeviceToHost, s.stream);
  for (void* p : allocs) (void)hipFreeAsync(p, s.stream);
  if (e == hipSuccess && rc == RNJ_OK) e = hipEventRecord(s.done, s.stream);
  if (e != hipSuccess || rc != RNJ_OK) {
    const std::string msg = e != hipSuccess ? std::string("batch launch: ") + hipGetErrorString(e)
                                            : std::string(last_error_cstr());
    fail_batch(s, e != hipSuccess ? RNJ_E_HIP : rc, msg);
    (void)hipStreamSynchronize(s.stream);
    used -= s.bytes;
    s.busy = false;
    s.jobs.clear();
    cv_done.notify_all();
  }
}

void rnj_server::retire(Slot& s) {
  const Clock::time_point now = Clock::now();
  for (size_t k = 0; k < s.jobs.size(); ++k) {
    ServedJob* j = s.jobs[k];
    j->result.assign(s.pinned + s.offsets[k], s.pinned + s.offsets[k] + pairs_of(j->n));
    j->chars.clear();
    j->chars.shrink_to_fit();
    j->status = RNJ_OK;
    j->state = State::Done;
    j->t_done = now;
  }
  used -= s.bytes;
  s.busy = false;
  s.jobs.clear();
  cv_done.notify_all();
}

void rnj_server::run() {
  (void)hipSetDevice(ctx->device);
  std::unique_lock<std::mutex> lock(mu);
  for (;;) {
    bool any_busy = false;
    for (Slot& s : slots) {
      if (!s.busy) continue;
      const hipError_t q = hipEventQuery(s.done);
      if (q == hipSuccess) retire(s);
      else if (q != hipErrorNotReady) {
        fail_batch(s, RNJ_E_HIP, std::string("batch failed: ") + hipGetErrorString(q));
        used -= s.bytes;
        s.busy = false;
        s.jobs.clear();
        cv_done.notify_all();
      } else {
        any_busy = true;
      }
    }
    // admission: FIFO, as many head jobs as fit the budget, one batch per free slot
    for (Slot& s : slots) {
      if (s.busy || queue.empty()) continue;
      std::vector<ServedJob*> batch;
      size_t add = 0;
      while (!queue.empty() && batch.size() < cfg.max_batch_jobs) {
        ServedJob* j = queue.front();
        if (j->need > cfg.hbm_budget) {  // can never fit
          queue.pop_front();
          j->status = RNJ_E_ADMISSION;
          j->error = "job needs more HBM than the server budget";
          j->state = State::Done;
          j->t_done = Clock::now();
          cv_done.notify_all();
          continue;
        }
        if (used + add + j->need > cfg.hbm_budget) break;
        add += j->need;
        batch.push_back(j);
        queue.pop_front();
      }
      if (batch.empty()) break;
      launch(s, batch);
      if (s.busy) any_busy = true;
    }
    if (stopping && queue.empty() && !any_busy) break;
    if (any_busy)
      cv_work.wait_for(lock, std::chrono::microseconds(100));
    else
      cv_work.wait(lock, [this] { return stopping || !queue.empty(); });
  }
}

extern "C" {

int rnj_server_create(rnj_context* ctx, const rnj_server_config* cfg, rnj_server** out) {
  if (!ctx || !out) return fail(RNJ_E_INVALID, "null argument");
  RNJ_HIP(hipSetDevice(ctx->device));
  auto* srv = new rnj_server();
  srv->ctx = ctx;
  if (cfg) srv->cfg = *cfg;
  if (srv->cfg.slots == 0) srv->cfg.slots = 2;
  if (srv->cfg.max_batch_jobs == 0) srv->cfg.max_batch_jobs = 64;
  if (srv->cfg.hbm_budget == 0) {
    size_t free_b = 0, total_b = 0;
    hipError_t e = hipMemGetInfo(&free_b, &total_b);
    if (e != hipSuccess) {
      delete srv;
      return hip_fail(e, "hipMemGetInfo");
    }
    srv->cfg.hbm_budget = (size_t)(0.8 * (double)free_b);
  }
  srv->slots.resize(srv->cfg.slots);
  for (Slot& s : srv->slots) {
    hipError_t e = hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&s.done, hipEventDisableTiming);
    if (e != hipSuccess) {
      for (Slot& t : srv->slots) {
        if (t.stream) (void)hipStreamDestroy(t.stream);
        if (t.done) (void)hipEventDestroy(t.done);
      }
      delete srv;
      return hip_fail(e, "rnj_server_create");
    }
  }
  srv->worker = std::thread([srv] { srv->run(); });
  *out = srv;
  return RNJ_OK;
}

int rnj_server_submit(rnj_server* srv, const char* chars, size_t pitch, unsigned n, unsigned L,
                      int model, uint64_t* ticket) {
  if (!srv || !chars || !ticket) return fail(RNJ_E_INVALID, "null argument");
  if (n < 2) return fail(RNJ_E_INVALID, "need at least 2 sequences");
  if (L < 1 || pitch < L) return fail(RNJ_E_INVALID, "bad sequence length or pitch");
  if (model != RNJ_JC && model != RNJ_K2P) return fail(RNJ_E_INVALID, "model must be RNJ_JC or RNJ_K2P");
  auto* j = new ServedJob();
  j->n = n;
  j->L = L;
  j->model = model;
  j->chars.resize((size_t)n * L);
  for (unsigned t = 0; t < n; ++t) std::memcpy(&j->chars[(size_t)t * L], chars + (size_t)t * pitch, L);
  j->need = job_bytes(n, L);
  j->t_submit = Clock::now();
  {
    std::lock_guard<std::mutex> lock(srv->mu);
    if (srv->stopping) {
      delete j;
      return fail(RNJ_E_INVALID, "server is shutting down");
    }
    j->id = srv->next_id++;
    srv->jobs[j->id] = j;
    srv->queue.push_back(j);
    *ticket = j->id;
  }
  srv->cv_work.notify_one();
  return RNJ_OK;
}

int rnj_server_poll(rnj_server* srv, uint64_t ticket) {
  if (!srv) return fail(RNJ_E_INVALID, "null server");
  std::lock_guard<std::mutex> lock(srv->mu);
  auto it = srv->jobs.find(ticket);
  if (it == srv->jobs.end()) return fail(RNJ_E_INVALID, "unknown ticket");
  return it->second->state == State::Done ? RNJ_OK : RNJ_E_NOT_READY;
}

int rnj_server_wait(rnj_server* srv, uint64_t ticket, float* tri_out, rnj_job_stats* stats) {
  if (!srv) return fail(RNJ_E_INVALID, "null server");
  std::unique_lock<std::mutex> lock(srv->mu);
  auto it = srv->jobs.find(ticket);
  if (it == srv->jobs.end()) return fail(RNJ_E_INVALID, "unknown ticket");
  ServedJob* j = it->second;
  srv->cv_done.wait(lock, [j] { return j->state == State::Done; });
  srv->jobs.erase(it);
  lock.unlock();
  const int status = j->status;
  if (status == RNJ_OK && tri_out) std::memcpy(tri_out, j->result.data(), j->result.size() * 4);
  if (stats) {
    auto sec = [](Clock::duration d) { return std::chrono::duration<double>(d).count(); };
    const bool admitted = j->t_admit.time_since_epoch().count() != 0;
    stats->queued_s = admitted ? sec(j->t_admit - j->t_submit) : sec(j->t_done - j->t_submit);
    stats->device_s = admitted ? sec(j->t_done - j->t_admit) : 0.0;
    stats->total_s = sec(j->t_done - j->t_submit);
    stats->batch_id = j->batch_id;
    stats->batch_jobs = j->batch_jobs;
  }
  if (status != RNJ_OK) set_error(j->error);
  delete j;
  return status;
}

int rnj_server_cancel(rnj_server* srv, uint64_t ticket) {
  if (!srv) return fail(RNJ_E_INVALID, "null server");
  std::lock_guard<std::mutex> lock(srv->mu);
  auto it = srv->jobs.find(ticket);
  if (it == srv->jobs.end()) return fail(RNJ_E_INVALID, "unknown ticket");
  ServedJob* j = it->second;
  if (j->state != State::Queued) return fail(RNJ_E_INVALID, "job already admitted");
  for (auto q = srv->queue.begin(); q != srv->queue.end(); ++q)
    if (*q == j) {
      srv->queue.erase(q);
      break;
    }
  j->status = RNJ_E_CANCELLED;
  j->error = "cancelled";
  j->state = State::Done;
  j->t_done = Clock::now();
  srv->cv_done.notify_all();
  return RNJ_OK;
}

int rnj_server_destroy(rnj_server* srv) {
  if (!srv) return RNJ_OK;
  {
    std::lock_guard<std::mutex> lock(srv->mu);
    srv->stopping = true;
    // jobs not yet admitted are cancelled; admitted batches finish
    for (ServedJob* j : srv->queue) {
      j->status = RNJ_E_CANCELLED;
      j->error = "server destroyed";
      j->state = State::Done;
    }
    srv->queue.clear();
  }
  srv->cv_work.notify_all();
  srv->cv_done.notify_all();
  if (srv->worker.joinable()) srv->worker.join();
  for (Slot& s : srv->slots) {
    if (s.stream) (void)hipStreamSynchronize(s.stream);
    if (s.pinned) (void)hipHostFree(s.pinned);
    if (s.stream) (void)hipStreamDestroy(s.stream);
    if (s.done) (void)hipEventDestroy(s.done);
  }
  for (auto& kv : srv->jobs) delete kv.second;
  delete srv;
  return RNJ_OK;
}

}  // extern "C"
