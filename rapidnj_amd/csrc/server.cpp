// Served distance-matrix jobs: a request loop that admits jobs by HBM budget and batches them.
//
// Jobs come two ways: characters on the host (rnj_server_submit: uploaded and packed by the
// server), or an alignment already packed in HBM (rnj_server_submit_resident: nothing to upload;
// the triangle goes to the caller's device buffer, or to a server buffer that rnj_server_wait
// copies out).
//
// Admission. A submitted job is admitted when its device memory (characters, packed planes,
// triangle, saturated-pair list) fits what the budget has left and no earlier job is waiting.
// An admitted job is uploaded from the caller's buffer on the submitting thread (library threads
// copy it into pinned staging chunks whose DMAs run on the server's upload stream, then the
// packing kernel) and becomes ready when that stream reaches its event; the caller's buffer is
// free when submit returns, without waiting for the DMAs. A job that does not fit is copied
// to host memory and queued; the worker admits queued jobs in FIFO order as batches retire and
// memory frees up.
//
// Batching. The worker forms a batch of up to max_batch_jobs ready jobs per free stream slot
// and enqueues on that slot's stream: a wait for each job's upload, ONE multi-job tile launch
// and the saturation patch. Up to `slots` batches are on the device at once, so uploads of new
// jobs overlap the counting of running batches. A finished job's triangle stays resident in HBM
// until rnj_server_wait copies it, once, straight into the caller's buffer on the caller's
// thread; its memory returns to the budget then. Each job's result is bit-identical to running it
// alone: the counts are exact integers and every job keeps its own maximum and saturated-pair
// list.
#include <algorithm>
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

enum class State { Queued, Uploading, Ready, Running, Done };

struct ServedJob {
  uint64_t id = 0;
  unsigned n = 0, L = 0, n_pad = 0, W = 0;
  int model = 0;
  int kind = kDNA;                 // jobs of one kind (and code width) share a launch
  unsigned code_bits = 0;
  const Alignment* aln = nullptr;  // resident job: the caller's alignment (planes in HBM)
  bool own_tri = true;             // the triangle is a server block (else the caller's buffer)
  std::vector<char> chars;  // host copy, only while queued
  State state = State::Queued;
  int status = RNJ_OK;
  std::string error;
  Clock::time_point t_submit, t_admit, t_ready, t_launch, t_done;
  unsigned batch_id = 0, batch_jobs = 0;
  size_t need = 0;
  uint8_t* chars_dev = nullptr;  // device blocks from the server's cache: characters and planes
  uint32_t* planes = nullptr;    // from admission until the job's batch retires,
  float* tri = nullptr;          // the triangle until rnj_server_wait has copied it
  hipEvent_t ready = nullptr;  // recorded after packing
};

struct Slot {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  bool busy = false;
  unsigned batch_id = 0;
  std::vector<ServedJob*> jobs;
  std::string error;
  int error_status = RNJ_OK;
};

size_t pairs_of(unsigned n) { return (size_t)n * (n - 1) / 2; }

// Counting work of a job in pair-words (pairs x 32-site words). A batch takes jobs while their
// sum stays within max_batch_work (a job above it runs alone). Measured on MI355X with the
// 512-job stream resident in HBM: several slots with one job per launch keep the device busiest
// (4 slots: 66 ms; 2 slots: 81 ms), while batches of medium and large jobs ran slower (16-64 jobs
// per launch: 85-90 ms). Batching pays for small jobs, whose launch costs more than their
// counting; the default, 2^28 pair-words (about 50 us of counting), groups only those.
unsigned long long job_work(unsigned n, unsigned W) { return (unsigned long long)pairs_of(n) * W; }
constexpr unsigned long long kDefaultBatchWork = 1ull << 28;
constexpr unsigned kMaxSlots = 4;

size_t tri_bytes(unsigned n) { return pairs_of(n) * 4; }
size_t chars_bytes(unsigned n, unsigned L) { return (size_t)n * L; }
size_t planes_bytes(unsigned n, unsigned L) {
  return (size_t)3 * words_for_sites(L) * round_up(n, kTaxaPad) * 4;
}

// Device blocks by size class, kept after a job for the next job of a similar size. HIP's
// stream-ordered allocator costs 0.1-0.3 ms per call when sizes vary from call to call (three
// calls per job: more than a small job's whole run). A block goes back to the cache only once
// its last use on the device is known to be complete. Size classes: four per power of two
// (at most 25 % above the request), at least 64 KiB. Idle blocks beyond `cap` bytes are freed
// at once; a failed allocation frees every idle block and tries again.
class DeviceCache {
 public:
  static size_t bucket(size_t bytes) {
    constexpr size_t kMin = (size_t)64 << 10;
    if (bytes <= kMin) return kMin;
    const int lg = 63 - __builtin_clzll((unsigned long long)(bytes - 1));
    const size_t step = (size_t)1 << (lg - 2);
    return (bytes + step - 1) / step * step;
  }
  void set_cap(size_t cap) { cap_ = cap; }
  hipError_t get(size_t bytes, void** out) {
    const size_t b = bucket(bytes);
    {
      std::lock_guard<std::mutex> lock(mu_);
      auto it = free_.find(b);
      if (it != free_.end()) {
        *out = it->second;
        free_.erase(it);
        idle_ -= b;
        return hipSuccess;
      }
    }
    hipError_t e = hipMalloc(out, b);
    if (e == hipErrorOutOfMemory) {
      (void)hipGetLastError();
      trim();
      e = hipMalloc(out, b);
    }
    if (e != hipSuccess) *out = nullptr;
    return e;
  }
  void put(void* p, size_t bytes) {
    if (!p) return;
    const size_t b = bucket(bytes);
    {
      std::lock_guard<std::mutex> lock(mu_);
      if (idle_ + b <= cap_) {
        free_.emplace(b, p);
        idle_ += b;
        return;
      }
    }
    (void)hipFree(p);
  }
  void trim() {
    std::multimap<size_t, void*> blocks;
    {
      std::lock_guard<std::mutex> lock(mu_);
      blocks.swap(free_);
      idle_ = 0;
    }
    for (auto& kv : blocks) (void)hipFree(kv.second);
  }

 private:
  std::mutex mu_;
  std::multimap<size_t, void*> free_;
  size_t idle_ = 0, cap_ = 0;
};

// HBM a job holds from admission until its batch retires (in the cache's size classes; its
// saturation records live in the launch's workspace).
size_t job_bytes(unsigned n, unsigned L) {
  return DeviceCache::bucket(chars_bytes(n, L)) + DeviceCache::bucket(planes_bytes(n, L)) +
         DeviceCache::bucket(tri_bytes(n)) + saturation_bytes(n);
}

// The same for a resident job: its triangle when the server holds it, and its saturation records.
size_t resident_job_bytes(unsigned n, bool own_tri) {
  return (own_tri ? DeviceCache::bucket(tri_bytes(n)) : 0) + saturation_bytes(n);
}

}  // namespace

struct rnj_server {
  Context* ctx = nullptr;
  rnj_server_config cfg{};
  std::vector<Slot> slots;
  hipStream_t upload_stream = nullptr;
  hipStream_t download_stream = nullptr;
  Clock::time_point t0 = Clock::now();  // the stats' launched_at / finished_at count from here
  std::mutex mu;
  std::mutex upload_mu;  // one upload at a time on the upload stream
  Stager uploader;       // pinned staging for the callers' characters (submitting threads)
  Stager downloader;     // and for their results (waiting threads)
  DeviceCache blocks;    // jobs' device buffers
  std::condition_variable cv_work, cv_done;
  std::deque<ServedJob*> queue;  // waiting for admission (host copies)
  std::deque<ServedJob*> ready;  // admitted: uploading or uploaded, in arrival order
  std::map<uint64_t, ServedJob*> jobs;
  size_t used = 0;
  uint64_t next_id = 1;
  unsigned next_batch = 1;
  bool stopping = false;
  std::thread worker;

  void run();
  int upload(ServedJob* j, const char* chars, size_t pitch);
  bool launch(Slot& s);
  void release_job_memory(ServedJob* j);
};

// Upload and pack one admitted job (no server lock held). chars: n rows at `pitch` on the host.
// Returns when the caller's buffer has been read; the copies and the packing run on the upload
// stream and the job's ready event follows them.
int rnj_server::upload(ServedJob* j, const char* chars, size_t pitch) {
  if (j->aln) {  // resident: its planes are complete; only a server triangle to take
    const hipError_t e = j->own_tri ? blocks.get(tri_bytes(j->n), (void**)&j->tri) : hipSuccess;
    if (e == hipSuccess) return RNJ_OK;
    j->error = std::string("triangle: ") + hipGetErrorString(e);
    return RNJ_E_HIP;
  }
  std::lock_guard<std::mutex> ul(upload_mu);
  hipStream_t s = upload_stream;
  hipError_t e = blocks.get(chars_bytes(j->n, j->L), (void**)&j->chars_dev);
  if (e == hipSuccess) e = blocks.get(planes_bytes(j->n, j->L), (void**)&j->planes);
  if (e == hipSuccess) e = blocks.get(tri_bytes(j->n), (void**)&j->tri);
  // through the pinned staging chunks; returns once the caller's rows have been read
  if (e == hipSuccess) e = uploader.upload(j->chars_dev, chars, pitch, j->L, j->n, s);
  if (e == hipSuccess)
    e = launch_pack_chars(j->chars_dev, j->L, j->n, j->L, j->n_pad, j->W, j->planes, s);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&j->ready, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventRecord(j->ready, s);
  if (e != hipSuccess) {
    (void)hipStreamSynchronize(s);
    j->error = std::string("upload: ") + hipGetErrorString(e);
    return RNJ_E_HIP;
  }
  return RNJ_OK;
}

// Every device block of the job back to the cache. Callers ensure that no work on the device
// still uses them.
void rnj_server::release_job_memory(ServedJob* j) {
  blocks.put(j->chars_dev, chars_bytes(j->n, j->L));
  blocks.put(j->planes, planes_bytes(j->n, j->L));
  if (j->own_tri) blocks.put(j->tri, tri_bytes(j->n));
  j->chars_dev = nullptr;
  j->planes = nullptr;
  j->tri = nullptr;
}

// Memory a job holds after its batch ran: only a server triangle (released by rnj_server_wait).
static size_t held_after_batch(const ServedJob* j) {
  return j->own_tri ? DeviceCache::bucket(tri_bytes(j->n)) : 0;
}

// Enqueue the batch in s.jobs on slot s (worker thread, no server lock). Returns false with the
// slot's error set if enqueueing failed.
bool rnj_server::launch(Slot& s) {
  hipError_t e = hipSuccess;
  std::vector<JobSpec> specs;
  // longest rows first: their tiles are dispatched first, and the launch ends on short ones
  std::vector<ServedJob*> order(s.jobs);
  std::stable_sort(order.begin(), order.end(),
                   [](const ServedJob* a, const ServedJob* b) { return a->W > b->W; });
  for (ServedJob* j : order) {
    if (e == hipSuccess && j->ready) e = hipStreamWaitEvent(s.stream, j->ready, 0);
    if (j->aln) {
      specs.push_back(alignment_spec(j->aln, j->model, kTriangle, j->tri, 0));
      continue;
    }
    JobSpec js{};
    js.planes = j->planes;
    js.n = j->n;
    js.n_pad = j->n_pad;
    js.W_pad = j->W;
    js.W = j->W;
    js.W_live = (j->L + 31) / 32;
    js.model = j->model;
    js.layout = kTriangle;
    js.out = j->tri;
    specs.push_back(js);
  }
  int rc = RNJ_OK;
  if (e == hipSuccess) rc = launch_jobs(ctx, specs, false, false, s.stream, false);
  if (e == hipSuccess && rc == RNJ_OK) e = hipEventRecord(s.done, s.stream);
  if (e == hipSuccess && rc == RNJ_OK) return true;
  s.error = e != hipSuccess ? std::string("batch launch: ") + hipGetErrorString(e)
                            : std::string(last_error_cstr());
  s.error_status = e != hipSuccess ? RNJ_E_HIP : rc;
  (void)hipStreamSynchronize(s.stream);
  return false;
}

void rnj_server::run() {
  (void)hipSetDevice(ctx->device);
  std::unique_lock<std::mutex> lock(mu);
  for (;;) {
    // 1. batches that finished on the device: their jobs are done at once (before any host
    //    work of this pass), and the slot takes a new batch in step 3
    std::vector<Slot*> failed;
    bool finished = false;
    for (Slot& s : slots) {
      if (!s.busy) continue;
      const hipError_t q = hipEventQuery(s.done);
      if (q == hipSuccess) {
        const Clock::time_point now = Clock::now();
        for (ServedJob* j : s.jobs) {
          j->status = RNJ_OK;
          j->state = State::Done;
          j->t_done = now;
          used -= j->need - held_after_batch(j);
          // the batch waited for the job's upload and packing: characters and planes are free
          blocks.put(j->chars_dev, chars_bytes(j->n, j->L));
          blocks.put(j->planes, planes_bytes(j->n, j->L));
          j->chars_dev = nullptr;
          j->planes = nullptr;
        }
        s.busy = false;
        s.jobs.clear();
        finished = true;
      } else if (q != hipErrorNotReady) {
        s.error = std::string("batch failed: ") + hipGetErrorString(q);
        s.error_status = RNJ_E_HIP;
        failed.push_back(&s);
      }
    }
    // 2. queued jobs that now fit (FIFO); they are uploaded below, outside the lock
    std::vector<ServedJob*> to_upload;
    while (!queue.empty()) {
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
      if (used + j->need > cfg.hbm_budget) break;
      queue.pop_front();
      used += j->need;
      j->state = State::Uploading;
      j->t_admit = Clock::now();
      to_upload.push_back(j);
    }
    // 3. batches of ready jobs (arrival order) for free slots
    std::vector<Slot*> launching;
    for (Slot& s : slots) {
      if (s.busy || ready.empty() || ready.front()->state != State::Ready) continue;
      if (cfg.batch_window_us && !stopping) {  // gather more jobs while the window is open
        size_t n_ready = 0;
        for (ServedJob* j : ready) n_ready += j->state == State::Ready;
        const auto age = Clock::now() - ready.front()->t_ready;
        if (n_ready < cfg.max_batch_jobs && age < std::chrono::microseconds(cfg.batch_window_us))
          break;
      }
      s.jobs.clear();
      const int kind = ready.front()->kind;
      const unsigned bits = ready.front()->code_bits;
      unsigned long long work = 0;
      while (!ready.empty() && ready.front()->state == State::Ready &&
             ready.front()->kind == kind && ready.front()->code_bits == bits &&
             s.jobs.size() < cfg.max_batch_jobs) {
        ServedJob* j = ready.front();
        const unsigned long long w = job_work(j->n, j->W);
        if (!s.jobs.empty() && work + w > cfg.max_batch_work) break;
        work += w;
        ready.pop_front();
        j->state = State::Running;
        s.jobs.push_back(j);
      }
      s.batch_id = next_batch++;
      for (ServedJob* j : s.jobs) {
        j->batch_id = s.batch_id;
        j->batch_jobs = (unsigned)s.jobs.size();
      }
      s.busy = true;
      launching.push_back(&s);
    }
    bool busy = false;
    for (Slot& s : slots) busy = busy || s.busy;
    if (finished) cv_done.notify_all();
    const bool idle = queue.empty() && ready.empty() && to_upload.empty() && failed.empty();
    if (stopping && idle && !busy) break;
    lock.unlock();

    // 4. host-side work without the lock
    for (ServedJob* j : to_upload) {
      j->status = upload(j, j->chars.data(), j->L);
      j->chars.clear();
      j->chars.shrink_to_fit();
    }
    for (Slot* s : launching) {
      if (!launch(*s)) failed.push_back(s);
      const Clock::time_point t = Clock::now();
      for (ServedJob* j : s->jobs) j->t_launch = t;
    }

    lock.lock();
    const Clock::time_point now = Clock::now();
    for (ServedJob* j : to_upload) {
      if (j->status == RNJ_OK) {
        j->state = State::Ready;
        j->t_ready = now;
        ready.push_back(j);
      } else {
        release_job_memory(j);  // the failed upload synchronised its stream
        used -= j->need;
        j->state = State::Done;
        j->t_done = now;
      }
    }
    for (Slot* s : failed) {
      (void)hipStreamSynchronize(s->stream);  // nothing of the batch still runs
      for (ServedJob* j : s->jobs) {
        j->status = s->error_status;
        j->error = s->error;
        j->state = State::Done;
        j->t_done = now;
        release_job_memory(j);
        used -= j->need;
      }
      s->busy = false;
      s->jobs.clear();
    }
    if (!failed.empty() || !to_upload.empty()) cv_done.notify_all();
    bool busy_now = false;
    for (Slot& s : slots) busy_now = busy_now || s.busy;
    if (busy_now || !ready.empty())  // a batch, an upload or a batch window in flight
      cv_work.wait_for(lock, std::chrono::microseconds(50));
    else if (!stopping && queue.empty())
      cv_work.wait(lock, [this] { return stopping || !queue.empty() || !ready.empty(); });
  }
}

extern "C" {

int rnj_server_create(rnj_context* ctx, const rnj_server_config* cfg, rnj_server** out) {
  if (!ctx || !out) return fail(RNJ_E_INVALID, "null argument");
  RNJ_HIP(hipSetDevice(ctx->device));
  auto* srv = new rnj_server();
  srv->ctx = ctx;
  if (cfg) srv->cfg = *cfg;
  // at most as many slots as the device's hardware queues per process (4 on MI355X boxes):
  // more streams share queues and gave no more overlap
  if (srv->cfg.slots == 0) srv->cfg.slots = kMaxSlots;
  srv->cfg.slots = std::min(srv->cfg.slots, kMaxSlots);
  if (srv->cfg.max_batch_work == 0) srv->cfg.max_batch_work = kDefaultBatchWork;
  if (srv->cfg.max_batch_jobs == 0) srv->cfg.max_batch_jobs = 64;
  srv->cfg.max_batch_jobs = std::min(srv->cfg.max_batch_jobs, kMaxLaunchJobs);
  if (srv->cfg.hbm_budget == 0) {
    size_t free_b = 0, total_b = 0;
    hipError_t e = hipMemGetInfo(&free_b, &total_b);
    if (e != hipSuccess) {
      delete srv;
      return hip_fail(e, "hipMemGetInfo");
    }
    srv->cfg.hbm_budget = (size_t)(0.8 * (double)free_b);
  }
  srv->blocks.set_cap(srv->cfg.hbm_budget / 4);
  srv->slots.resize(srv->cfg.slots);
  hipError_t e = hipStreamCreateWithFlags(&srv->upload_stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&srv->download_stream, hipStreamNonBlocking);
  for (Slot& s : srv->slots) {
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&s.done, hipEventDisableTiming);
  }
  if (e != hipSuccess) {
    for (Slot& t : srv->slots) {
      if (t.stream) (void)hipStreamDestroy(t.stream);
      if (t.done) (void)hipEventDestroy(t.done);
    }
    if (srv->upload_stream) (void)hipStreamDestroy(srv->upload_stream);
    if (srv->download_stream) (void)hipStreamDestroy(srv->download_stream);
    delete srv;
    return hip_fail(e, "rnj_server_create");
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
  if (srv->ctx->char_rules != RNJ_CHARS_BITS)
    return fail(RNJ_E_INVALID,
                "served jobs count under the bit-string rules only (the server's context has the "
                "naive character rules set)");
  if (alignment_kind(srv->ctx, chars, L) != kDNA)
    return fail(RNJ_E_NOT_DNA, "served jobs are DNA alignments only");
  auto* j = new ServedJob();
  j->n = n;
  j->L = L;
  j->n_pad = round_up(n, kTaxaPad);
  j->W = words_for_sites(L);
  j->model = model;
  j->need = job_bytes(n, L);
  j->t_submit = Clock::now();
  bool admit_now = false;
  {
    std::lock_guard<std::mutex> lock(srv->mu);
    if (srv->stopping) {
      delete j;
      return fail(RNJ_E_INVALID, "server is shutting down");
    }
    j->id = srv->next_id++;
    srv->jobs[j->id] = j;
    *ticket = j->id;
    if (srv->queue.empty() && j->need <= srv->cfg.hbm_budget &&
        srv->used + j->need <= srv->cfg.hbm_budget) {
      admit_now = true;
      srv->used += j->need;
      j->state = State::Uploading;
      j->t_admit = j->t_submit;
      srv->ready.push_back(j);  // keeps arrival order; batched once Ready
    }
  }
  if (admit_now) {
    (void)hipSetDevice(srv->ctx->device);
    const int rc = srv->upload(j, chars, pitch);
    std::lock_guard<std::mutex> lock(srv->mu);
    if (rc == RNJ_OK) {
      j->state = State::Ready;
      j->t_ready = Clock::now();
    } else {
      for (auto it = srv->ready.begin(); it != srv->ready.end(); ++it)
        if (*it == j) {
          srv->ready.erase(it);
          break;
        }
      srv->release_job_memory(j);  // the failed upload synchronised its stream
      srv->used -= j->need;
      j->status = rc;
      j->state = State::Done;
      j->t_done = Clock::now();
      srv->cv_done.notify_all();
    }
  } else {
    j->chars.resize((size_t)n * L);
    for (unsigned t = 0; t < n; ++t)
      std::memcpy(&j->chars[(size_t)t * L], chars + (size_t)t * pitch, L);
    std::lock_guard<std::mutex> lock(srv->mu);
    srv->queue.push_back(j);
  }
  srv->cv_work.notify_one();
  return RNJ_OK;
}

int rnj_server_submit_resident(rnj_server* srv, const rnj_alignment* aln, int model, float* tri_dev,
                               uint64_t* ticket) {
  if (!srv || !aln || !ticket) return fail(RNJ_E_INVALID, "null argument");
  if (!aln->planes) return fail(RNJ_E_INVALID, "null alignment");
  if (aln->ctx != srv->ctx) return fail(RNJ_E_INVALID, "alignment belongs to another context");
  if (aln->n < 2) return fail(RNJ_E_INVALID, "need at least 2 sequences");
  if (model != RNJ_JC && model != RNJ_K2P) return fail(RNJ_E_INVALID, "model must be RNJ_JC or RNJ_K2P");
  auto* j = new ServedJob();
  j->n = aln->n;
  j->L = aln->L;
  j->n_pad = aln->n_pad;
  j->W = aln->W;
  j->model = model;
  j->kind = aln->kind;
  j->code_bits = aln->kind == kProtein ? aln->code_bits : 0;
  j->aln = aln;
  j->own_tri = tri_dev == nullptr;
  j->tri = tri_dev;
  j->need = resident_job_bytes(j->n, j->own_tri);
  j->t_submit = Clock::now();
  bool admit_now = false;
  {
    std::lock_guard<std::mutex> lock(srv->mu);
    if (srv->stopping) {
      delete j;
      return fail(RNJ_E_INVALID, "server is shutting down");
    }
    j->id = srv->next_id++;
    srv->jobs[j->id] = j;
    *ticket = j->id;
    if (srv->queue.empty() && j->need <= srv->cfg.hbm_budget &&
        srv->used + j->need <= srv->cfg.hbm_budget) {
      admit_now = true;
      srv->used += j->need;
      j->state = State::Uploading;
      j->t_admit = j->t_submit;
      srv->ready.push_back(j);  // keeps arrival order; batched once Ready
    } else {
      srv->queue.push_back(j);  // nothing to copy: the alignment stays the caller's
    }
  }
  if (admit_now) {
    const int rc = srv->upload(j, nullptr, 0);  // a server triangle, if any
    std::lock_guard<std::mutex> lock(srv->mu);
    if (rc == RNJ_OK) {
      j->state = State::Ready;
      j->t_ready = Clock::now();
    } else {
      for (auto it = srv->ready.begin(); it != srv->ready.end(); ++it)
        if (*it == j) {
          srv->ready.erase(it);
          break;
        }
      srv->used -= j->need;
      j->status = rc;
      j->state = State::Done;
      j->t_done = Clock::now();
      srv->cv_done.notify_all();
    }
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
  int status = j->status;
  const Clock::time_point copy_start = Clock::now();
  if (j->tri) {  // finished: the triangle is resident; copy it once into the caller's buffer
    hipError_t e = hipSuccess;
    if (status == RNJ_OK && tri_out)
      e = srv->downloader.download(tri_out, j->tri, tri_bytes(j->n), srv->download_stream);
    if (j->own_tri) {
      srv->blocks.put(j->tri, tri_bytes(j->n));  // the download is complete (or its stream drained)
      std::lock_guard<std::mutex> relock(srv->mu);
      srv->used -= held_after_batch(j);
    }
    j->tri = nullptr;
    srv->cv_work.notify_one();  // memory came back: queued jobs may fit now
    if (e != hipSuccess) {
      status = RNJ_E_HIP;
      j->error = std::string("download: ") + hipGetErrorString(e);
    }
  }
  const Clock::time_point copied = Clock::now();
  if (stats) {
    auto sec = [](Clock::duration d) { return std::chrono::duration<double>(d).count(); };
    const bool admitted = j->t_admit.time_since_epoch().count() != 0;
    stats->queued_s = admitted ? sec(j->t_admit - j->t_submit) : sec(j->t_done - j->t_submit);
    stats->device_s = admitted ? sec(copied - j->t_admit) : 0.0;
    stats->total_s = sec(copied - j->t_submit);
    stats->batch_id = j->batch_id;
    stats->batch_jobs = j->batch_jobs;
    const bool ran = j->t_launch.time_since_epoch().count() != 0;
    stats->upload_s = admitted && ran ? sec(j->t_ready - j->t_admit) : 0.0;
    stats->launch_s = ran ? sec(j->t_launch - j->t_ready) : 0.0;
    stats->run_s = ran ? sec(j->t_done - j->t_launch) : 0.0;
    stats->download_s = sec(copied - copy_start);
    stats->launched_at = ran ? sec(j->t_launch - srv->t0) : -1.0;
    stats->finished_at = ran ? sec(j->t_done - srv->t0) : -1.0;
  }
  if (status != RNJ_OK) set_error(j->error);
  if (j->ready) (void)hipEventDestroy(j->ready);
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
    // jobs not yet admitted are cancelled; admitted ones finish
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
  for (Slot& s : srv->slots)
    if (s.stream) (void)hipStreamSynchronize(s.stream);
  for (hipStream_t st : {srv->upload_stream, srv->download_stream})
    if (st) (void)hipStreamSynchronize(st);
  for (auto& kv : srv->jobs) {  // results never waited for
    srv->release_job_memory(kv.second);
    if (kv.second->ready) (void)hipEventDestroy(kv.second->ready);
    delete kv.second;
  }
  srv->blocks.trim();
  for (Slot& s : srv->slots) {
    if (s.stream) (void)hipStreamDestroy(s.stream);
    if (s.done) (void)hipEventDestroy(s.done);
  }
  for (hipStream_t st : {srv->upload_stream, srv->download_stream})
    if (st) (void)hipStreamDestroy(st);
  delete srv;
  return RNJ_OK;
}

}  // extern "C"
