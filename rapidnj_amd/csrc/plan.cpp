// Captured batched-bootstrap plans: the whole multi-replicate launch as one HIP graph.
//
// rnj_bootstrap_plan_create allocates every buffer the launch needs, derives the alignment's
// per-site codes once, runs the launch eagerly once (which also settles the tile configuration)
// and then captures it into a graph on a stream of its own:
//   H2D copy of the column weights (from the plan's pinned buffer)
//   -> bootstrap_index (expanded column lists) -> bootstrap_gather (replicate planes)
//   -> statistics reset -> tile kernel (-> finish kernel) -> saturation patch
// (the job table, which never changes, is uploaded once by the eager run).
// Every replicate uses the word count of the plan's maximum replicate length; the gather writes
// zero words past each replicate's own length, so shorter replicates count exactly their sites.
// The body is captured twice, once per pinned weight buffer. rnj_bootstrap_plan_run copies new
// weights into the buffer of its turn (after the replay that last read it, two runs back, has
// finished) and replays that buffer's graph behind the previous replay (a device-side wait, so
// runs on different streams still follow one another while the host prepares the next run).
// The graph holds no memset node: statistics are cleared by a kernel.
#include <cstring>

#include "runtime.hpp"

using namespace rnj;

struct rnj_plan {
  Context* ctx = nullptr;
  const rnj_alignment* aln = nullptr;
  unsigned R = 0, L = 0, Lmax = 0, Wr = 0;
  uint16_t* h_weights[2] = {nullptr, nullptr};  // pinned, one per captured graph
  uint16_t* d_weights = nullptr;
  unsigned* d_idx = nullptr;
  unsigned* d_len = nullptr;
  void* d_codes = nullptr;
  uint32_t* d_planes = nullptr;
  size_t rep_words = 0;
  Workspace ws;
  std::vector<JobSpec> specs;
  LaunchShape shape;
  int config = 0;
  bool raw = false, exact = false;
  hipStream_t stream = nullptr;
  hipGraph_t graph[2] = {nullptr, nullptr};
  hipGraphExec_t exec[2] = {nullptr, nullptr};
  hipEvent_t done[2] = {nullptr, nullptr};  // recorded after the replay of each graph
  int turn = 0;                             // graph of the next run
};

namespace {

void destroy_plan(rnj_plan* p) {
  if (!p) return;
  if (p->ctx) (void)hipSetDevice(p->ctx->device);
  if (p->stream) (void)hipStreamSynchronize(p->stream);
  for (int k = 0; k < 2; ++k) {
    if (p->done[k]) (void)hipEventSynchronize(p->done[k]);
    if (p->exec[k]) (void)hipGraphExecDestroy(p->exec[k]);
    if (p->graph[k]) (void)hipGraphDestroy(p->graph[k]);
    if (p->h_weights[k]) (void)hipHostFree(p->h_weights[k]);
  }
  for (void* d : {(void*)p->d_weights, (void*)p->d_idx, (void*)p->d_len, (void*)p->d_codes,
                  (void*)p->d_planes})
    if (d) (void)hipFree(d);
  free_workspace_public(&p->ws);
  for (hipEvent_t ev : p->done)
    if (ev) (void)hipEventDestroy(ev);
  if (p->stream) (void)hipStreamDestroy(p->stream);
  delete p;
}

// The launch body with pinned weight buffer `buf`, enqueued on p->stream (eagerly or under
// capture). The job table never changes: the eager run uploads it, captures leave the copy out.
int enqueue_plan(rnj_plan* p, int buf, bool upload_table) {
  const rnj_alignment* a = p->aln;
  RNJ_HIP(hipMemcpyAsync(p->d_weights, p->h_weights[buf], (size_t)p->R * p->L * sizeof(uint16_t),
                         hipMemcpyHostToDevice, p->stream));
  RNJ_HIP(launch_bootstrap_index(p->d_weights, p->R, p->L, p->d_idx, p->Lmax, p->d_len, p->stream));
  RNJ_HIP(launch_replicate_gather(a, p->d_codes, p->d_idx, p->Lmax, p->d_len, p->R, p->Wr,
                                  p->d_planes, p->rep_words, p->stream));
  return enqueue_launch(p->ctx, p->specs, p->shape, p->config, &p->ws, p->raw, p->exact,
                        p->stream, upload_table);
}

int check_weights(const rnj_plan* p, const uint16_t* weights) {
  for (unsigned r = 0; r < p->R; ++r) {
    unsigned long long sum = 0;
    for (unsigned c = 0; c < p->L; ++c) sum += weights[(size_t)r * p->L + c];
    if (sum > p->Lmax)
      return fail(RNJ_E_INVALID, "replicate " + std::to_string(r) + " has " + std::to_string(sum) +
                                     " columns, more than the plan's maximum " +
                                     std::to_string(p->Lmax));
  }
  return RNJ_OK;
}

}  // namespace

extern "C" {

int rnj_bootstrap_plan_create(rnj_context* ctx, const rnj_alignment* aln, int model, unsigned R,
                              unsigned max_length, int layout, float* out, size_t matrix_stride,
                              size_t ld, int flags, rnj_plan** plan) {
  if (!ctx || !aln || !out || !plan) return fail(RNJ_E_INVALID, "null argument");
  if (aln->ctx != ctx) return fail(RNJ_E_INVALID, "alignment belongs to another context");
  if (model != RNJ_JC && model != RNJ_K2P) return fail(RNJ_E_INVALID, "model must be RNJ_JC or RNJ_K2P");
  if (layout != RNJ_TRIANGLE && layout != RNJ_SQUARE) return fail(RNJ_E_INVALID, "unknown layout");
  if (R == 0) return fail(RNJ_E_INVALID, "a plan needs at least one replicate");
  if (R > kMaxLaunchJobs) return fail(RNJ_E_INVALID, "more than 65535 replicates in one plan");
  const unsigned n = aln->n, L = aln->L;
  if (layout == RNJ_SQUARE && ld < n) return fail(RNJ_E_INVALID, "ld smaller than n");
  const size_t min_stride = layout == RNJ_SQUARE ? (size_t)n * ld : (size_t)n * (n - 1) / 2;
  if (matrix_stride < min_stride) return fail(RNJ_E_INVALID, "matrix_stride too small");
  RNJ_HIP(hipSetDevice(ctx->device));
  auto* p = new rnj_plan();
  p->ctx = ctx;
  p->aln = aln;
  p->R = R;
  p->L = L;
  p->Lmax = max_length ? max_length : L;
  p->Wr = words_for_sites(p->Lmax);
  p->raw = (flags & RNJ_FLAG_RAW) != 0;
  p->exact = (flags & RNJ_FLAG_EXACT) != 0;
  p->rep_words = replicate_words(aln, p->Wr);
  hipError_t e = hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking);
  for (int k = 0; k < 2; ++k) {
    if (e == hipSuccess) e = hipEventCreateWithFlags(&p->done[k], hipEventDisableTiming);
    if (e == hipSuccess)
      e = hipHostMalloc((void**)&p->h_weights[k], (size_t)R * L * sizeof(uint16_t),
                        hipHostMallocDefault);
  }
  if (e == hipSuccess) e = hipMalloc((void**)&p->d_weights, (size_t)R * L * sizeof(uint16_t));
  if (e == hipSuccess) e = hipMalloc((void**)&p->d_idx, (size_t)R * p->Lmax * sizeof(unsigned));
  if (e == hipSuccess) e = hipMalloc((void**)&p->d_len, (size_t)R * sizeof(unsigned));
  if (e == hipSuccess) e = hipMalloc(&p->d_codes, alignment_codes_bytes(aln));
  if (e == hipSuccess) e = hipMalloc((void**)&p->d_planes, (size_t)R * p->rep_words * 4);
  if (e == hipSuccess) e = launch_alignment_codes(aln, p->d_codes, p->stream);
  if (e != hipSuccess) {
    destroy_plan(p);
    return hip_fail(e, "rnj_bootstrap_plan_create");
  }
  for (unsigned r = 0; r < R; ++r)  // every replicate holds at most Lmax sites; a protein
    // replicate's own length is read on the device, as it changes with the weights of each run
    p->specs.push_back(replicate_spec(aln, p->d_planes + (size_t)r * p->rep_words, p->Wr, p->Lmax,
                                      p->d_len + r, model, layout, out + (size_t)r * matrix_stride,
                                      ld));
  p->shape = shape_of(p->specs);
  int rc = prepare_workspace(&p->ws, R);
  // the original alignment as every replicate: a valid eager run that settles the configuration
  for (size_t k = 0; k < (size_t)R * L; ++k) p->h_weights[0][k] = 1;
  if (rc == RNJ_OK)
    rc = select_config(ctx, p->specs, p->shape, &p->ws, p->exact, p->stream, true, &p->config);
  if (rc == RNJ_OK) rc = enqueue_plan(p, 0, true);
  if (rc == RNJ_OK && hipStreamSynchronize(p->stream) != hipSuccess)
    rc = fail(RNJ_E_HIP, "eager plan run failed");
  for (int k = 0; k < 2 && rc == RNJ_OK; ++k) {
    e = hipStreamBeginCapture(p->stream, hipStreamCaptureModeRelaxed);
    if (e != hipSuccess) {
      rc = hip_fail(e, "hipStreamBeginCapture");
      break;
    }
    const int body = enqueue_plan(p, k, false);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(p->stream, &g);
    p->graph[k] = g;
    if (body != RNJ_OK) rc = body;
    else if (e != hipSuccess) rc = hip_fail(e, "hipStreamEndCapture");
    if (rc == RNJ_OK) {
      e = hipGraphInstantiate(&p->exec[k], p->graph[k], nullptr, nullptr, 0);
      if (e != hipSuccess) rc = hip_fail(e, "hipGraphInstantiate");
    }
  }
  if (rc != RNJ_OK) {
    destroy_plan(p);
    return rc;
  }
  *plan = p;
  return RNJ_OK;
}

int rnj_bootstrap_plan_run(rnj_plan* p, const uint16_t* weights, void* stream) {
  if (!p || !weights) return fail(RNJ_E_INVALID, "null argument");
  int rc = check_weights(p, weights);
  if (rc != RNJ_OK) return rc;
  RNJ_HIP(hipSetDevice(p->ctx->device));
  const int k = p->turn;
  RNJ_HIP(hipEventSynchronize(p->done[k]));  // the replay two runs back has read buffer k
  std::memcpy(p->h_weights[k], weights, (size_t)p->R * p->L * sizeof(uint16_t));
  hipStream_t s = (hipStream_t)stream;
  RNJ_HIP(hipStreamWaitEvent(s, p->done[1 - k], 0));  // after the previous replay, on any stream
  RNJ_HIP(hipGraphLaunch(p->exec[k], s));
  RNJ_HIP(hipEventRecord(p->done[k], s));
  p->turn = 1 - k;
  return RNJ_OK;
}

int rnj_plan_tile_config(const rnj_plan* p) { return p ? p->config : -1; }

int rnj_plan_destroy(rnj_plan* p) {
  destroy_plan(p);
  return RNJ_OK;
}

}  // extern "C"
