// Captured multi-alignment launches: the distance matrices of several alignments of any taxa and
// site counts as ONE tile launch, captured once into a HIP graph and replayed.
//
// rnj_batch_plan_create builds the launch's job table (one job per alignment: the tile kernel
// finds its job by binary search over the jobs' first work items, so jobs of different n, L and
// model share one grid), lets the autotuner settle the tile configuration on an eager run, which
// also uploads the table, and captures the launch without the table copy on a stream of its own:
//   statistics reset -> tile kernel (-> finish kernel for split jobs) -> saturation patch.
// The graph holds no memset node. The alignments' planes are read where they lie, so a replay
// after rnj_alignment_update_chars computes the new characters' matrices; the alignments and the
// output buffers must outlive the plan. Each replay follows the previous one on the device (an
// event wait), so replays on different streams do not overlap on the plan's scratch.
#include "runtime.hpp"

using namespace rnj;

struct rnj_batch_plan {
  Context* ctx = nullptr;
  std::vector<JobSpec> specs;
  LaunchShape shape;
  int config = 0;
  bool raw = false, exact = false;
  Workspace ws;
  hipStream_t stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipEvent_t done = nullptr;  // recorded after each replay
};

namespace {

void destroy_batch_plan(rnj_batch_plan* p) {
  if (!p) return;
  if (p->ctx) (void)hipSetDevice(p->ctx->device);
  if (p->stream) (void)hipStreamSynchronize(p->stream);
  if (p->done) (void)hipEventSynchronize(p->done);
  if (p->exec) (void)hipGraphExecDestroy(p->exec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  free_workspace_public(&p->ws);
  if (p->done) (void)hipEventDestroy(p->done);
  if (p->stream) (void)hipStreamDestroy(p->stream);
  delete p;
}

}  // namespace

extern "C" {

int rnj_batch_plan_create(rnj_context* ctx, const rnj_alignment* const* alns, const int* models,
                          unsigned count, int layout, float* const* outs, const size_t* lds,
                          int flags, rnj_batch_plan** plan) {
  if (!ctx || !alns || !models || !outs || !plan) return fail(RNJ_E_INVALID, "null argument");
  if (count == 0) return fail(RNJ_E_INVALID, "a batch plan needs at least one alignment");
  if (count > kMaxLaunchJobs) return fail(RNJ_E_INVALID, "more than 65535 alignments in one plan");
  if (layout != RNJ_TRIANGLE && layout != RNJ_SQUARE) return fail(RNJ_E_INVALID, "unknown layout");
  if (layout == RNJ_SQUARE && !lds) return fail(RNJ_E_INVALID, "square layout needs row strides");
  std::vector<JobSpec> specs;
  for (unsigned k = 0; k < count; ++k) {
    const rnj_alignment* a = alns[k];
    if (!a || !a->planes) return fail(RNJ_E_INVALID, "null alignment in the batch");
    if (a->ctx != ctx) return fail(RNJ_E_INVALID, "alignment belongs to another context");
    if (models[k] != RNJ_JC && models[k] != RNJ_K2P)
      return fail(RNJ_E_INVALID, "model must be RNJ_JC or RNJ_K2P");
    if (!outs[k]) return fail(RNJ_E_INVALID, "null output buffer in the batch");
    const size_t ld = layout == RNJ_SQUARE ? lds[k] : 0;
    if (layout == RNJ_SQUARE && ld < a->n) return fail(RNJ_E_INVALID, "ld smaller than n");
    specs.push_back(alignment_spec(a, models[k], layout, outs[k], ld));
  }
  LaunchShape shape = shape_of(specs);
  if (shape.mixed) return fail(RNJ_E_INVALID, mixed_message(shape));
  RNJ_HIP(hipSetDevice(ctx->device));
  auto* p = new rnj_batch_plan();
  p->ctx = ctx;
  p->specs = std::move(specs);
  p->shape = shape;
  p->raw = (flags & RNJ_FLAG_RAW) != 0;
  p->exact = (flags & RNJ_FLAG_EXACT) != 0;
  hipError_t e = hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&p->done, hipEventDisableTiming);
  if (e != hipSuccess) {
    destroy_batch_plan(p);
    return hip_fail(e, "rnj_batch_plan_create");
  }
  // the alignments may have been packed on other streams of the context
  (void)hipStreamSynchronize(ctx->stream);
  int rc = prepare_workspace(&p->ws, count);
  if (rc == RNJ_OK)
    rc = select_config(ctx, p->specs, p->shape, &p->ws, p->exact, p->stream, true, &p->config);
  if (rc == RNJ_OK)  // eager run: uploads the job table the graph keeps using
    rc = enqueue_launch(ctx, p->specs, p->shape, p->config, &p->ws, p->raw, p->exact, p->stream);
  if (rc == RNJ_OK && hipStreamSynchronize(p->stream) != hipSuccess)
    rc = fail(RNJ_E_HIP, "eager batch plan run failed");
  if (rc == RNJ_OK) {
    e = hipStreamBeginCapture(p->stream, hipStreamCaptureModeRelaxed);
    if (e != hipSuccess) rc = hip_fail(e, "hipStreamBeginCapture");
  }
  if (rc == RNJ_OK) {
    const int body = enqueue_launch(ctx, p->specs, p->shape, p->config, &p->ws, p->raw, p->exact,
                                    p->stream, false);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(p->stream, &g);
    p->graph = g;
    if (body != RNJ_OK) rc = body;
    else if (e != hipSuccess) rc = hip_fail(e, "hipStreamEndCapture");
  }
  if (rc == RNJ_OK) {
    e = hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) rc = hip_fail(e, "hipGraphInstantiate");
  }
  if (rc != RNJ_OK) {
    destroy_batch_plan(p);
    return rc;
  }
  *plan = p;
  return RNJ_OK;
}

int rnj_batch_plan_run(rnj_batch_plan* p, void* stream) {
  if (!p) return fail(RNJ_E_INVALID, "null plan");
  RNJ_HIP(hipSetDevice(p->ctx->device));
  hipStream_t s = (hipStream_t)stream;
  RNJ_HIP(hipStreamWaitEvent(s, p->done, 0));  // after the previous replay, on any stream
  RNJ_HIP(hipGraphLaunch(p->exec, s));
  RNJ_HIP(hipEventRecord(p->done, s));
  return RNJ_OK;
}

int rnj_batch_plan_tile_config(const rnj_batch_plan* p) { return p ? p->config : -1; }

int rnj_batch_plan_destroy(rnj_batch_plan* p) {
  destroy_batch_plan(p);
  return RNJ_OK;
}

}  // extern "C"
