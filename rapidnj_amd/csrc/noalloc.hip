// rnj_distances_device_noalloc: rnj_distances_device for callers that cannot allocate device
// memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
//
// A distance launch needs scratch beyond its input planes and output matrix: the job table, the
// job's maximum, the per-wave saturation records and, for launches cut into word ranges, the
// partial counts. Here all of it lives in `scratch`, a file-scope device array of
// kNoAllocScratchBytes: every GPU the library runs on holds its own copy until the process ends.
// Host code never uses the array's name as an address (in host code it is only a stand-in for the
// device symbol): scratch_setup, called by rnj_create once per device with that device current,
// takes the device address with hipGetSymbolAddress and keeps it per device, and the entry point
// hands that address to the kernels.
//
// The entry point calls no hipMalloc, hipFree or other allocating routine, and does not allocate
// host memory on its success path: the job is planned on the stack and its description reaches
// the device as a kernel argument. What must start at zero is cleared by the launch itself (the
// job maximum, by reset_stats_kernel); the saturation records and partial counts are written in
// full before they are read. Every call on one GPU shares that GPU's array, so calls that
// overlap on one GPU (on different streams or host threads) race; callers order them.
#include "runtime.hpp"

static __device__ __align__(256) unsigned char scratch[rnj::kNoAllocScratchBytes];

namespace rnj {
namespace {

constexpr int kMaxScratchDevices = 64;
constexpr size_t kScratchAlign = 256;
std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[kMaxScratchDevices] = {};

size_t align_up(size_t x) { return (x + kScratchAlign - 1) / kScratchAlign * kScratchAlign; }

// Offsets of one job's scratch regions, and the bytes they take.
struct ScratchMap {
  JobPlan plan;
  size_t stats = 0, sat = 0, partial = 0, total = 0;
};

ScratchMap scratch_map(const JobSpec& s, int config, unsigned cus) {
  ScratchMap m;
  const unsigned tiles = job_tiles(s, config);
  m.plan = plan_job(s, config, tiles, split_want(tiles, cus));
  m.stats = align_up(sizeof(JobDesc));
  m.sat = m.stats + kScratchAlign;
  m.partial = m.sat + align_up((size_t)m.plan.sat_units * 66 * sizeof(uint32_t));
  m.total = m.partial + m.plan.partial_elems * sizeof(uint32_t);
  return m;
}

}  // namespace

int scratch_setup(int device) {
  if (device < 0 || device >= kMaxScratchDevices)
    return fail(RNJ_E_INVALID, "device index beyond the library's scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device]) return RNJ_OK;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
  if (!p) return fail(RNJ_E_HIP, "hipGetSymbolAddress(scratch) gave no address");
  g_scratch_ptr[device] = static_cast<unsigned char*>(p);
  return RNJ_OK;
}

}  // namespace rnj

using namespace rnj;

extern "C" int rnj_distances_device_noalloc(rnj_context* ctx, const rnj_alignment* aln,
                                            int model, int layout, float* out, size_t ld,
                                            int flags, float* max_out, void* stream) {
  if (!ctx) return fail(RNJ_E_INVALID, "null context");
  if (!aln || !aln->planes) return fail(RNJ_E_INVALID, "null alignment");
  if (aln->ctx != ctx) return fail(RNJ_E_INVALID, "alignment belongs to another context");
  if (model != RNJ_JC && model != RNJ_K2P) return fail(RNJ_E_INVALID, "model must be RNJ_JC or RNJ_K2P");
  if (!out) return fail(RNJ_E_INVALID, "null output");
  if (layout != RNJ_TRIANGLE && layout != RNJ_SQUARE) return fail(RNJ_E_INVALID, "unknown layout");
  if (layout == RNJ_SQUARE && ld < aln->n) return fail(RNJ_E_INVALID, "ld smaller than n");
  (void)hipGetLastError();
  RNJ_HIP(hipSetDevice(ctx->device));
  unsigned char* base = nullptr;
  if (ctx->device >= 0 && ctx->device < kMaxScratchDevices) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    base = g_scratch_ptr[ctx->device];
  }
  if (!base) return fail(RNJ_E_INVALID, "no scratch address for this device");

  JobSpec spec = alignment_spec(aln, model, layout, out, ld);
  LaunchShape sh;
  sh.kind = spec.kind;
  sh.code_bits = spec.code_bits;
  sh.jc_only = model == RNJ_JC;
  sh.n_key = spec.n;
  sh.W_key = spec.W;
  const unsigned cus = (unsigned)ctx->prop.multiProcessorCount;
  // the configuration the launch would get anyway (never tuned here: tuning times launches);
  // when its scratch does not fit, the default configuration, whose records are the smallest
  int config = current_config(ctx, sh);
  ScratchMap m = scratch_map(spec, config, cus);
  if (m.total > kNoAllocScratchBytes && config != kDefaultTileConfig) {
    config = kDefaultTileConfig;
    m = scratch_map(spec, config, cus);
  }
  if (m.total > kNoAllocScratchBytes)
    return fail(RNJ_E_INVALID,
                "a " + std::to_string(spec.n) + "-taxon matrix needs " +
                    std::to_string(m.total) + " bytes of launch scratch; the no-allocation entry "
                    "point has " + std::to_string(kNoAllocScratchBytes) +
                    " (about 22000 taxa); use rnj_distances_device");
  if ((unsigned long long)m.plan.tiles * m.plan.splits >= (1ull << 31) ||
      (spec.n + kTileConfigs[config].tile - 1) / kTileConfigs[config].tile > 65535u)
    return fail(RNJ_E_INVALID, "launch too large");

  JobDesc* d_job = reinterpret_cast<JobDesc*>(base);
  unsigned* d_max = reinterpret_cast<unsigned*>(base + m.stats);
  uint32_t* d_sat = reinterpret_cast<uint32_t*>(base + m.sat);
  uint32_t* d_partial = reinterpret_cast<uint32_t*>(base + m.partial);
  JobDesc d;
  describe_job(d, spec, config, m.plan, 0, 0, d_partial, d_max, d_sat, m.plan.sat_units, 0);
  const bool raw = (flags & RNJ_FLAG_RAW) != 0, exact = (flags & RNJ_FLAG_EXACT) != 0;
  hipStream_t s = (hipStream_t)stream;
  hipError_t e = launch_store_job(d, d_job, s);
  if (e == hipSuccess) e = launch_reset_stats(d_job, 1, s);
  if (e == hipSuccess)
    e = launch_counting(config, sh, exact, d_job, 1, m.plan.tiles * m.plan.splits,
                        m.plan.splits > 1 ? m.plan.tiles : 0, s);
  if (e == hipSuccess) e = launch_saturation(d_job, 1, m.plan.sat_units, config, !raw, s);
  if (e == hipSuccess && max_out) e = launch_copy_max(d_max, max_out, s);
  if (e != hipSuccess) return hip_fail(e, "rnj_distances_device_noalloc launch");
  return RNJ_OK;
}
