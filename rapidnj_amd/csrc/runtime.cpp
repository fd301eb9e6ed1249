// Host runtime of librapidnj_amd.so: contexts, alignments resident in HBM, multi-job launches,
// the handle-based C ABI and the reference's own device entry points.
#include "runtime.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <thread>
#include <condition_variable>
#include <unistd.h>
#include <sys/stat.h>
#include <fcntl.h>
#include <cstdlib>
#include <cstring>

namespace rnj {

// ---- errors ---------------------------------------------------------------------------------
namespace {
thread_local std::string g_last_error;
}

void set_error(const std::string& msg) { g_last_error = msg; }
const char* last_error_cstr() { return g_last_error.c_str(); }

int fail(int status, const std::string& msg) {
  g_last_error = msg;
  return status;
}

int hip_fail(hipError_t e, const char* what) {
  g_last_error = std::string(what) + ": " + hipGetErrorString(e);
  return (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? RNJ_E_NOMEM : RNJ_E_HIP;
}

// ---- sizes ----------------------------------------------------------------------------------
unsigned round_up(unsigned x, unsigned m) { return (x + m - 1) / m * m; }

unsigned words_for_sites(unsigned sites) {
  const unsigned w = (sites + 31) / 32;
  return round_up(w == 0 ? 1 : w, kWordPad);
}

size_t saturation_bytes(unsigned n) {
  size_t most = 0;
  for (int c = 0; c < kNumTileConfigs; ++c) {
    const unsigned long long nb = (n + kTileConfigs[c].tile - 1) / kTileConfigs[c].tile;
    most = std::max<size_t>(most, sat_units(nb * (nb + 1) / 2, c) * 66 * sizeof(uint32_t));
  }
  return most;
}

int alloc_alignment(Context* ctx, unsigned n, unsigned L, unsigned sites, Alignment** out,
                    int kind, unsigned code_bits) {
  auto* a = new rnj_alignment();
  a->ctx = ctx;
  a->n = n;
  a->L = L;
  a->n_pad = round_up(n, kTaxaPad);
  a->W = words_for_sites(sites);
  a->kind = kind;
  a->code_bits = code_bits;
  hipError_t e = hipMalloc((void**)&a->planes, a->bytes());
  if (e == hipSuccess && kind == kProtein) {
    a->gaps = a->planes + (size_t)a->nplanes() * a->W * a->n_pad;
    e = hipMemsetAsync(a->gaps, 0, (size_t)a->n_pad * sizeof(uint32_t), ctx->stream);
  }
  if (e != hipSuccess) {
    if (a->planes) (void)hipFree(a->planes);
    delete a;
    return hip_fail(e, "hipMalloc(alignment planes)");
  }
  {
    std::lock_guard<std::mutex> lock(ctx->mu);
    ctx->alignments.insert(a);
  }
  *out = a;
  return RNJ_OK;
}

// ---- workspaces -------------------------------------------------------------------------------
namespace {

void free_workspace(Workspace* ws) {
  if (ws->h_jobs) (void)hipHostFree(ws->h_jobs);
  if (ws->d_jobs) (void)hipFree(ws->d_jobs);
  if (ws->d_stats) (void)hipFree(ws->d_stats);
  if (ws->d_sat) (void)hipFree(ws->d_sat);
  if (ws->d_partial) (void)hipFree(ws->d_partial);
  ws->h_jobs = nullptr;
  ws->d_jobs = nullptr;
  ws->d_stats = nullptr;
  ws->d_sat = nullptr;
  ws->d_partial = nullptr;
  ws->cap_jobs = 0;
  ws->cap_units = 0;
  ws->cap_partial = 0;
}

// Saturation records for `units` wave units (words, block coordinates, 64 masks each): nothing
// needs clearing, since every epilogue wave writes its word and a flagged wave the rest.
int grow_sat(Workspace* ws, size_t units) {
  if (units <= ws->cap_units) return RNJ_OK;
  const size_t cap = std::max(units, ws->cap_units + ws->cap_units / 2);
  if (ws->d_sat) (void)hipFree(ws->d_sat);
  ws->d_sat = nullptr;
  ws->cap_units = 0;
  RNJ_HIP(hipMalloc((void**)&ws->d_sat, cap * 66 * sizeof(uint32_t)));
  ws->cap_units = cap;
  return RNJ_OK;
}

int grow_partial(Workspace* ws, size_t elems) {
  if (elems <= ws->cap_partial) return RNJ_OK;
  const size_t cap = std::max(elems, ws->cap_partial + ws->cap_partial / 2);
  if (ws->d_partial) (void)hipFree(ws->d_partial);
  ws->d_partial = nullptr;
  ws->cap_partial = 0;
  RNJ_HIP(hipMalloc((void**)&ws->d_partial, cap * sizeof(uint32_t)));
  ws->cap_partial = cap;
  return RNJ_OK;
}

int grow_workspace(Workspace* ws, unsigned njobs) {
  if (njobs > ws->cap_jobs) {
    const unsigned cap = std::max(njobs, 2 * ws->cap_jobs);
    if (ws->h_jobs) (void)hipHostFree(ws->h_jobs);
    if (ws->d_jobs) (void)hipFree(ws->d_jobs);
    if (ws->d_stats) (void)hipFree(ws->d_stats);
    ws->h_jobs = nullptr;
    ws->d_jobs = nullptr;
    ws->d_stats = nullptr;
    ws->cap_jobs = 0;
    RNJ_HIP(hipHostMalloc((void**)&ws->h_jobs, cap * sizeof(JobDesc), hipHostMallocDefault));
    RNJ_HIP(hipMalloc((void**)&ws->d_jobs, cap * sizeof(JobDesc)));
    RNJ_HIP(hipMalloc((void**)&ws->d_stats, cap * sizeof(unsigned)));
    ws->cap_jobs = cap;
  }
  return RNJ_OK;
}

int acquire_workspace(Context* ctx, unsigned njobs, Workspace** out) {
  std::lock_guard<std::mutex> lock(ctx->mu);
  Workspace* ws = nullptr;
  for (auto& w : ctx->pool) {
    if (w->in_use) continue;
    if (w->recorded && hipEventQuery(w->done) != hipSuccess) continue;
    ws = w.get();
    break;
  }
  if (!ws) {
    ctx->pool.emplace_back(new Workspace());
    ws = ctx->pool.back().get();
    RNJ_HIP(hipEventCreateWithFlags(&ws->done, hipEventDisableTiming));
  }
  int rc = grow_workspace(ws, njobs);
  if (rc != RNJ_OK) return rc;
  ws->in_use = true;
  *out = ws;
  return RNJ_OK;
}

void release_workspace(Context* ctx, Workspace* ws, hipStream_t stream) {
  std::lock_guard<std::mutex> lock(ctx->mu);
  if (hipEventRecord(ws->done, stream) == hipSuccess) ws->recorded = true;
  ws->in_use = false;
}

int ilog2(unsigned long long x) {
  int r = 0;
  while (x > 1) { x >>= 1; ++r; }
  return r;
}

}  // namespace

unsigned job_tiles(const JobSpec& s, int config) {
  const unsigned nb = (s.n + kTileConfigs[config].tile - 1) / kTileConfigs[config].tile;
  // a kRows job also covers the transposed block triangle (job_tile)
  return nb * (nb + 1) / 2 * (s.layout == kRows ? 2u : 1u);
}

unsigned split_want(unsigned long long total_tiles, unsigned cus) {
  const unsigned long long target = 8ull * std::max(1u, cus);
  if (total_tiles == 0 || total_tiles >= target) return 1;
  return (unsigned)std::min<unsigned long long>(16, (target + total_tiles - 1) / total_tiles);
}

JobPlan plan_job(const JobSpec& s, int config, unsigned tiles, unsigned want) {
  const TileConfig& c = kTileConfigs[config];
  JobPlan jp;
  jp.tiles = tiles;
  const unsigned chunks = s.W / c.chunk;
  unsigned S = std::max(1u, std::min(want, chunks / 2));
  const unsigned cps = (chunks + S - 1) / S;  // chunks per split
  jp.splits = (chunks + cps - 1) / cps;
  jp.wps = cps * c.chunk;
  jp.partial_elems = jp.splits > 1 ? (size_t)tiles * jp.splits * 3 * c.tile * c.tile : 0;
  jp.sat_units = s.model != kCounts ? sat_units(tiles, config) : 0;
  return jp;
}

void describe_job(JobDesc& d, const JobSpec& s, int config, const JobPlan& jp,
                  unsigned long long item_base, unsigned long long fin_base, uint32_t* partial,
                  unsigned* maxbits, uint32_t* sat, size_t cap_units, size_t sat_off) {
  std::memset(&d, 0, sizeof d);
  d.planes = s.planes;
  d.plane_stride = (unsigned long long)s.W_pad * s.n_pad;
  d.n = s.n;
  d.n_pad = s.n_pad;
  d.W = s.W;
  d.W_live = s.W_live ? std::min(s.W_live, s.W) : s.W;
  d.nblk = (s.n + kTileConfigs[config].tile - 1) / kTileConfigs[config].tile;
  d.tiles = jp.tiles;
  d.item_base = (unsigned)item_base;
  d.splits = jp.splits;
  d.words_per_split = jp.wps;
  d.fin_base = (unsigned)fin_base;
  d.partial = jp.splits > 1 ? partial : nullptr;
  d.model = s.model;
  d.layout = s.layout;
  d.ld = s.ld;
  d.out = s.out;
  d.out_ts = s.out_ts;
  d.out_tv = s.out_tv;
  d.out_valid = s.out_valid;
  d.gaps = s.gaps;
  d.L = s.L;
  d.bias_rule = s.bias_rule;
  d.len_dev = s.len_dev;
  d.maxbits = maxbits;
  d.sat_word = sat + sat_off;
  d.sat_tile = sat + cap_units + sat_off;
  d.sat_mask = sat + 2 * cap_units + 64 * sat_off;
}

namespace {

// Work decomposition of one launch for one tile configuration. When the launch has fewer tiles
// than about 8 per CU, every job's words are cut into up to 16 ranges (each at least two LDS
// chunks), so that the device fills; the split tiles are finished by the finish kernel.
struct LaunchPlan {
  std::vector<unsigned> splits, wps, tiles;
  std::vector<size_t> partial_off, sat_off;
  size_t partial_elems = 0;
  unsigned long long items = 0, finish_items = 0;
  unsigned long long sat_units = 0, max_units = 0;  // wave units: all jobs, the largest job
};

LaunchPlan plan_launch(const std::vector<JobSpec>& specs, int config, unsigned cus) {
  LaunchPlan P;
  unsigned long long total_tiles = 0;
  for (const JobSpec& s : specs) {
    P.tiles.push_back(job_tiles(s, config));
    total_tiles += P.tiles.back();
  }
  const unsigned want = split_want(total_tiles, cus);
  for (size_t k = 0; k < specs.size(); ++k) {
    const JobPlan jp = plan_job(specs[k], config, P.tiles[k], want);
    P.splits.push_back(jp.splits);
    P.wps.push_back(jp.wps);
    P.partial_off.push_back(P.partial_elems);
    P.partial_elems += jp.partial_elems;
    if (jp.splits > 1) P.finish_items += jp.tiles;
    P.items += (unsigned long long)jp.tiles * jp.splits;
    P.sat_off.push_back(P.sat_units);
    P.sat_units += jp.sat_units;
    P.max_units = std::max(P.max_units, jp.sat_units);
  }
  return P;
}

// Fill the job table for `plan` and bind scratch.
void number_jobs(std::vector<JobSpec>& specs, int config, const LaunchPlan& P, Workspace* ws) {
  unsigned long long base = 0, fin = 0;
  for (size_t k = 0; k < specs.size(); ++k) {
    JobPlan jp;
    jp.tiles = P.tiles[k];
    jp.splits = P.splits[k];
    jp.wps = P.wps[k];
    describe_job(ws->h_jobs[k], specs[k], config, jp, base, fin, ws->d_partial + P.partial_off[k],
                 ws->d_stats + k, ws->d_sat, ws->cap_units, P.sat_off[k]);
    base += (unsigned long long)jp.tiles * jp.splits;
    if (jp.splits > 1) fin += jp.tiles;
  }
}

}  // namespace

void free_workspace_public(Workspace* ws) { free_workspace(ws); }

// Configuration before (or without) tuning, from measurements on MI355X: the 32-taxon tile with
// 8-word chunks was fastest or within 1 % of it on every shape timed (K2P 2000 x 10000: 0.181 ms,
// against 0.191 ms for t32_m4x4_c16 and 0.218 ms for t64_m4x4_c8; K2P 20000 x 5000: 5.59 ms,
// against 5.74 ms for t64_m4x4_c16; JC 20000 x 5000: 4.49 ms, against 4.48 ms for t64_m4x4_c8;
// Kimura protein 20000 x 1250: 2.35 ms, against 2.43 ms for t64_m4x4_c16).
static int default_config(unsigned n, unsigned W, bool jc_only) {
  (void)n;
  (void)W;
  (void)jc_only;
  return kDefaultTileConfig;
}

// Autotune classes: DNA launches by whether every job is JC (the 6-instruction loop), protein
// launches by code width.
static int tune_variant(const LaunchShape& sh) {
  if (sh.kind == kNaive) return 8;
  return sh.kind == kProtein ? 16 + (int)sh.code_bits : (int)sh.jc_only;
}

hipError_t launch_counting(int config, const LaunchShape& sh, bool exact,
                                  const JobDesc* jobs, unsigned njobs, unsigned items,
                                  unsigned finish_items, hipStream_t stream) {
  if (sh.kind == kProtein)
    return launch_protein_tiles(config, sh.code_bits, exact, jobs, njobs, items, finish_items,
                                stream);
  const int mode = sh.kind == kNaive ? kModeNaive : (sh.jc_only ? kModeJC6 : kModeK2P);
  return launch_tiles(config, mode, exact, jobs, njobs, items, finish_items, stream);
}

// The configuration the next launch of this shape class uses (without tuning it).
int current_config(Context* ctx, const LaunchShape& sh) {
  if (ctx->tile_config >= 0) return ctx->tile_config;
  std::lock_guard<std::mutex> lock(ctx->mu);
  auto it = ctx->tuned.find(std::make_tuple(ilog2(sh.n_key), ilog2(sh.W_key), tune_variant(sh)));
  if (it != ctx->tuned.end()) return it->second;
  return default_config(sh.n_key, sh.W_key, sh.jc_only);
}

int choose_config(Context* ctx, unsigned n, unsigned W, bool jc_only) {
  if (ctx->tile_config >= 0) return ctx->tile_config;
  std::lock_guard<std::mutex> lock(ctx->mu);
  auto it = ctx->tuned.find(std::make_tuple(ilog2(n), ilog2(W), (int)jc_only));
  if (it != ctx->tuned.end()) return it->second;
  return default_config(n, W, jc_only);
}

// Time every configuration on this launch (stream-synchronous) and remember the fastest for the
// launch's shape class. Only for launches big enough that the timing means something and only
// outside stream capture. The device is warmed first with two launches of the default (a cold
// device runs its first milliseconds at lower clocks), then every configuration is timed in two
// rounds of two launches, and each keeps its fastest launch, so that a clock change during the
// sweep does not decide the choice.
static int autotune(Context* ctx, std::vector<JobSpec>& specs, Workspace* ws,
                    const LaunchShape& sh, bool exact, hipStream_t stream, int* best) {
  const bool jc_only = sh.jc_only;
  const unsigned n_key = sh.n_key, W_key = sh.W_key;
  hipEvent_t e0, e1;
  RNJ_HIP(hipEventCreate(&e0));
  RNJ_HIP(hipEventCreate(&e1));
  *best = default_config(n_key, W_key, jc_only);
  const unsigned cus = (unsigned)ctx->prop.multiProcessorCount;
  std::vector<float> ms(kNumTileConfigs, 1e30f);
  auto run = [&](int c, bool timed) -> int {
    const LaunchPlan P = plan_launch(specs, c, cus);
    if (P.items >= (1ull << 31)) return RNJ_OK;
    RNJ_HIP_STATUS(grow_partial(ws, P.partial_elems));
    RNJ_HIP_STATUS(grow_sat(ws, P.sat_units));
    number_jobs(specs, c, P, ws);
    RNJ_HIP(hipMemcpyAsync(ws->d_jobs, ws->h_jobs, specs.size() * sizeof(JobDesc),
                           hipMemcpyHostToDevice, stream));
    RNJ_HIP(launch_reset_stats(ws->d_jobs, (unsigned)specs.size(), stream));
    RNJ_HIP(hipEventRecord(e0, stream));
    RNJ_HIP(launch_counting(c, sh, exact, ws->d_jobs, (unsigned)specs.size(), (unsigned)P.items,
                            (unsigned)P.finish_items, stream));
    RNJ_HIP(hipEventRecord(e1, stream));
    RNJ_HIP(hipEventSynchronize(e1));
    float t = 0;
    RNJ_HIP(hipEventElapsedTime(&t, e0, e1));
    if (timed) ms[c] = std::min(ms[c], t);
    return RNJ_OK;
  };
  int rc = RNJ_OK;
  for (int k = 0; k < 2 && rc == RNJ_OK; ++k) rc = run(*best, false);
  for (int round = 0; round < 2 && rc == RNJ_OK; ++round)
    for (int c = 0; c < kNumTileConfigs && rc == RNJ_OK; ++c) {
      if (round == 0) rc = run(c, false);  // first launch of a configuration: code and caches warm
      for (int k = 0; k < 2 && rc == RNJ_OK; ++k) rc = run(c, true);
    }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc != RNJ_OK) return rc;
  float best_ms = 1e30f;
  for (int c = 0; c < kNumTileConfigs; ++c)
    if (ms[c] < best_ms) {
      best_ms = ms[c];
      *best = c;
    }
  // Within 1.5 % of the fastest (timing noise), take the larger tile edge: it re-reads each
  // taxon block from L2/HBM half as often (at 20000 x 5000, 0.78 GB fetched instead of 1.5 GB).
  for (int c = 0; c < kNumTileConfigs; ++c) {
    if (ms[c] > best_ms * 1.015f) continue;
    const unsigned tc = kTileConfigs[c].tile, tb = kTileConfigs[*best].tile;
    if (tc > tb || (tc == tb && ms[c] < ms[*best])) *best = c;
  }
  std::lock_guard<std::mutex> lock(ctx->mu);
  ctx->tuned[std::make_tuple(ilog2(n_key), ilog2(W_key), tune_variant(sh))] = *best;
  return RNJ_OK;
}

LaunchShape shape_of(const std::vector<JobSpec>& specs) {
  LaunchShape sh;
  if (!specs.empty()) {
    sh.kind = specs[0].kind;
    sh.code_bits = specs[0].code_bits;
  }
  for (size_t k = 0; k < specs.size(); ++k) {
    const JobSpec& s = specs[k];
    if (!sh.mixed && (s.kind != sh.kind || s.code_bits != sh.code_bits)) {
      sh.mixed = true;
      sh.mixed_kind = s.kind;
      sh.mixed_bits = s.code_bits;
    }
    if (s.model != kJC) sh.jc_only = false;
    sh.n_key = std::max(sh.n_key, s.n);
    sh.W_key = std::max(sh.W_key, s.W);
    sh.work += (unsigned long long)s.n * s.n * s.W;
  }
  return sh;
}

std::string mixed_message(const LaunchShape& sh) {
  auto name = [](int kind, unsigned bits) -> std::string {
    if (kind == kNaive) return "DNA under the naive character rules";
    if (kind == kDNA) return "DNA under the bit-string rules";
    return "protein with " + std::to_string(bits) + " code bits";
  };
  return "one launch cannot mix " + name(sh.kind, sh.code_bits) + " and " +
         name(sh.mixed_kind, sh.mixed_bits) +
         " (alignments of one launch share the count loop: one character-rule set for DNA, one "
         "code width for protein)";
}

int select_config(Context* ctx, std::vector<JobSpec>& specs, LaunchShape& sh, Workspace* ws,
                  bool exact, hipStream_t stream, bool allow_tune, int* config) {
  *config = ctx->tile_config;
  if (*config >= 0) return RNJ_OK;
  {
    std::lock_guard<std::mutex> lock(ctx->mu);
    auto it = ctx->tuned.find(std::make_tuple(ilog2(sh.n_key), ilog2(sh.W_key), tune_variant(sh)));
    if (it != ctx->tuned.end()) {
      *config = it->second;
      return RNJ_OK;
    }
  }
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  (void)hipStreamIsCapturing(stream, &cap);
  // tune only launches worth >= 2^28 taxon^2-words (~0.1 ms of counting): smaller ones keep
  // the default
  if (allow_tune && cap == hipStreamCaptureStatusNone && sh.work >= (1ull << 28))
    return autotune(ctx, specs, ws, sh, exact, stream, config);
  *config = default_config(sh.n_key, sh.W_key, sh.jc_only);
  return RNJ_OK;
}

int prepare_workspace(Workspace* ws, unsigned njobs) { return grow_workspace(ws, njobs); }

int enqueue_launch(Context* ctx, std::vector<JobSpec>& specs, LaunchShape& sh, int config,
                   Workspace* ws, bool raw, bool exact, hipStream_t stream, bool upload_table) {
  if (sh.mixed) return fail(RNJ_E_INVALID, mixed_message(sh));
  if (specs.size() > kMaxLaunchJobs) return fail(RNJ_E_INVALID, "more than 65535 jobs in one launch");
  const LaunchPlan P = plan_launch(specs, config, (unsigned)ctx->prop.multiProcessorCount);
  if (P.items >= (1ull << 31)) return fail(RNJ_E_INVALID, "launch too large: more than 2^31 pair tiles");
  for (const JobSpec& s : specs)  // saturation records keep block coordinates in 16 bits each
    if ((s.n + kTileConfigs[config].tile - 1) / kTileConfigs[config].tile > 65535u)
      return fail(RNJ_E_INVALID, "too many taxa for one distance launch");
  RNJ_HIP_STATUS(grow_partial(ws, P.partial_elems));
  RNJ_HIP_STATUS(grow_sat(ws, P.sat_units));
  number_jobs(specs, config, P, ws);
  hipError_t e = hipSuccess;
  if (upload_table)
    e = hipMemcpyAsync(ws->d_jobs, ws->h_jobs, specs.size() * sizeof(JobDesc),
                       hipMemcpyHostToDevice, stream);
  if (e == hipSuccess) e = launch_reset_stats(ws->d_jobs, (unsigned)specs.size(), stream);
  if (e == hipSuccess)
    e = launch_counting(config, sh, exact, ws->d_jobs, (unsigned)specs.size(),
                        (unsigned)P.items, (unsigned)P.finish_items, stream);
  if (e == hipSuccess)
    e = launch_saturation(ws->d_jobs, (unsigned)specs.size(), P.max_units, config, !raw, stream);
  for (size_t k = 0; e == hipSuccess && k < specs.size(); ++k)
    if (specs[k].max_out)
      e = hipMemcpyAsync(specs[k].max_out, ws->d_stats + k, sizeof(float),
                         hipMemcpyDeviceToDevice, stream);
  if (e != hipSuccess) return hip_fail(e, "launch");
  return RNJ_OK;
}

int launch_jobs(Context* ctx, std::vector<JobSpec>& specs, bool raw, bool exact,
                hipStream_t stream, bool allow_tune) {
  if (specs.empty()) return RNJ_OK;
  LaunchShape sh = shape_of(specs);
  Workspace* ws = nullptr;
  int rc = acquire_workspace(ctx, (unsigned)specs.size(), &ws);
  if (rc != RNJ_OK) return rc;
  int config = 0;
  rc = select_config(ctx, specs, sh, ws, exact, stream, allow_tune, &config);
  if (rc == RNJ_OK) rc = enqueue_launch(ctx, specs, sh, config, ws, raw, exact, stream);
  release_workspace(ctx, ws, stream);
  return rc;
}

// dataloaderFasta.hpp:54-76 on the first sequence: bytes below 'A' (signed chars, so also bytes
// >= 128) and ACGTUN of either case classify as DNA, anything else as protein.
bool first_row_is_dna(const char* row, unsigned L) {
  for (unsigned k = 0; k < L; ++k) {
    const signed char c = (signed char)row[k];
    if (c < 65) continue;
    switch (c) {
      case 'A': case 'a': case 'C': case 'c': case 'G': case 'g': case 'T': case 't':
      case 'U': case 'u': case 'N': case 'n':
        continue;
      default:
        return false;
    }
  }
  return true;
}

int alignment_kind(const Context* ctx, const char* row0, unsigned L) {
  int kind;
  if (ctx->input_type == RNJ_TYPE_DNA) kind = kDNA;
  else if (ctx->input_type == RNJ_TYPE_PROTEIN) kind = kProtein;
  else kind = first_row_is_dna(row0, L) ? kDNA : kProtein;
  return (kind == kDNA && ctx->char_rules == RNJ_CHARS_NAIVE) ? kNaive : kind;
}

// The characters dataloader::resolveChar maps to the gap for protein alignments (the ones
// bitDistanceProtein never counts as residues).
static bool protein_gap_byte(unsigned v) {
  switch (v) {
    case '-': case '.': case 'X': case 'x': case 'z': case 'Z': case 'b': case 'B': case 'J':
    case 'j': case '?':
      return true;
    default:
      return false;
  }
}

int pack_alignment(Context* ctx, const uint8_t* chars, size_t pitch, unsigned n, unsigned L,
                   int kind, hipStream_t s, Alignment** out, int gap_rule) {
  Alignment* a = nullptr;
  if (kind == kDNA || kind == kNaive) {
    RNJ_HIP_STATUS(alloc_alignment(ctx, n, L, L, &a, kind, 0));
    hipError_t e =
        launch_pack_chars(chars, pitch, n, L, a->n_pad, a->W, a->planes, s, kind == kNaive);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
      rnj_alignment_free(static_cast<rnj_alignment*>(a));
      return hip_fail(e, "pack DNA alignment");
    }
    *out = a;
    return RNJ_OK;
  }
  // protein: census of the byte values, residue codes 1..S in byte order, gaps 0
  uint32_t* scratch = nullptr;  // present[8], then codes[256]
  hipError_t e = hipMalloc((void**)&scratch, 8 * sizeof(uint32_t) + 256);
  if (e == hipSuccess) e = hipMemsetAsync(scratch, 0, 8 * sizeof(uint32_t), s);
  if (e == hipSuccess) e = launch_symbol_census(chars, pitch, n, L, scratch, s);
  uint32_t present[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  if (e == hipSuccess)
    e = hipMemcpyAsync(present, scratch, sizeof present, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  uint8_t codes[256];
  unsigned S = 0;
  for (unsigned v = 0; v < 256; ++v) {
    const bool here = (present[v >> 5] >> (v & 31)) & 1u;
    const bool gap = gap_rule == kGapBelow64 ? v < 64 : protein_gap_byte(v);
    codes[v] = (here && !gap) ? (uint8_t)++S : (uint8_t)0;
  }
  const unsigned B = S < 32 ? 5 : (S < 64 ? 6 : 8);
  int rc = RNJ_OK;
  if (e == hipSuccess) rc = alloc_alignment(ctx, n, L, L, &a, kProtein, B);
  if (e == hipSuccess && rc == RNJ_OK)
    a->hamming = gap_rule == kGapBelow64 || ctx->char_rules == RNJ_CHARS_NAIVE;
  uint8_t* dcodes = reinterpret_cast<uint8_t*>(scratch + 8);
  if (e == hipSuccess && rc == RNJ_OK)
    e = hipMemcpyAsync(dcodes, codes, sizeof codes, hipMemcpyHostToDevice, s);
  if (e == hipSuccess && rc == RNJ_OK)
    e = launch_pack_protein(chars, pitch, n, L, a->n_pad, a->W, dcodes, B, a->planes, a->gaps, s);
  if (e == hipSuccess && rc == RNJ_OK) e = hipStreamSynchronize(s);
  if (scratch) (void)hipFree(scratch);
  if (rc == RNJ_OK && e != hipSuccess) rc = hip_fail(e, "pack protein alignment");
  if (rc != RNJ_OK) {
    if (a) rnj_alignment_free(static_cast<rnj_alignment*>(a));
    return rc;
  }
  *out = a;
  return RNJ_OK;
}

}  // namespace rnj

using namespace rnj;

// ---- helpers for the C ABI ---------------------------------------------------------------------
namespace {

// The caller's stream; NULL is HIP's null stream, which orders with every blocking stream (the
// convention of HIP libraries, and what torch's default stream handle is).
hipStream_t stream_of(rnj_context* ctx, void* stream) {
  (void)ctx;
  return (hipStream_t)stream;
}

// Every entry point starts here: it also clears the thread's last-error slot, so that an error
// left by an unrelated earlier HIP call is not reported by this call's launch checks.
int set_device(rnj_context* ctx) {
  (void)hipGetLastError();
  RNJ_HIP(hipSetDevice(ctx->device));
  return RNJ_OK;
}


int check_model(int model) {
  if (model != RNJ_JC && model != RNJ_K2P) return fail(RNJ_E_INVALID, "model must be RNJ_JC or RNJ_K2P");
  return RNJ_OK;
}

int check_aln(rnj_context* ctx, const rnj_alignment* aln) {
  if (!ctx) return fail(RNJ_E_INVALID, "null context");
  if (!aln || !aln->planes) return fail(RNJ_E_INVALID, "null alignment");
  if (aln->ctx != ctx) return fail(RNJ_E_INVALID, "alignment belongs to another context");
  return RNJ_OK;
}

}  // namespace

JobSpec rnj::alignment_spec(const Alignment* a, int model, int layout, float* out, size_t ld) {
  JobSpec s{};
  s.kind = a->kind;
  s.code_bits = a->code_bits;
  s.gaps = a->gaps;
  s.L = a->L;
  // bitDistanceProtein walks 128 * (L/128 + 1) positions and subtracts from L rounded up to 128,
  // so its valid count is 128 short when L is a multiple of 128 (bitDistanceProtein.cpp:31-61);
  // the Hamming counter of the -f path counts the residue pairs exactly
  s.bias_rule = (a->kind == kProtein && !a->hamming) ? 1u : 0u;
  s.planes = a->planes;
  s.n = a->n;
  s.n_pad = a->n_pad;
  s.W_pad = a->W;
  s.W = a->W;
  s.W_live = (a->L + 31) / 32;
  s.model = model;
  s.layout = layout;
  s.ld = ld;
  s.out = out;
  return s;
}

namespace {
JobSpec spec_for(const rnj_alignment* a, int model, int layout, float* out, size_t ld) {
  return alignment_spec(a, model, layout, out, ld);
}
}  // namespace

size_t rnj::alignment_codes_bytes(const Alignment* a) {
  return a->kind == kProtein ? protein_codes_bytes(a->n_pad, a->L) : codes_bytes(a->n_pad, a->L);
}

size_t rnj::replicate_words(const Alignment* a, unsigned Wr) {
  return ((size_t)a->nplanes() * Wr + (a->kind == kProtein ? 1 : 0)) * a->n_pad;
}

hipError_t rnj::launch_alignment_codes(const Alignment* a, void* codes, hipStream_t s) {
  if (a->kind == kProtein)
    return launch_protein_codes(a->planes, a->n_pad, a->W, a->L, a->code_bits, (uint8_t*)codes, s);
  return launch_planes_to_codes(a->planes, a->n_pad, a->W, a->L, (uint32_t*)codes, s,
                                a->kind == kNaive);
}

hipError_t rnj::launch_replicate_gather(const Alignment* a, const void* codes,
                                        const unsigned* col_index, unsigned idx_stride,
                                        const unsigned* lengths, unsigned R, unsigned Wr,
                                        uint32_t* planes, size_t rep_words, hipStream_t s) {
  if (a->kind == kProtein)
    return launch_protein_gather((const uint8_t*)codes, a->n_pad, a->code_bits, col_index,
                                 idx_stride, lengths, R, Wr, planes, rep_words, s);
  // the DNA gather lays replicates out back to back, rep_words apart
  return launch_bootstrap_gather((const uint32_t*)codes, a->n_pad, col_index, idx_stride, lengths,
                                 R, Wr, planes, s, a->kind == kNaive);
}

JobSpec rnj::replicate_spec(const Alignment* a, const uint32_t* planes, unsigned Wr, unsigned len,
                            const unsigned* len_dev, int model, int layout, float* out, size_t ld) {
  JobSpec j{};
  j.kind = a->kind;
  j.code_bits = a->code_bits;
  j.planes = planes;
  j.n = a->n;
  j.n_pad = a->n_pad;
  j.W_pad = Wr;
  j.W = std::min(words_for_sites(len), Wr);
  j.W_live = (len + 31) / 32;
  j.model = model;
  j.layout = layout;
  j.ld = ld;
  j.out = out;
  if (a->kind == kProtein) {
    j.gaps = planes + (size_t)a->nplanes() * Wr * a->n_pad;
    j.L = len;
    j.len_dev = len_dev;
    j.bias_rule = a->hamming ? 0u : 1u;
  }
  return j;
}

extern "C" {

const char* rnj_last_error(void) { return last_error_cstr(); }

const char* rnj_status_string(int status) {
  switch (status) {
    case RNJ_OK: return "RNJ_OK";
    case RNJ_E_INVALID: return "RNJ_E_INVALID";
    case RNJ_E_HIP: return "RNJ_E_HIP";
    case RNJ_E_NOMEM: return "RNJ_E_NOMEM";
    case RNJ_E_NOT_DNA: return "RNJ_E_NOT_DNA";
    case RNJ_E_ADMISSION: return "RNJ_E_ADMISSION";
    case RNJ_E_CANCELLED: return "RNJ_E_CANCELLED";
    case RNJ_E_NOT_READY: return "RNJ_E_NOT_READY";
    case RNJ_E_NO_DEVICE: return "RNJ_E_NO_DEVICE";
    default: return "unknown status";
  }
}

int rnj_api_version(void) { return RNJ_API_VERSION; }

int rnj_create(int device, rnj_context** out) {
  if (!out) return fail(RNJ_E_INVALID, "null output pointer");
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
    return fail(RNJ_E_NO_DEVICE, "no HIP device");
  if (device < 0 || device >= count) return fail(RNJ_E_INVALID, "device index out of range");
  auto* ctx = new rnj_context();
  ctx->device = device;
  hipError_t e = hipSetDevice(device);
  if (e == hipSuccess) e = hipGetDeviceProperties(&ctx->prop, device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
  if (e == hipSuccess) {
    // keep freed stream-ordered allocations cached in the pool instead of returning them
    hipMemPool_t pool;
    if (hipDeviceGetDefaultMemPool(&pool, device) == hipSuccess) {
      uint64_t thr = UINT64_MAX;
      (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &thr);
    }
  }
  if (e != hipSuccess) {
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return hip_fail(e, "rnj_create");
  }
  // the device address of rnj_distances_device_noalloc's scratch array, once per device
  const int rc = scratch_setup(device);
  if (rc != RNJ_OK) {
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return rc;
  }
  *out = ctx;
  return RNJ_OK;
}

int rnj_synchronize(rnj_context* ctx) {
  if (!ctx) return fail(RNJ_E_INVALID, "null context");
  RNJ_HIP(hipSetDevice(ctx->device));
  RNJ_HIP(hipDeviceSynchronize());
  return RNJ_OK;
}

int rnj_destroy(rnj_context* ctx) {
  if (!ctx) return RNJ_OK;
  (void)hipSetDevice(ctx->device);
  (void)hipDeviceSynchronize();
  for (Alignment* a : ctx->alignments) {
    if (a->planes) (void)hipFree(a->planes);
    if (a->chars_dev) (void)hipFree(a->chars_dev);
    delete a;
  }
  ctx->alignments.clear();
  for (auto& kv : ctx->scratch_free) (void)hipFree(kv.second);
  ctx->scratch_free.clear();
  for (int k = 0; k < 2; ++k) {
    if (ctx->pin[k]) (void)hipHostFree(ctx->pin[k]);
    if (ctx->pin_ev[k]) (void)hipEventDestroy(ctx->pin_ev[k]);
  }
  for (auto& w : ctx->pool) {
    free_workspace_public(w.get());
    if (w->done) (void)hipEventDestroy(w->done);
  }
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return RNJ_OK;
}

int rnj_tile_config_count(void) { return kNumTileConfigs; }

int rnj_get_tile_config_for(rnj_context* ctx, const rnj_alignment* aln, int model, int* config) {
  int rc = check_aln(ctx, aln);
  if (rc == RNJ_OK) rc = check_model(model);
  if (rc != RNJ_OK) return rc;
  if (!config) return fail(RNJ_E_INVALID, "null argument");
  std::vector<JobSpec> jobs{spec_for(aln, model, kTriangle, nullptr, 0)};
  *config = current_config(ctx, shape_of(jobs));
  return RNJ_OK;
}

int rnj_valu_issue_rates(rnj_context* ctx, double* bitwise, double* bcnt) {
  if (!ctx || !bitwise || !bcnt) return fail(RNJ_E_INVALID, "null argument");
  const int rc = set_device(ctx);
  if (rc != RNJ_OK) return rc;
  const unsigned blocks = (unsigned)std::max(1, ctx->prop.multiProcessorCount) * 8;  // 8 waves/SIMD
  const int iters = 2000;
  hipStream_t s = ctx->stream;
  uint32_t* out = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipMalloc((void**)&out, (size_t)blocks * 256 * sizeof(uint32_t));
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  double rate[2] = {0, 0};
  for (int op = 0; op < 2 && e == hipSuccess; ++op) {
    e = launch_issue_probe(op, out, blocks, iters / 4, s);  // warm
    if (e == hipSuccess) e = hipEventRecord(e0, s);
    if (e == hipSuccess) e = launch_issue_probe(op, out, blocks, iters, s);
    if (e == hipSuccess) e = hipEventRecord(e1, s);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e == hipSuccess && ms > 0.0f)
      rate[op] = (double)blocks * 256 * iters * 32 / ((double)ms * 1e-3);
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (out) (void)hipFree(out);
  RNJ_HIP(e);
  *bitwise = rate[0];
  *bcnt = rate[1];
  return RNJ_OK;
}

static_assert(RNJ_MIX_K2P == kMixK2P && RNJ_MIX_JC == kMixJC && RNJ_MIX_NAIVE == kMixNaive &&
                  RNJ_MIX_PROTEIN == kMixProtein,
              "the ABI's mix values are the probe's");

int rnj_valu_mix_rate(rnj_context* ctx, int mix, int planes, double* rate) {
  if (!ctx || !rate) return fail(RNJ_E_INVALID, "null argument");
  if (mix < RNJ_MIX_K2P || mix > RNJ_MIX_PROTEIN)
    return fail(RNJ_E_INVALID, "mix must be RNJ_MIX_K2P, _JC, _NAIVE or _PROTEIN");
  if (mix == RNJ_MIX_PROTEIN && planes != 5 && planes != 6 && planes != 8)
    return fail(RNJ_E_INVALID, "protein code planes must be 5, 6 or 8");
  const int rc = set_device(ctx);
  if (rc != RNJ_OK) return rc;
  // two rounds of 8 waves per SIMD (work-groups of 4 waves, 8 per CU at a time)
  const unsigned blocks = (unsigned)std::max(1, ctx->prop.multiProcessorCount) * 16;
  const int iters = 1000;
  hipStream_t s = ctx->stream;
  uint32_t* out = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipMalloc((void**)&out, (size_t)blocks * 256 * sizeof(uint32_t));
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e == hipSuccess) e = launch_mix_probe(mix, planes, out, blocks, iters / 4, s);  // warm
  float best = 0.0f;
  for (int rep = 0; rep < 3 && e == hipSuccess; ++rep) {
    e = hipEventRecord(e0, s);
    if (e == hipSuccess) e = launch_mix_probe(mix, planes, out, blocks, iters, s);
    if (e == hipSuccess) e = hipEventRecord(e1, s);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e == hipSuccess && ms > 0.0f && (best == 0.0f || ms < best)) best = ms;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (out) (void)hipFree(out);
  RNJ_HIP(e);
  // lane-ops: threads x iterations x 16 pairs x instructions per pair
  *rate = best > 0.0f ? (double)blocks * 256 * iters * 16 * mix_ops_per_pair(mix, planes) /
                            ((double)best * 1e-3)
                      : 0.0;
  return RNJ_OK;
}

int rnj_set_char_rules(rnj_context* ctx, int rules) {
  if (!ctx) return fail(RNJ_E_INVALID, "null context");
  if (rules != RNJ_CHARS_BITS && rules != RNJ_CHARS_NAIVE)
    return fail(RNJ_E_INVALID, "rules must be RNJ_CHARS_BITS or RNJ_CHARS_NAIVE");
  ctx->char_rules = rules;
  return RNJ_OK;
}

int rnj_set_input_type(rnj_context* ctx, int type) {
  if (!ctx) return fail(RNJ_E_INVALID, "null context");
  if (type != RNJ_TYPE_AUTO && type != RNJ_TYPE_DNA && type != RNJ_TYPE_PROTEIN)
    return fail(RNJ_E_INVALID, "type must be RNJ_TYPE_AUTO, RNJ_TYPE_DNA or RNJ_TYPE_PROTEIN");
  ctx->input_type = type;
  return RNJ_OK;
}

const char* rnj_tile_config_name(int config) {
  if (config < 0 || config >= kNumTileConfigs) return nullptr;
  return kTileConfigs[config].name;
}

int rnj_set_tile_config(rnj_context* ctx, int config) {
  if (!ctx) return fail(RNJ_E_INVALID, "null context");
  if (config < -1 || config >= kNumTileConfigs) return fail(RNJ_E_INVALID, "no such tile configuration");
  ctx->tile_config = config;
  return RNJ_OK;
}

int rnj_get_tile_config(rnj_context* ctx, unsigned n, unsigned L, int model, int* config) {
  if (!ctx || !config) return fail(RNJ_E_INVALID, "null argument");
  *config = choose_config(ctx, n, words_for_sites(L), model == RNJ_JC);
  return RNJ_OK;
}

// ---- alignments -----------------------------------------------------------------------------
int rnj_alignment_from_chars(rnj_context* ctx, const char* chars, size_t pitch, unsigned n,
                             unsigned L, int chars_on_device, int skip_type_check,
                             rnj_alignment** out) {
  if (!ctx || !chars || !out) return fail(RNJ_E_INVALID, "null argument");
  if (n < 2) return fail(RNJ_E_INVALID, "need at least 2 sequences");
  if (L < 1) return fail(RNJ_E_INVALID, "empty sequences");
  if (pitch < L) return fail(RNJ_E_INVALID, "pitch smaller than sequence length");
  int rc = set_device(ctx);
  if (rc != RNJ_OK) return rc;
  int kind = kDNA;
  if (!skip_type_check) {
    std::vector<char> first(L);
    if (chars_on_device) RNJ_HIP(hipMemcpy(first.data(), chars, L, hipMemcpyDeviceToHost));
    else std::memcpy(first.data(), chars, L);
    kind = alignment_kind(ctx, first.data(), L);
  }
  if (kind == kDNA && ctx->char_rules == RNJ_CHARS_NAIVE) kind = kNaive;
  const uint8_t* dchars = (const uint8_t*)chars;
  size_t dpitch = pitch;
  void* staging = nullptr;
  if (!chars_on_device) {  // rows land back to back through the pinned staging chunks
    hipError_t e = hipMalloc(&staging, (size_t)n * L);
    if (e == hipSuccess)
      e = ctx->staging.upload((uint8_t*)staging, chars, pitch, L, n, ctx->stream);
    if (e != hipSuccess) {
      if (staging) {
        (void)hipStreamSynchronize(ctx->stream);
        (void)hipFree(staging);
      }
      return hip_fail(e, "rnj_alignment_from_chars upload");
    }
    dchars = (const uint8_t*)staging;
    dpitch = L;
  }
  Alignment* a = nullptr;
  rc = pack_alignment(ctx, dchars, dpitch, n, L, kind, ctx->stream, &a);
  if (staging) {
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(staging);
  }
  if (rc != RNJ_OK) return rc;
  *out = static_cast<rnj_alignment*>(a);
  return RNJ_OK;
}

int rnj_alignment_from_rows(rnj_context* ctx, const char* const* rows, unsigned n, unsigned L,
                            rnj_alignment** out) {
  if (!ctx || !rows || !out) return fail(RNJ_E_INVALID, "null argument");
  if (n < 2) return fail(RNJ_E_INVALID, "need at least 2 sequences");
  std::vector<char> buf((size_t)n * L);
  for (unsigned t = 0; t < n; ++t) {
    if (!rows[t]) return fail(RNJ_E_INVALID, "null sequence row");
    std::memcpy(buf.data() + (size_t)t * L, rows[t], L);
  }
  return rnj_alignment_from_chars(ctx, buf.data(), L, n, L, 0, 0, out);
}

int rnj_alignment_update_chars(rnj_context* ctx, rnj_alignment* aln, const char* chars,
                               size_t pitch, int chars_on_device, void* stream) {
  int rc = check_aln(ctx, aln);
  if (rc != RNJ_OK) return rc;
  if (!chars) return fail(RNJ_E_INVALID, "null characters");
  if (aln->kind != kDNA)
    return fail(RNJ_E_NOT_DNA, "only DNA alignments are updated in place (a protein alignment's "
                               "residue numbering comes from its own characters)");
  if (pitch < aln->L) return fail(RNJ_E_INVALID, "pitch smaller than sequence length");
  rc = set_device(ctx);
  if (rc != RNJ_OK) return rc;
  hipStream_t s = stream_of(ctx, stream);
  const uint8_t* src = reinterpret_cast<const uint8_t*>(chars);
  size_t src_pitch = pitch;
  if (!chars_on_device) {  // through the pinned staging into the alignment's character buffer
    if (!aln->chars_dev)
      RNJ_HIP(hipMalloc((void**)&aln->chars_dev, (size_t)aln->n * aln->L));
    RNJ_HIP(ctx->staging.upload(aln->chars_dev, chars, pitch, aln->L, aln->n, s));
    src = aln->chars_dev;
    src_pitch = aln->L;
  }
  RNJ_HIP(launch_pack_chars(src, src_pitch, aln->n, aln->L, aln->n_pad, aln->W, aln->planes, s));
  return RNJ_OK;
}

int rnj_alignment_from_bitstrings(rnj_context* ctx, const unsigned int* const* bitStrings,
                                  const unsigned int* const* gapFilters, unsigned n,
                                  unsigned bitStringsCount, rnj_alignment** out) {
  if (!ctx || !bitStrings || !gapFilters || !out) return fail(RNJ_E_INVALID, "null argument");
  if (n < 2) return fail(RNJ_E_INVALID, "need at least 2 sequences");
  const unsigned blocks = bitStringsCount / 6;     // bitDistanceGap.cpp:19
  const unsigned words16 = blocks * 6 * 4;         // 16-site u32 words counted per taxon
  if (words16 == 0) return fail(RNJ_E_INVALID, "bitStringsCount below 6 counts no site");
  int rc = set_device(ctx);
  if (rc != RNJ_OK) return rc;
  const size_t stride = words16;
  std::vector<uint32_t> hb((size_t)n * stride), hg((size_t)n * stride);
  for (unsigned t = 0; t < n; ++t) {
    if (!bitStrings[t] || !gapFilters[t]) return fail(RNJ_E_INVALID, "null bit string row");
    std::memcpy(&hb[(size_t)t * stride], bitStrings[t], stride * 4);
    std::memcpy(&hg[(size_t)t * stride], gapFilters[t], stride * 4);
  }
  Alignment* a = nullptr;
  rc = alloc_alignment(ctx, n, words16 * 16, words16 * 16, &a);
  if (rc != RNJ_OK) return rc;
  uint32_t *db = nullptr, *dg = nullptr;
  hipError_t e = hipMalloc((void**)&db, hb.size() * 4);
  if (e == hipSuccess) e = hipMalloc((void**)&dg, hg.size() * 4);
  if (e == hipSuccess) e = hipMemcpy(db, hb.data(), hb.size() * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dg, hg.data(), hg.size() * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = launch_pack_legacy(db, dg, stride, words16, n, a->n_pad, a->W, a->planes, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (db) (void)hipFree(db);
  if (dg) (void)hipFree(dg);
  if (e != hipSuccess) {
    rnj_alignment_free(static_cast<rnj_alignment*>(a));
    return hip_fail(e, "rnj_alignment_from_bitstrings");
  }
  *out = static_cast<rnj_alignment*>(a);
  return RNJ_OK;
}

int rnj_alignment_free(rnj_alignment* aln) {
  if (!aln) return RNJ_OK;
  if (aln->ctx) {
    (void)hipSetDevice(aln->ctx->device);
    std::lock_guard<std::mutex> lock(aln->ctx->mu);
    aln->ctx->alignments.erase(aln);
  }
  if (aln->planes || aln->chars_dev) {
    (void)hipDeviceSynchronize();
    if (aln->planes) (void)hipFree(aln->planes);
    if (aln->chars_dev) (void)hipFree(aln->chars_dev);
  }
  delete aln;
  return RNJ_OK;
}

using TextParser = int (*)(Context*, const uint8_t*, size_t, hipMemcpyKind, hipStream_t,
                           Alignment**);

// A text alignment in host or device memory through `parse` (fasta_to_alignment or
// stockholm_to_alignment).
static int alignment_from_text(rnj_context* ctx, const char* data, size_t len, int data_on_device,
                               rnj_alignment** out, TextParser parse) {
  if (!ctx || (!data && len) || !out) return fail(RNJ_E_INVALID, "null argument");
  int rc = set_device(ctx);
  if (rc != RNJ_OK) return rc;
  // device input may still be in flight on another stream of the caller's
  if (data_on_device) RNJ_HIP(hipDeviceSynchronize());
  // the parser reads 16-byte words: host input is uploaded, unaligned device input is copied
  const bool aligned = (reinterpret_cast<uintptr_t>(data) & 15u) == 0;
  const hipMemcpyKind stage = !data_on_device ? hipMemcpyHostToDevice
                              : aligned       ? hipMemcpyDefault
                                              : hipMemcpyDeviceToDevice;
  Alignment* a = nullptr;
  rc = parse(ctx, reinterpret_cast<const uint8_t*>(data), len, stage, ctx->stream, &a);
  if (rc != RNJ_OK) return rc;
  *out = static_cast<rnj_alignment*>(a);
  return RNJ_OK;
}

int rnj_alignment_from_fasta(rnj_context* ctx, const char* data, size_t len, int data_on_device,
                             rnj_alignment** out) {
  return alignment_from_text(ctx, data, len, data_on_device, out, fasta_to_alignment);
}

int rnj_alignment_from_stockholm(rnj_context* ctx, const char* data, size_t len,
                                 int data_on_device, rnj_alignment** out) {
  return alignment_from_text(ctx, data, len, data_on_device, out, stockholm_to_alignment);
}

// A text alignment file: read through the context's two pinned chunks and uploaded while
// reading, then parsed on the device by `parse`.
static int alignment_from_text_file(rnj_context* ctx, const char* path, rnj_alignment** out,
                                    TextParser parse) {
  if (!ctx || !path || !out) return fail(RNJ_E_INVALID, "null argument");
  int rc = set_device(ctx);
  if (rc != RNJ_OK) return rc;
  const int fd = ::open(path, O_RDONLY);
  if (fd < 0) return fail(RNJ_E_INVALID, std::string("cannot read file: ") + path);
  struct stat st;
  if (::fstat(fd, &st) != 0) {
    ::close(fd);
    return fail(RNJ_E_INVALID, std::string("cannot size file: ") + path);
  }
  const size_t len = (size_t)st.st_size;
  // Two pinned chunks kept by the context: several threads read one chunk (page-cache copies
  // scale with threads) while the other uploads.
  std::lock_guard<std::mutex> file_lock(ctx->pin_mu);
  hipError_t e = ensure_pinned(ctx);
  uint8_t* dev = nullptr;
  if (e == hipSuccess) e = hipMalloc((void**)&dev, len ? len : 1);
  const unsigned hw = std::thread::hardware_concurrency();
  const unsigned nthreads = std::max(1u, std::min(8u, hw / 2));
  size_t done = 0;
  int k = 0;
  bool short_read = false;
  while (e == hipSuccess && done < len) {
    e = hipEventSynchronize(ctx->pin_ev[k]);  // the upload that last used this chunk is done
    if (e != hipSuccess) break;
    const size_t want = std::min(kPinBytes, len - done);
    const size_t part = (want + nthreads - 1) / nthreads;
    std::vector<std::thread> readers;
    std::atomic<bool> bad{false};
    for (unsigned t = 0; t < nthreads && t * part < want; ++t) {
      readers.emplace_back([&, t] {
        size_t at = t * part;
        const size_t stop = std::min(want, at + part);
        while (at < stop) {
          const ssize_t got = ::pread(fd, ctx->pin[k] + at, stop - at, (off_t)(done + at));
          if (got <= 0) {
            bad = true;
            return;
          }
          at += (size_t)got;
        }
      });
    }
    for (auto& r : readers) r.join();
    if (bad) {
      short_read = true;
      break;
    }
    e = hipMemcpyAsync(dev + done, ctx->pin[k], want, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipEventRecord(ctx->pin_ev[k], ctx->stream);
    done += want;
    k ^= 1;
  }
  ::close(fd);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess || short_read) {
    if (dev) (void)hipFree(dev);
    return short_read ? fail(RNJ_E_INVALID, std::string("short read: ") + path)
                      : hip_fail(e, "alignment file upload");
  }
  Alignment* a = nullptr;
  rc = parse(ctx, dev, len, hipMemcpyDefault, ctx->stream, &a);
  (void)hipStreamSynchronize(ctx->stream);
  (void)hipFree(dev);
  if (rc != RNJ_OK) return rc;
  *out = static_cast<rnj_alignment*>(a);
  return RNJ_OK;
}

int rnj_alignment_from_fasta_file(rnj_context* ctx, const char* path, rnj_alignment** out) {
  return alignment_from_text_file(ctx, path, out, fasta_to_alignment);
}

int rnj_alignment_from_stockholm_file(rnj_context* ctx, const char* path, rnj_alignment** out) {
  return alignment_from_text_file(ctx, path, out, stockholm_to_alignment);
}

unsigned rnj_alignment_name_count(const rnj_alignment* aln) {
  return aln ? (unsigned)aln->names.size() : 0u;
}

const char* rnj_alignment_name(const rnj_alignment* aln, unsigned k) {
  if (!aln || k >= aln->names.size()) return nullptr;
  return aln->names[k].c_str();
}

int rnj_alignment_info(const rnj_alignment* aln, unsigned* n, unsigned* L, unsigned* W,
                       size_t* bytes) {
  if (!aln) return fail(RNJ_E_INVALID, "null alignment");
  if (n) *n = aln->n;
  if (L) *L = aln->L;
  if (W) *W = aln->W;
  if (bytes) *bytes = aln->bytes();
  return RNJ_OK;
}

int rnj_alignment_type(const rnj_alignment* aln, int* type, unsigned* planes) {
  if (!aln) return fail(RNJ_E_INVALID, "null alignment");
  if (type) *type = aln->kind == kProtein ? RNJ_TYPE_PROTEIN : RNJ_TYPE_DNA;
  if (planes) *planes = aln->nplanes();
  return RNJ_OK;
}

int rnj_alignment_planes(const rnj_alignment* aln, uint32_t* host_planes) {
  if (!aln || !host_planes) return fail(RNJ_E_INVALID, "null argument");
  RNJ_HIP(hipSetDevice(aln->ctx->device));
  RNJ_HIP(hipStreamSynchronize(aln->ctx->stream));
  for (unsigned p = 0; p < aln->nplanes(); ++p)
    RNJ_HIP(hipMemcpy2D(host_planes + (size_t)p * aln->W * aln->n, aln->n * sizeof(uint32_t),
                        aln->planes + (size_t)p * aln->W * aln->n_pad,
                        aln->n_pad * sizeof(uint32_t), aln->n * sizeof(uint32_t), aln->W,
                        hipMemcpyDeviceToHost));
  return RNJ_OK;
}

// ---- one matrix ---------------------------------------------------------------------------------
int rnj_distances_device(rnj_context* ctx, const rnj_alignment* aln, int model, int layout,
                         float* out, size_t ld, int flags, float* max_out, void* stream) {
  int rc = check_aln(ctx, aln);
  if (rc == RNJ_OK) rc = check_model(model);
  if (rc != RNJ_OK) return rc;
  if (!out) return fail(RNJ_E_INVALID, "null output");
  if (layout != RNJ_TRIANGLE && layout != RNJ_SQUARE) return fail(RNJ_E_INVALID, "unknown layout");
  if (layout == RNJ_SQUARE && ld < aln->n) return fail(RNJ_E_INVALID, "ld smaller than n");
  rc = set_device(ctx);
  if (rc != RNJ_OK) return rc;
  std::vector<JobSpec> jobs{spec_for(aln, model, layout, out, ld)};
  jobs[0].max_out = max_out;
  return launch_jobs(ctx, jobs, (flags & RNJ_FLAG_RAW) != 0, (flags & RNJ_FLAG_EXACT) != 0,
                     stream_of(ctx, stream));
}

int rnj_copy_to_host(rnj_context* ctx, void* dst, const void* src, size_t bytes, void* stream) {
  if (!ctx || ((!dst || !src) && bytes)) return fail(RNJ_E_INVALID, "null argument");
  const int rc = set_device(ctx);
  if (rc != RNJ_OK) return rc;
  RNJ_HIP(ctx->staging.download(dst, src, bytes, stream_of(ctx, stream)));
  return RNJ_OK;
}

int rnj_counts_device(rnj_context* ctx, const rnj_alignment* aln, uint32_t* ts, uint32_t* tv,
                      uint32_t* valid, void* stream) {
  int rc = check_aln(ctx, aln);
  if (rc != RNJ_OK) return rc;
  if (!ts || !tv || !valid) return fail(RNJ_E_INVALID, "null output");
  rc = set_device(ctx);
  if (rc != RNJ_OK) return rc;
  std::vector<JobSpec> jobs{spec_for(aln, kCounts, kTriangle, nullptr, 0)};
  jobs[0].out_ts = ts;
  jobs[0].out_tv = tv;
  jobs[0].out_valid = valid;
  return launch_jobs(ctx, jobs, true, false, stream_of(ctx, stream));
}

}  // extern "C"

// ---- host hand-off ------------------------------------------------------------------------------
namespace rnj {
hipError_t ensure_pinned(Context* ctx) {
  hipError_t e = hipSuccess;
  for (int k = 0; k < 2 && e == hipSuccess; ++k) {
    if (!ctx->pin[k]) e = hipHostMalloc((void**)&ctx->pin[k], kPinBytes, hipHostMallocDefault);
    if (e == hipSuccess && !ctx->pin_ev[k])
      e = hipEventCreateWithFlags(&ctx->pin_ev[k], hipEventDisableTiming);
  }
  return e;
}
}  // namespace rnj

// The post-processed square matrix is computed in HBM once. The rows handed over (full rows, or
// the lower triangle with the diagonal that rapidNJMem keeps, distMatrixReader.cpp:26-35) lie
// back to back on the device (the lower triangle packed by launch_lower_pack), and go to the
// host in slabs of whole rows through the context's two pinned buffers: the copy of slab k+1
// overlaps the delivery of slab k. Delivery runs on `threads` library threads, each taking every
// threads-th row of a slab (copies from pinned into the caller's pageable rows are the slow
// part); with one thread the rows arrive in order.
namespace {
struct RowShape {
  bool lower;
  unsigned n;
  size_t off(unsigned r) const { return lower ? (size_t)r * (r + 1) / 2 : (size_t)r * n; }
  size_t len(unsigned r) const { return lower ? (size_t)r + 1 : (size_t)n; }
};
}  // namespace

// row_side: every entry (i, j) counted with taxon i as the row's taxon and the maximum taken over
// the whole square (the disk-matrix path; differs from the symmetric matrix only for naive-rule
// alignments).
template <class Deliver>
static int rows_handoff(rnj_context* ctx, const rnj_alignment* aln, int model, float* max_out,
                        bool lower, bool row_side, unsigned threads, Deliver deliver) {
  int rc = check_aln(ctx, aln);
  if (rc == RNJ_OK) rc = check_model(model);
  if (rc != RNJ_OK) return rc;
  rc = set_device(ctx);
  if (rc != RNJ_OK) return rc;
  const unsigned n = aln->n;
  const RowShape shape{lower, n};
  const size_t cap = kPinBytes / sizeof(float);
  if ((size_t)n > cap) return fail(RNJ_E_INVALID, "rows longer than the pinned staging buffer");
  hipStream_t s = ctx->stream;
  float *dmat = nullptr, *dlow = nullptr, *dmax = nullptr;
  RNJ_HIP(hipMallocAsync((void**)&dmat, (size_t)n * n * sizeof(float), s));
  RNJ_HIP(hipMallocAsync((void**)&dmax, sizeof(float), s));
  std::vector<JobSpec> jobs{
      spec_for(aln, model, row_side && aln->kind == kNaive ? kRows : kSquare, dmat, n)};
  jobs[0].max_out = dmax;
  rc = launch_jobs(ctx, jobs, false, true, s);
  hipError_t e = hipSuccess;
  const float* src = dmat;
  if (rc == RNJ_OK && lower) {
    e = hipMallocAsync((void**)&dlow, shape.off(n) * sizeof(float), s);
    if (e == hipSuccess) e = launch_lower_pack(dmat, n, n, dlow, s);
    src = dlow;
  }
  // slabs of whole rows
  std::vector<std::pair<unsigned, unsigned>> slabs;
  for (unsigned r0 = 0, r = 0; r <= n; ++r) {
    if (r == n || shape.off(r + 1) - shape.off(r0) > cap) {
      if (r > r0 || r == n) slabs.emplace_back(r0, r == n ? n : r);
      if (r == n) break;
      r0 = r;
    }
  }
  std::lock_guard<std::mutex> pin_lock(ctx->pin_mu);
  if (rc == RNJ_OK && e == hipSuccess) e = ensure_pinned(ctx);
  std::mutex m;
  std::condition_variable cv;
  long long issued = -1;
  std::vector<unsigned> consumed(slabs.size(), 0);
  bool stop = rc != RNJ_OK || e != hipSuccess;
  int deliver_rc = RNJ_OK;
  const unsigned T = std::max(1u, threads);
  auto worker = [&](unsigned t) {
    for (size_t k = 0; k < slabs.size(); ++k) {
      {
        std::unique_lock<std::mutex> lk(m);
        cv.wait(lk, [&] { return issued >= (long long)k || stop; });
        if (stop) return;
      }
      int wrc = RNJ_OK;
      if (hipEventSynchronize(ctx->pin_ev[k & 1]) != hipSuccess)
        wrc = fail(RNJ_E_HIP, "row hand-off copy failed");
      const float* base = reinterpret_cast<const float*>(ctx->pin[k & 1]);
      const size_t off0 = shape.off(slabs[k].first);
      for (unsigned r = slabs[k].first + t; wrc == RNJ_OK && r < slabs[k].second; r += T)
        wrc = deliver(r, base + (shape.off(r) - off0), shape.len(r));
      std::lock_guard<std::mutex> lk(m);
      if (wrc != RNJ_OK) {
        if (deliver_rc == RNJ_OK) deliver_rc = wrc;
        stop = true;
      }
      ++consumed[k];
      cv.notify_all();
      if (stop) return;
    }
  };
  std::vector<std::thread> pool;
  if (!stop)
    for (unsigned t = 0; t < T; ++t) pool.emplace_back(worker, t);
  for (size_t k = 0; !stop && k < slabs.size(); ++k) {
    if (k >= 2) {
      std::unique_lock<std::mutex> lk(m);
      cv.wait(lk, [&] { return consumed[k - 2] == T || stop; });
      if (stop) break;
    }
    const size_t first = shape.off(slabs[k].first), cnt = shape.off(slabs[k].second) - first;
    e = hipMemcpyAsync(ctx->pin[k & 1], src + first, cnt * sizeof(float), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipEventRecord(ctx->pin_ev[k & 1], s);
    std::lock_guard<std::mutex> lk(m);
    if (e != hipSuccess) stop = true;
    else issued = (long long)k;
    cv.notify_all();
  }
  for (auto& th : pool) th.join();
  if (rc == RNJ_OK && e == hipSuccess && deliver_rc == RNJ_OK && max_out)
    e = hipMemcpyAsync(max_out, dmax, sizeof(float), hipMemcpyDeviceToHost, s);
  (void)hipStreamSynchronize(s);
  (void)hipFreeAsync(dmat, s);
  if (dlow) (void)hipFreeAsync(dlow, s);
  (void)hipFreeAsync(dmax, s);
  if (rc != RNJ_OK) return rc;
  if (deliver_rc != RNJ_OK) return deliver_rc;
  if (e != hipSuccess) return hip_fail(e, "row hand-off");
  return RNJ_OK;
}

static unsigned copy_threads() {
  return std::max(1u, std::min(8u, std::thread::hardware_concurrency() / 2));
}

extern "C" {

int rnj_distances_rows(rnj_context* ctx, const rnj_alignment* aln, int model, float** rows,
                       float* max_out) {
  if (!aln || !rows) return fail(RNJ_E_INVALID, "null argument");
  for (unsigned i = 0; i < aln->n; ++i)
    if (!rows[i]) return fail(RNJ_E_INVALID, "null row");
  return rows_handoff(ctx, aln, model, max_out, false, false, copy_threads(),
                      [rows](unsigned row, const float* v, size_t len) {
                        std::memcpy(rows[row], v, len * sizeof(float));
                        return (int)RNJ_OK;
                      });
}

int rnj_distances_rows_lower(rnj_context* ctx, const rnj_alignment* aln, int model, float** rows,
                             float* max_out) {
  if (!aln || !rows) return fail(RNJ_E_INVALID, "null argument");
  for (unsigned i = 0; i < aln->n; ++i)
    if (!rows[i]) return fail(RNJ_E_INVALID, "null row");
  return rows_handoff(ctx, aln, model, max_out, true, false, copy_threads(),
                      [rows](unsigned row, const float* v, size_t len) {
                        std::memcpy(rows[row], v, len * sizeof(float));
                        return (int)RNJ_OK;
                      });
}

int rnj_distances_stream_rows(rnj_context* ctx, const rnj_alignment* aln, int model,
                              rnj_row_sink sink, void* user, float* max_out) {
  if (!aln || !sink) return fail(RNJ_E_INVALID, "null argument");
  return rows_handoff(ctx, aln, model, max_out, false, true, 1,
                      [sink, user](unsigned row, const float* v, size_t len) {
                        if (sink(user, row, v, (unsigned)len) != 0)
                          return fail(RNJ_E_CANCELLED, "row sink stopped the stream");
                        return (int)RNJ_OK;
                      });
}

// ---- bootstrap ------------------------------------------------------------------------------------
int rnj_bootstrap_device(rnj_context* ctx, const rnj_alignment* aln, int model,
                         const uint16_t* weights, unsigned R, int layout, float* out,
                         size_t matrix_stride, size_t ld, int flags, void* stream) {
  int rc = check_aln(ctx, aln);
  if (rc == RNJ_OK) rc = check_model(model);
  if (rc != RNJ_OK) return rc;
  if (!weights || !out) return fail(RNJ_E_INVALID, "null argument");
  if (R == 0) return RNJ_OK;
  if (R > kMaxLaunchJobs) return fail(RNJ_E_INVALID, "more than 65535 replicates in one call");
  if (layout != RNJ_TRIANGLE && layout != RNJ_SQUARE) return fail(RNJ_E_INVALID, "unknown layout");
  const unsigned n = aln->n, L = aln->L;
  const size_t min_stride = layout == RNJ_SQUARE ? (size_t)n * ld : (size_t)n * (n - 1) / 2;
  if (layout == RNJ_SQUARE && ld < n) return fail(RNJ_E_INVALID, "ld smaller than n");
  if (matrix_stride < min_stride) return fail(RNJ_E_INVALID, "matrix_stride too small");
  rc = set_device(ctx);
  if (rc != RNJ_OK) return rc;
  hipStream_t s = stream_of(ctx, stream);
  // replicate lengths (sum of weights) size the column lists and the replicate planes
  std::vector<unsigned> len(R);
  unsigned maxlen = 1;
  for (unsigned r = 0; r < R; ++r) {
    unsigned long long sum = 0;
    for (unsigned c = 0; c < L; ++c) sum += weights[(size_t)r * L + c];
    if (sum >= (1ull << 31)) return fail(RNJ_E_INVALID, "replicate weights sum too large");
    len[r] = (unsigned)sum;
    maxlen = std::max(maxlen, len[r]);
  }
  const unsigned Wr = words_for_sites(maxlen);
  const size_t rep_words = replicate_words(aln, Wr);
  uint16_t* dw = nullptr;
  unsigned *didx = nullptr, *dlen = nullptr;
  void* dcodes = nullptr;
  uint32_t* dplanes = nullptr;
  hipError_t e = hipMallocAsync((void**)&dw, (size_t)R * L * sizeof(uint16_t), s);
  if (e == hipSuccess) e = hipMallocAsync((void**)&didx, (size_t)R * maxlen * sizeof(unsigned), s);
  if (e == hipSuccess) e = hipMallocAsync((void**)&dlen, (size_t)R * sizeof(unsigned), s);
  if (e == hipSuccess) e = hipMallocAsync(&dcodes, alignment_codes_bytes(aln), s);
  if (e == hipSuccess) e = hipMallocAsync((void**)&dplanes, (size_t)R * rep_words * 4, s);
  if (e == hipSuccess)
    e = hipMemcpyAsync(dw, weights, (size_t)R * L * sizeof(uint16_t), hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = launch_alignment_codes(aln, dcodes, s);
  if (e == hipSuccess) e = launch_bootstrap_index(dw, R, L, didx, maxlen, dlen, s);
  if (e == hipSuccess)
    e = launch_replicate_gather(aln, dcodes, didx, maxlen, dlen, R, Wr, dplanes, rep_words, s);
  if (e == hipSuccess) {
    std::vector<JobSpec> jobs(R);
    for (unsigned r = 0; r < R; ++r)
      jobs[r] = replicate_spec(aln, dplanes + (size_t)r * rep_words, Wr, len[r], nullptr, model,
                               layout, out + (size_t)r * matrix_stride, ld);
    rc = launch_jobs(ctx, jobs, (flags & RNJ_FLAG_RAW) != 0, (flags & RNJ_FLAG_EXACT) != 0, s);
  } else {
    rc = hip_fail(e, "rnj_bootstrap_device");
  }
  if (dw) (void)hipFreeAsync(dw, s);
  if (didx) (void)hipFreeAsync(didx, s);
  if (dlen) (void)hipFreeAsync(dlen, s);
  if (dcodes) (void)hipFreeAsync(dcodes, s);
  if (dplanes) (void)hipFreeAsync(dplanes, s);
  return rc;
}

}  // extern "C"

// ---- the reference's device boundary ---------------------------------------------------------------
// Module-wide state like the reference's (kimuraDNA_gpu.cu:178-191); errors print and exit(1) as
// cutilSafeCall does (cutil_inline_runtime.h:159-166).
namespace {
struct LegacyState {
  rnj_context* ctx = nullptr;
  unsigned n = 0, bsc = 0, stride = 0, words16 = 0;
  uint32_t* d_bits = nullptr;
  uint32_t* d_gaps = nullptr;
  rnj_alignment* aln = nullptr;
};
LegacyState g_k2p, g_counts;

void legacy_die(const char* where) {
  std::fprintf(stderr, "rapidnj_amd: %s failed: %s\n", where, last_error_cstr());
  std::exit(1);
}

void legacy_check(int rc, const char* where) {
  if (rc != RNJ_OK) legacy_die(where);
}

void legacy_init(LegacyState& st, unsigned n, unsigned bsc, unsigned bsStride, const char* where) {
  if (!st.ctx) legacy_check(rnj_create(0, &st.ctx), where);
  if (n < 2) {
    set_error("need at least 2 sequences");
    legacy_die(where);
  }
  if (st.aln) { rnj_alignment_free(st.aln); st.aln = nullptr; }
  if (st.d_bits) (void)hipFree(st.d_bits);
  if (st.d_gaps) (void)hipFree(st.d_gaps);
  st.d_bits = st.d_gaps = nullptr;
  st.n = n;
  st.bsc = bsc;
  st.stride = bsStride;
  st.words16 = std::min(bsStride, bsc * 4);  // kimuraDNA_gpu.cu:245
  const size_t bytes = (size_t)n * bsStride * sizeof(uint32_t);
  hipError_t e = hipMalloc((void**)&st.d_bits, bytes);
  if (e == hipSuccess) e = hipMalloc((void**)&st.d_gaps, bytes);
  if (e != hipSuccess) legacy_check(hip_fail(e, "hipMalloc"), where);
  Alignment* a = nullptr;
  legacy_check(alloc_alignment(st.ctx, n, st.words16 * 16, st.words16 * 16, &a), where);
  st.aln = static_cast<rnj_alignment*>(a);
}

void legacy_store(LegacyState& st, const unsigned* bits, const unsigned* gaps, const char* where) {
  if (!st.ctx) {
    set_error("called before initialisation");
    legacy_die(where);
  }
  const size_t bytes = (size_t)st.n * st.stride * sizeof(uint32_t);
  hipError_t e = hipMemcpy(st.d_bits, bits, bytes, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(st.d_gaps, gaps, bytes, hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = launch_pack_legacy(st.d_bits, st.d_gaps, st.stride, st.words16, st.n, st.aln->n_pad,
                           st.aln->W, st.aln->planes, st.ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(st.ctx->stream);
  if (e != hipSuccess) legacy_check(hip_fail(e, where), where);
}
}  // namespace

extern "C" {

void initialiseDNA2_gpu(unsigned int sequenceCount, unsigned int bitStringCount,
                        unsigned int bsStride) {
  legacy_init(g_k2p, sequenceCount, bitStringCount, bsStride, "initialiseDNA2_gpu");
}

void storeDataDNA2_gpu(unsigned int* bitStrings, unsigned int* gapFilters) {
  legacy_store(g_k2p, bitStrings, gapFilters, "storeDataDNA2_gpu");
}

void computeDistancesDNA2_gpu(float* results) {
  const char* where = "computeDistancesDNA2_gpu";
  if (!g_k2p.ctx) {
    set_error("called before initialisation");
    legacy_die(where);
  }
  const unsigned n = g_k2p.n;
  float* d = nullptr;
  hipError_t e = hipMalloc((void**)&d, (size_t)n * n * sizeof(float));
  if (e != hipSuccess) legacy_check(hip_fail(e, "hipMalloc"), where);
  legacy_check(rnj_distances_device(g_k2p.ctx, g_k2p.aln, RNJ_K2P, RNJ_SQUARE, d, n,
                                    RNJ_FLAG_RAW | RNJ_FLAG_EXACT,
                                    nullptr, nullptr),
               where);
  e = hipMemcpy(results, d, (size_t)n * n * sizeof(float), hipMemcpyDeviceToHost);
  (void)hipFree(d);
  if (e != hipSuccess) legacy_check(hip_fail(e, where), where);
}

void initialiseDNA_gpu(unsigned int sequenceCount, unsigned int bitStringCount,
                       unsigned int bsStride) {
  legacy_init(g_counts, sequenceCount, bitStringCount, bsStride, "initialiseDNA_gpu");
}

void storeDataDNA_gpu(unsigned int* bitStrings, unsigned int* gapFilters) {
  legacy_store(g_counts, bitStrings, gapFilters, "storeDataDNA_gpu");
}

void computeDistancesDNA_gpu(unsigned int* ts, unsigned int* tv, unsigned int* gaps) {
  const char* where = "computeDistancesDNA_gpu";
  if (!g_counts.ctx) {
    set_error("called before initialisation");
    legacy_die(where);
  }
  const unsigned n = g_counts.n;
  const size_t m = (size_t)n * (n - 1) / 2;
  uint32_t* d = nullptr;
  hipError_t e = hipMalloc((void**)&d, 3 * m * sizeof(uint32_t));
  if (e != hipSuccess) legacy_check(hip_fail(e, "hipMalloc"), where);
  legacy_check(rnj_counts_device(g_counts.ctx, g_counts.aln, d, d + m, d + 2 * m, nullptr), where);
  std::vector<uint32_t> h(3 * m);
  e = hipMemcpy(h.data(), d, 3 * m * sizeof(uint32_t), hipMemcpyDeviceToHost);
  (void)hipFree(d);
  if (e != hipSuccess) legacy_check(hip_fail(e, where), where);
  unsigned int* outs[3] = {ts, tv, gaps};
  for (int q = 0; q < 3; ++q) {
    const uint32_t* src = h.data() + q * m;
    unsigned int* dst = outs[q];
    size_t k = 0;
    for (size_t i = 0; i < n; ++i) {
      dst[i * n + i] = 0;
      for (size_t j = i + 1; j < n; ++j, ++k) {
        dst[i * n + j] = src[k];
        dst[j * n + i] = src[k];
      }
    }
  }
}

void getResultsDNA_gpu(unsigned int* ts, unsigned int* tv, unsigned int* gaps) {
  (void)ts; (void)tv; (void)gaps;
  if (g_counts.ctx) legacy_check(rnj_synchronize(g_counts.ctx), "getResultsDNA_gpu");
}

}  // extern "C"

// ---- the reference's protein device boundary (gpu/distanceCalculatorProtein_gpu.cu:139-229) ----
namespace {
struct ProteinLegacyState {
  rnj_context* ctx = nullptr;
  unsigned n = 0, stride = 0, padded = 0;
};
ProteinLegacyState g_protein;
}  // namespace

extern "C" {

void initialiseProtein_gpu(unsigned int sequenceCount, unsigned int bsStride,
                           unsigned int paddedLength) {
  const char* where = "initialiseProtein_gpu";
  ProteinLegacyState& st = g_protein;
  if (!st.ctx) legacy_check(rnj_create(0, &st.ctx), where);
  if (sequenceCount < 2) {
    set_error("need at least 2 sequences");
    legacy_die(where);
  }
  if (paddedLength == 0 || (size_t)paddedLength > (size_t)bsStride * 4) {
    set_error("paddedLength must be between 1 and 4 * bsStride characters");
    legacy_die(where);
  }
  st.n = sequenceCount;
  st.stride = bsStride;
  st.padded = paddedLength;
}

void computeDistancesProtein_gpu(unsigned int* bitStrings, unsigned int* dist, unsigned int* gaps) {
  const char* where = "computeDistancesProtein_gpu";
  ProteinLegacyState& st = g_protein;
  if (!st.ctx) {
    set_error("called before initialisation");
    legacy_die(where);
  }
  if (!bitStrings || !dist || !gaps) {
    set_error("null argument");
    legacy_die(where);
  }
  Context* ctx = st.ctx;
  const unsigned n = st.n, P = st.padded;
  const size_t m = (size_t)n * (n - 1) / 2;
  legacy_check(set_device(st.ctx), where);
  // the staged rows (4 characters per word, dataloader::initGPUData) as n x P characters
  uint8_t* dchars = nullptr;
  uint32_t* d = nullptr;
  hipError_t e = hipMalloc((void**)&dchars, (size_t)n * P);
  if (e == hipSuccess) e = hipMalloc((void**)&d, 3 * std::max<size_t>(m, 1) * sizeof(uint32_t));
  if (e == hipSuccess)
    e = ctx->staging.upload(dchars, reinterpret_cast<const char*>(bitStrings),
                            (size_t)st.stride * 4, P, n, ctx->stream);
  if (e != hipSuccess) legacy_check(hip_fail(e, where), where);
  Alignment* a = nullptr;
  legacy_check(pack_alignment(ctx, dchars, P, n, P, kProtein, ctx->stream, &a, kGapBelow64), where);
  rnj_alignment* aln = static_cast<rnj_alignment*>(a);
  const int rc = rnj_counts_device(st.ctx, aln, d, d + m, d + 2 * m, ctx->stream);
  std::vector<uint32_t> h(3 * m);
  if (rc == RNJ_OK)
    e = hipMemcpyAsync(h.data(), d, 3 * m * sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream);
  if (rc == RNJ_OK && e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  rnj_alignment_free(aln);
  (void)hipFree(dchars);
  (void)hipFree(d);
  legacy_check(rc, where);
  if (e != hipSuccess) legacy_check(hip_fail(e, where), where);
  // mismatches and the positions where either character is a gap, of the P positions walked
  size_t k = 0;
  for (size_t i = 0; i < n; ++i) {
    dist[i * n + i] = 0;
    gaps[i * n + i] = 0;
    for (size_t j = i + 1; j < n; ++j, ++k) {
      const uint32_t gp = P - h[2 * m + k];
      dist[i * n + j] = dist[j * n + i] = h[k];
      gaps[i * n + j] = gaps[j * n + i] = gp;
    }
  }
}

void getResultsProtein_gpu(unsigned int* dist, unsigned int* gaps) {
  (void)dist; (void)gaps;
  if (g_protein.ctx) legacy_check(rnj_synchronize(g_protein.ctx), "getResultsProtein_gpu");
}

}  // extern "C"
