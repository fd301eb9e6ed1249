// Host runtime shared by runtime.cpp (context, alignments, launches, C ABI) and server.cpp
// (served jobs). Not part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <map>
#include <set>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/rapidnj_amd.h"
#include "internal.hpp"

namespace rnj {

// ---- errors ---------------------------------------------------------------------------------
void set_error(const std::string& msg);
const char* last_error_cstr();
int fail(int status, const std::string& msg);
int hip_fail(hipError_t e, const char* what);

#define RNJ_HIP(call)                                   \
  do {                                                  \
    hipError_t _e = (call);                             \
    if (_e != hipSuccess) return hip_fail(_e, #call);   \
  } while (0)

#define RNJ_HIP_STATUS(call)                            \
  do {                                                  \
    const int _rc = (call);                             \
    if (_rc != RNJ_OK) return _rc;                      \
  } while (0)

// ---- scratch for one launch -----------------------------------------------------------------
// Job table (pinned host copy + device copy), per-job statistics (max bits, saturated count) and
// the saturated-pair list. Reused once the event recorded after its last launch has completed.
struct Workspace {
  JobDesc* h_jobs = nullptr;
  JobDesc* d_jobs = nullptr;
  unsigned cap_jobs = 0;
  unsigned* d_stats = nullptr;  // [cap_jobs]: maxbits
  uint32_t* d_sat = nullptr;    // [cap_units] wave words, [cap_units] tiles, [cap_units][64] masks
  size_t cap_units = 0;
  uint32_t* d_partial = nullptr;  // partial counts of split tiles
  size_t cap_partial = 0;         // u32 elements
  hipEvent_t done = nullptr;
  bool recorded = false;
  bool in_use = false;
};

// ---- copies between pageable host memory and the device (upload.cpp) ------------------------
// Pinned staging chunks: library threads fill (or empty) one chunk while a neighbouring chunk's
// DMA runs. Calls on one Stager are serialised.
class CopyPool;
class Stager {
 public:
  Stager();
  Stager(const Stager&) = delete;
  Stager& operator=(const Stager&) = delete;
  ~Stager();
  // `rows` rows of `len` bytes (row t at src + t * pitch, host memory) into dst (device memory,
  // rows back to back), enqueued on `stream`. Returns once every source byte has been read;
  // work enqueued on `stream` afterwards sees the data.
  hipError_t upload(uint8_t* dst, const char* src, size_t pitch, size_t len, size_t rows,
                    hipStream_t stream);
  // `bytes` from device memory into host memory, through DMAs on `stream` (after the work
  // already enqueued there). Returns when the bytes are in dst_host.
  hipError_t download(void* dst_host, const void* src_dev, size_t bytes, hipStream_t stream);
  static constexpr int kChunks = 4;
  static constexpr size_t kChunkBytes = (size_t)8 << 20;
  static constexpr size_t kParallelBytes = (size_t)1 << 20;  // smaller chunks: one thread copies

 private:
  hipError_t ensure();
  std::mutex mu_;
  uint8_t* chunk_[kChunks] = {};
  hipEvent_t ev_[kChunks] = {};
  int next_ = 0;
  unsigned parts_ = 1;
  std::unique_ptr<CopyPool> pool_;
};

struct Context {
  int device = 0;
  hipStream_t stream = nullptr;
  int tile_config = -1;  // -1: autotune per shape class
  std::mutex mu;
  std::vector<std::unique_ptr<Workspace>> pool;
  // autotune cache: (log2 n bucket, log2 W bucket, jc_only) -> config
  std::map<std::tuple<int, int, int>, int> tuned;
  hipDeviceProp_t prop{};
  std::set<struct Alignment*> alignments;  // live alignments (freed by rnj_destroy if left)
  // device blocks of finished FASTA parses, by power-of-two size, reused by later parses
  std::mutex scratch_mu;
  std::multimap<size_t, void*> scratch_free;
  size_t scratch_bytes = 0;
  // two pinned staging buffers (FASTA file reads, row hand-offs), allocated on first use, kept
  std::mutex pin_mu;
  uint8_t* pin[2] = {nullptr, nullptr};
  hipEvent_t pin_ev[2] = {nullptr, nullptr};
  int input_type = RNJ_TYPE_AUTO;  // type of alignments read from characters / FASTA
  int char_rules = RNJ_CHARS_BITS;  // rules of DNA alignments read from characters
  Stager staging;  // characters and FASTA images from host memory, results to it
};

constexpr size_t kPinBytes = (size_t)32 << 20;  // bytes per pinned staging buffer

// Allocate the context's pinned staging buffers if needed (caller holds pin_mu).
hipError_t ensure_pinned(Context* ctx);

struct Alignment {
  Context* ctx = nullptr;
  unsigned n = 0, L = 0;     // taxa, sites
  unsigned n_pad = 0;        // multiple of kTaxaPad
  unsigned W = 0;            // 32-site words, multiple of kWordPad (= W_pad of the planes)
  int kind = kDNA;           // Kind
  unsigned code_bits = 0;    // protein: B code planes (5, 6 or 8)
  // protein whose valid count is the residue pairs exactly, as
  // simpleDistanceCalculator::calculateHammingDistance (simpleDistanceCalculator.cpp:40-55, the
  // `-f` path) and the dormant protein kernel count them, without bitDistanceProtein's
  // 128-position shortfall when L is a multiple of 128
  bool hamming = false;
  uint32_t* planes = nullptr;  // [nplanes()][W][n_pad], then (protein) gaps
  uint32_t* gaps = nullptr;    // protein: gap sites per taxon [n_pad], inside the planes block
  std::vector<std::string> names;  // FASTA record names in file order (alignments from FASTA)
  uint8_t* chars_dev = nullptr;      // n x L characters of in-place updates (allocated by the first)
  unsigned nplanes() const { return kind == kProtein ? 1 + code_bits : (kind == kNaive ? 4 : 3); }
  size_t bytes() const {
    return ((size_t)nplanes() * W + (kind == kProtein ? 1 : 0)) * n_pad * sizeof(uint32_t);
  }
};

// One matrix of a launch, before the runtime numbers tiles and binds scratch.
struct JobSpec {
  int kind;               // Kind; the jobs of one launch share kind and code_bits
  unsigned code_bits;     // protein
  const uint32_t* gaps;   // protein: gap sites per taxon
  unsigned L;             // protein: sites
  unsigned bias_rule;     // protein: 1 for bitDistanceProtein's valid count (JobDesc::bias_rule)
  const unsigned* len_dev;  // protein: the sites in device memory (bootstrap replicates), or null
  const uint32_t* planes;
  unsigned n, n_pad, W_pad, W;
  unsigned W_live;  // words holding sites (0: W); the DNA tile kernel counts no word beyond
  int model;   // Model
  int layout;  // Layout
  size_t ld;
  float* out;
  uint32_t *out_ts, *out_tv, *out_valid;
  float* max_out;  // device pointer or null
};

// The launch description of one alignment's matrix.
JobSpec alignment_spec(const Alignment* a, int model, int layout, float* out, size_t ld);

// Bootstrap replicates of an alignment (DNA under either character rules, or protein): per-site
// codes of the alignment for the column gathers, the replicates' planes (replicate r at
// planes + r * replicate_words(a, Wr), Wr words each; protein ones followed by their gap counts)
// and the launch description of replicate r, of `len` sites (a protein replicate's length is read
// from len_dev when that is set, the case of captured plans whose lengths change per run).
size_t alignment_codes_bytes(const Alignment* a);
size_t replicate_words(const Alignment* a, unsigned Wr);
hipError_t launch_alignment_codes(const Alignment* a, void* codes, hipStream_t s);
hipError_t launch_replicate_gather(const Alignment* a, const void* codes, const unsigned* col_index,
                                   unsigned idx_stride, const unsigned* lengths, unsigned R,
                                   unsigned Wr, uint32_t* planes, size_t rep_words, hipStream_t s);
JobSpec replicate_spec(const Alignment* a, const uint32_t* planes, unsigned Wr, unsigned len,
                       const unsigned* len_dev, int model, int layout, float* out, size_t ld);

unsigned round_up(unsigned x, unsigned m);
unsigned words_for_sites(unsigned sites);  // ceil(sites/32) rounded up to kWordPad
// Bytes of a launch workspace's saturation records for one job of n taxa, in the configuration
// that needs the most (one word per wave of the epilogue, 64 masks per flagged wave).
size_t saturation_bytes(unsigned n);
// Jobs in one launch: the statistics, maximum and patch kernels index jobs by grid y.
constexpr unsigned kMaxLaunchJobs = 65535;

// Allocate planes for an alignment of n taxa and `sites` counted sites (not initialised; the
// packers write every word, a protein alignment's gap counts are zeroed on ctx->stream).
int alloc_alignment(Context* ctx, unsigned n, unsigned L, unsigned sites, Alignment** out,
                    int kind = kDNA, unsigned code_bits = 0);

// dataloaderFasta.hpp:54-76 on the first sequence: DNA unless a byte >= 'A' is not ACGTUN.
bool first_row_is_dna(const char* row, unsigned L);

// Kind of an alignment whose first sequence is `row0`: the context's input type, or when that
// is RNJ_TYPE_AUTO the reference's discovery on the first sequence; DNA becomes kNaive under the
// context's naive character rules.
int alignment_kind(const Context* ctx, const char* row0, unsigned L);

// Pack n x L characters in device memory (row t at chars + t * pitch) into a new alignment of
// `kind`, enqueued on `stream` and synchronised. Protein alignments first take a census of the
// byte values to number the residues; under the context's naive character rules they count as
// the Hamming counter of the -f path does (Alignment::hamming).
// gap_rule (protein): kGapResolved, the characters dataloader::resolveChar maps to the gap
// (- . X x z Z b B J j ?); kGapBelow64, every byte below 64, as the reference's dormant protein
// kernel tests them (gpu/distanceCalculatorProtein_gpu.cu:76-90; its valid count is then the
// residue pairs exactly).
enum GapRule : int { kGapResolved = 0, kGapBelow64 = 1 };
int pack_alignment(Context* ctx, const uint8_t* chars, size_t pitch, unsigned n, unsigned L,
                   int kind, hipStream_t stream, Alignment** out, int gap_rule = kGapResolved);

// Parse a FASTA image and pack it (fasta.hip). `bytes` is in device memory, 16-byte aligned,
// unless `stage` is hipMemcpyHostToDevice or hipMemcpyDeviceToDevice: then it is first copied
// into a (cached) device buffer of the parse.
int fasta_to_alignment(Context* ctx, const uint8_t* bytes, size_t len, hipMemcpyKind stage,
                       hipStream_t stream, Alignment** out);
// The same for a (single-line) Stockholm image (fasta.hip).
int stockholm_to_alignment(Context* ctx, const uint8_t* bytes, size_t len, hipMemcpyKind stage,
                           hipStream_t stream, Alignment** out);

// Enqueue a multi-job launch on `stream`: reset statistics, tile kernel, saturation patch
// (unless raw), max copies. All jobs must share one tile configuration (chosen here). exact:
// the reference's expression with the library's FP64 log (else the leaner FP64 log).
// allow_tune: the first launch of an untuned shape class may time every configuration
// (synchronously); without it such launches use the size-based default.
int launch_jobs(Context* ctx, std::vector<JobSpec>& jobs, bool raw, bool exact,
                hipStream_t stream, bool allow_tune = true);

int choose_config(Context* ctx, unsigned n, unsigned W, bool jc_only);

// The pieces of launch_jobs, for callers that own their workspace (captured plans).
struct LaunchShape {
  bool jc_only = true;
  int kind = kDNA;
  unsigned code_bits = 0;
  bool mixed = false;  // jobs of different kinds or code widths (not launchable together)
  int mixed_kind = 0;  // the first job's kind and code width that differ from job 0's
  unsigned mixed_bits = 0;
  unsigned n_key = 0, W_key = 0;
  unsigned long long work = 0;
};
LaunchShape shape_of(const std::vector<JobSpec>& specs);
// Why the jobs of a mixed shape cannot share one launch (names the two kinds involved).
std::string mixed_message(const LaunchShape& sh);
int prepare_workspace(Workspace* ws, unsigned njobs);
int select_config(Context* ctx, std::vector<JobSpec>& specs, LaunchShape& sh, Workspace* ws,
                  bool exact, hipStream_t stream, bool allow_tune, int* config);
// upload_table: copy the job table to the device first (a caller replaying a captured launch
// whose table never changes uploads it once and captures without the copy).
int enqueue_launch(Context* ctx, std::vector<JobSpec>& specs, LaunchShape& sh, int config,
                   Workspace* ws, bool raw, bool exact, hipStream_t stream,
                   bool upload_table = true);
void free_workspace_public(Workspace* ws);

// One job's share of a launch (the pieces of plan_launch / number_jobs, for callers that plan a
// single job without the workspace machinery).
struct JobPlan {
  unsigned tiles = 0, splits = 1, wps = 0;  // pair tiles, word ranges per tile, words per range
  size_t partial_elems = 0;                 // u32 partial counts when split
  unsigned long long sat_units = 0;         // saturation wave units
};
unsigned job_tiles(const JobSpec& s, int config);
// Word ranges per tile that a launch of `total_tiles` tiles asks for (1: no split).
unsigned split_want(unsigned long long total_tiles, unsigned cus);
JobPlan plan_job(const JobSpec& s, int config, unsigned tiles, unsigned want);
// The kernels' description of one job: items from item_base, finish items from fin_base, its
// partial counts, maximum, and saturation records (regions of cap_units words, then tiles, then
// 64 masks per unit; this job's units from sat_off).
void describe_job(JobDesc& d, const JobSpec& s, int config, const JobPlan& jp,
                  unsigned long long item_base, unsigned long long fin_base, uint32_t* partial,
                  unsigned* maxbits, uint32_t* sat, size_t cap_units, size_t sat_off);
// The tile configuration launches get before (or without) tuning: t32_m4x4_c8.
constexpr int kDefaultTileConfig = 4;
// Device scratch of rnj_distances_device_noalloc (noalloc.hip): one array of this size per GPU.
constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;
// Takes the device address of that array on `device` (current) once; rnj_create calls it.
int scratch_setup(int device);
// The configuration a launch of this shape class gets without tuning (the context's forced
// configuration, a tuned one, else the default).
int current_config(Context* ctx, const LaunchShape& sh);
// The count kernels of a launch (tile kernel, and the finish kernel for split jobs).
hipError_t launch_counting(int config, const LaunchShape& sh, bool exact, const JobDesc* jobs,
                           unsigned njobs, unsigned items, unsigned finish_items,
                           hipStream_t stream);

}  // namespace rnj

struct rnj_context : rnj::Context {};
struct rnj_alignment : rnj::Alignment {};
