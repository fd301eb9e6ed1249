// Internal declarations shared by the HIP kernels (kernels.hip) and the runtime (runtime.cpp).
//
// Packed alignment in HBM ("bit planes"). For taxon t and 32-site word w:
//   planes[0][w][t]  H bits: 1 for a pyrimidine (C/T), the high bit of the reference's 2-bit code
//   planes[1][w][t]  L bits: 1 for G/C, the low bit of that code (A=00 G=01 T=10 C=11)
//   planes[2][w][t]  V bits: 1 where the site holds one of ACGTacgt
// with site s of word w at bit s % 32. Taxa are the fastest index, so a row of taxa for one word
// is one contiguous, coalesced run. n_pad is n rounded up to kTaxaPad and the padded taxa are
// all-zero; W_pad is the word count rounded up to kWordPad and padded words are zero, so padding
// contributes nothing to any count. For a taxon pair (i, j) and one word:
//   g = Vi & Vj, tv = (Hi ^ Hj) & g, ts = (Li ^ Lj) & (g & ~tv)
// which is the reference's per-word logic (bitDistanceGap.cpp:138-161) on split planes.
//
// Protein alignments use 1 + B planes instead (protein.hip): planes[0] the gap bits, planes[1..B]
// the bits of each site's residue code, plus a per-taxon count of gap sites.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace rnj {

constexpr unsigned kTaxaPad = 128;    // n_pad multiple (largest tile edge)
constexpr unsigned kWordPad = 16;     // W_pad multiple (largest LDS site chunk, in words)
constexpr unsigned kBlockThreads = 256;
constexpr uint32_t kSentinelBits = 0xBF800000u;  // -1.0f, the reference's saturation marker
constexpr uint32_t kSatFlag = 0x80000000u;       // a wave's saturation word: it holds saturated pairs

enum Model : int { kJC = 0, kK2P = 1, kCounts = 2 };
// kNaive: DNA under the reference's naive character rules (simpleDistanceCalculator, `-f`),
// packed with a fourth plane.
enum Kind : int { kDNA = 0, kProtein = 1, kNaive = 2 };
// Count loops of the DNA tile kernel: K2P (and counts), JC only, naive character rules.
enum CountMode : int { kModeK2P = 0, kModeJC6 = 1, kModeNaive = 2 };
// kRows (internal to the row stream): square rows in which every entry (i, j), i != j, is counted
// with taxon i as the row's taxon, as the reference's disk-matrix path computes them
// (KimuraDistance.cpp:300-320 with offset 0). A kRows job covers the upper triangle of tile blocks
// and then the transposed one (see job_tile); only naive-rule alignments, whose counts are not
// symmetric, need it.
enum Layout : int { kTriangle = 0, kSquare = 1, kRows = 2 };

// One distance matrix inside a launch. A launch covers jobs[0..njobs); the work items of job k
// (tiles of TI x TJ taxon pairs over the upper triangle of tile blocks, each cut into `splits`
// ranges of words) are numbered item_base .. item_base + tiles * splits - 1, item = tile * splits
// + split. A split job leaves exact partial counts in `partial` and the finish kernel, numbering
// that job's tiles from fin_base, sums them and applies the epilogue.
struct JobDesc {
  const uint32_t* planes;     // [3][W_pad][n_pad] (DNA) or [1 + B][W_pad][n_pad] (protein)
  unsigned long long plane_stride;  // W_pad * n_pad
  unsigned n, n_pad, W;       // W: words to process (multiple of kWordPad, <= W_pad)
  unsigned W_live;            // words holding sites (<= W): the DNA tile kernel stops there
  unsigned nblk;              // ceil(n / TI) for the launch's tile edge (TI == TJ)
  unsigned tiles;             // nblk * (nblk + 1) / 2
  unsigned item_base;
  unsigned splits;            // word ranges per tile (1 = finish in the tile kernel)
  unsigned words_per_split;   // multiple of the configuration's chunk
  unsigned fin_base;          // first finish-kernel item of this job
  uint32_t* partial;          // [tiles][splits][3][threads][micro_i * micro_j] when splits > 1
  int model;                  // Model
  int layout;                 // Layout (distances only)
  unsigned long long ld;      // row stride in elements for kSquare
  float* out;                 // distances
  uint32_t* out_ts;           // counts (kCounts), upper triangle
  uint32_t* out_tv;
  uint32_t* out_valid;
  unsigned* maxbits;          // bits of the largest distance (non-negative floats), job_max_kernel
  uint32_t* sat_word;         // [tiles * waves per work-group]: kSatFlag | wave maximum bits
  uint32_t* sat_tile;         // [tiles * waves per work-group]: bi << 16 | bj, flagged waves
  uint32_t* sat_mask;         // [tiles * waves per work-group][64]: micro-tile masks, flagged waves
  const uint32_t* gaps;       // protein: gap sites per taxon [n_pad]
  unsigned L;                 // protein: alignment length (sites)
  // protein: 1 for bitDistanceProtein's valid count (128 short of the residue pairs when the
  // length is a multiple of 128), 0 for the residue pairs
  unsigned bias_rule;
  const unsigned* len_dev;    // protein: the length in device memory (bootstrap replicates), or null
};

// Upper-triangle position of pair (i, j), i < j: row i holds j = i+1 .. n-1 (row-major).
__host__ __device__ inline unsigned long long tri_index(unsigned long long i, unsigned long long j,
                                                       unsigned long long n) {
  return i * n - i * (i + 1) / 2 + (j - i - 1);
}

// Tile-kernel configurations (pair-tile edge, per-thread micro-tile, LDS site chunk).
struct TileConfig {
  const char* name;
  unsigned tile;     // TI == TJ
  unsigned micro_i;  // pairs per thread along i
  unsigned micro_j;  // pairs per thread along j
  unsigned chunk;    // words of 32 sites staged in LDS per step
};
extern const TileConfig kTileConfigs[];
extern const int kNumTileConfigs;

// Launchers (kernels.hip). All enqueue on `stream` and return the launch status.
// naive: the four planes of the naive character rules (H, L, V, U) instead of H, L, V.
hipError_t launch_pack_chars(const uint8_t* chars, size_t pitch, unsigned n, unsigned L,
                             unsigned n_pad, unsigned W_pad, uint32_t* planes,
                             hipStream_t stream, bool naive = false);
hipError_t launch_pack_legacy(const uint32_t* bits, const uint32_t* gaps, size_t stride,
                              unsigned words16, unsigned n, unsigned n_pad, unsigned W_pad,
                              uint32_t* planes, hipStream_t stream);
hipError_t launch_reset_stats(const JobDesc* jobs, unsigned njobs, hipStream_t stream);
// Writes one job description (a kernel argument) to dst in device memory.
hipError_t launch_store_job(const JobDesc& d, JobDesc* dst, hipStream_t stream);
// Writes the float whose bits are *maxbits to dst (device memory).
hipError_t launch_copy_max(const unsigned* maxbits, float* dst, hipStream_t stream);
// exact: the reference's expression with the library's FP64 log; otherwise the leaner FP64 log
// (device_common.hpp, within ~2 ulp of double).
// mode: CountMode (kModeJC6 only when every job is JC; kModeNaive for naive-rule alignments).
hipError_t launch_tiles(int config, int mode, bool exact, const JobDesc* jobs, unsigned njobs,
                        unsigned total_items, unsigned finish_items, hipStream_t stream);
// Protein (protein.hip). census: sets bit v of present[8] for every byte value v in the n x L
// characters (present must start at zero). pack: characters -> gap plane + B code planes using
// codes[256] (code of each byte value, 0 for gaps), gaps[t] += gap sites of taxon t (gaps must
// start at zero). B is 5, 6 or 8.
hipError_t launch_symbol_census(const uint8_t* chars, size_t pitch, unsigned n, unsigned L,
                                uint32_t* present, hipStream_t stream);
hipError_t launch_pack_protein(const uint8_t* chars, size_t pitch, unsigned n, unsigned L,
                               unsigned n_pad, unsigned W_pad, const uint8_t* codes, unsigned B,
                               uint32_t* planes, uint32_t* gaps, hipStream_t stream);
hipError_t launch_protein_tiles(int config, unsigned B, bool exact, const JobDesc* jobs,
                                unsigned njobs, unsigned total_items, unsigned finish_items,
                                hipStream_t stream);
hipError_t launch_lower_pack(const float* sq, size_t ld, unsigned n, float* out,
                             hipStream_t stream);
unsigned tile_threads(int config);  // work-group size of a configuration
// Issue-rate probe (op 0: v_bitop3_b32, 1: v_bcnt_u32_b32): blocks x 256 threads, each running
// 16 dependency chains for iters x 32 instructions; out holds blocks x 256 u32.
hipError_t launch_issue_probe(int op, uint32_t* out, unsigned blocks, int iters, hipStream_t s);
// Count-loop instruction mixes timed by the mix probe (values of rnj_valu_mix_rate's `mix`).
enum CountMix : int { kMixK2P = 0, kMixJC = 1, kMixNaive = 2, kMixProtein = 3 };
int mix_ops_per_pair(int mix, int planes);
hipError_t launch_mix_probe(int mix, int planes, uint32_t* out, unsigned blocks, int iters,
                            hipStream_t s);

// After the tile (and finish) kernels of a launch: job_max_kernel folds each job's wave words
// into its maximum, then, unless raw, patch_kernel writes 2 * maximum over the saturated pairs
// the flagged waves marked. max_units: the largest job's wave-unit count; config: the launch's
// tile configuration (the geometry that maps a wave's mask bits back to pairs).
hipError_t launch_saturation(const JobDesc* jobs, unsigned njobs, unsigned long long max_units,
                             int config, bool patch, hipStream_t stream);
// Wave units of a job of `tiles` pair tiles in configuration `config` (one saturation word each).
unsigned long long sat_units(unsigned long long tiles, int config);
// Column gathers of bootstrap replicates read per-site codes: DNA one nibble per taxon
// (U<<3|V<<2|H<<1|L, U only under the naive rules), two taxa per byte, site-major:
// codes_bytes(n_pad, L) bytes; protein one byte per taxon, the residue code (0 = gap):
// protein_codes_bytes(n_pad, L).
inline size_t codes_bytes(unsigned n_pad, unsigned L) { return (size_t)L * n_pad / 2; }
inline size_t protein_codes_bytes(unsigned n_pad, unsigned L) { return (size_t)L * n_pad; }
hipError_t launch_planes_to_codes(const uint32_t* planes, unsigned n_pad, unsigned W_pad,
                                  unsigned L, uint32_t* codes, hipStream_t stream,
                                  bool naive = false);
hipError_t launch_bootstrap_index(const uint16_t* weights, unsigned R, unsigned L,
                                  unsigned* col_index, unsigned idx_stride, unsigned* lengths,
                                  hipStream_t stream);
// naive: four planes per replicate (H, L, V, U) instead of three.
hipError_t launch_bootstrap_gather(const uint32_t* codes, unsigned n_pad, const unsigned* col_index,
                                   unsigned idx_stride, const unsigned* lengths, unsigned R,
                                   unsigned W_pad, uint32_t* planes, hipStream_t stream,
                                   bool naive = false);
// Protein (protein.hip): per-site code bytes of an alignment of 1 + B planes, and the replicate
// gather: replicate r at planes + r * rep_words, its 1 + B planes of W_pad words, then its
// per-taxon gap counts [n_pad] (recounted over the replicate's sites).
hipError_t launch_protein_codes(const uint32_t* planes, unsigned n_pad, unsigned W_pad,
                                unsigned L, unsigned B, uint8_t* codes, hipStream_t stream);
hipError_t launch_protein_gather(const uint8_t* codes, unsigned n_pad, unsigned B,
                                 const unsigned* col_index, unsigned idx_stride,
                                 const unsigned* lengths, unsigned R, unsigned W_pad,
                                 uint32_t* planes, size_t rep_words, hipStream_t stream);

}  // namespace rnj
