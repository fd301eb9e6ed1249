// Device code of the pair-tile kernels: bit helpers, the JC / K2P corrections, tile and job
// mapping, and the per-pair epilogue. Included by one translation unit (internal linkage).
#pragma once

#include "internal.hpp"

#include <math.h>

#include <type_traits>

namespace rnj {
namespace {

// popcount(x) + acc as one v_bcnt_u32_b32. Written out because LLVM otherwise re-associates
// the per-word accumulations into v_bcnt + v_add3 trees (1.5 extra VALU ops per pair and word).
__device__ __forceinline__ uint32_t bcnt_add(uint32_t x, uint32_t acc) {
  uint32_t r;
  asm("v_bcnt_u32_b32 %0, %1, %2" : "=v"(r) : "v"(x), "v"(acc));
  return r;
}

// The count loops' instructions as volatile asm, which the compiler keeps in program order.
// v_bitop3_b32's table is over a = 0xF0, b = 0xCC, c = 0xAA: 0x28 = (a ^ b) & c,
// 0x82 = ~(a ^ b) & c, 0x7D = (a ^ b) | ~c, 0x80 = a & b & c, 0xF6 = a | (b ^ c).
__device__ __forceinline__ uint32_t ordered_and(uint32_t a, uint32_t b) {
  uint32_t r;
  asm volatile("v_and_b32 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
  return r;
}
template <unsigned kTable>
__device__ __forceinline__ uint32_t ordered_bitop3(uint32_t a, uint32_t b, uint32_t c) {
  uint32_t r;
  asm volatile("v_bitop3_b32 %0, %1, %2, %3 bitop3:%4" : "=v"(r) : "v"(a), "v"(b), "v"(c), "i"(kTable));
  return r;
}
__device__ __forceinline__ void ordered_bcnt_add(uint32_t x, uint32_t& acc) {
  asm volatile("v_bcnt_u32_b32 %0, %1, %0" : "+v"(acc) : "v"(x));
}

// Pointer in the global address space, so loads compile to global_load rather than flat_load
// (the job table hands over generic pointers).
template <class T>
using gptr = const __attribute__((address_space(1))) T*;

template <class T>
using gptr_mut = __attribute__((address_space(1))) T*;

// Native 4 x u32 vector (ext_vector_type) for 128-bit global and LDS accesses.
typedef unsigned int v4u __attribute__((ext_vector_type(4)));

// Chunk `chunk` of the i and j taxon blocks into registers: linear element lin = q + Q*(w + CW*p)
// of a block is word row (chunk*CW + w) of plane p, uint4 column q (E = planes * CW * Q).
template <int EP, int NT, int Q, int CW, int E>
__device__ __forceinline__ void fetch_chunk(v4u (&pI)[EP], v4u (&pJ)[EP], gptr<v4u> gI,
                                            gptr<v4u> gJ, unsigned tid, unsigned chunk,
                                            size_t row4, size_t plane4) {
#pragma unroll
  for (int e = 0; e < EP; ++e) {
    const unsigned lin = tid + e * NT;
    if (E % NT != 0 && lin >= (unsigned)E) break;
    const unsigned q = lin % Q, w = (lin / Q) % CW, p = lin / (Q * CW);
    const size_t off = p * plane4 + (size_t)(chunk * CW + w) * row4 + q;
    pI[e] = gI[off];
    pJ[e] = gJ[off];
  }
}

// Partial counts of one (tile, split): [3][NT][MI*MJ] u32, thread-major so that every thread
// stores and loads its own contiguous 4*MI*MJ bytes.
template <int NT, int MI, int MJ>
__device__ __forceinline__ uint32_t* partial_slot(const JobDesc& J, unsigned tile,
                                                  unsigned split) {
  return J.partial + ((size_t)tile * J.splits + split) * 3 * NT * MI * MJ;
}

// ---- corrections ---------------------------------------------------------------------------
// Kimura 2-parameter (KimuraDistance.cpp:339-348) and Jukes-Cantor with alpha = 4
// (JCdistance.cpp:170-183). What decides the result's class is kept exactly as the reference
// computes it:
//  * the proportions: (float)((double)a / b) of integers below 2^24 equals the correctly rounded
//    float quotient (a double rounding 53 -> 24 bits is innocuous for a quotient), so the float
//    division is used, and the double one above 2^24;
//  * the arguments of the logarithms, 1-2P-Q and 1-2Q (K2P) and 1-p/0.75 (JC), formed in double
//    from the float proportions in the reference's order (exact for K2P), and the saturation
//    tests on them, so the saturated pairs (-1) are the same;
//  * zero distances: ts = tv = 0 gives -0.0f in both models, as the reference's -c*log(1.0) does.
// The logarithms of those double arguments use log_table (above) unless kExact, which calls the
// library's compensated FP64 log on the reference's expression; k2p_value / jc_value state how
// the default path regroups them and what that costs in accuracy.
#pragma clang fp contract(off)
__device__ __forceinline__ float proportion(uint32_t a, uint32_t valid) {
  if (valid < (1u << 24)) return (float)a / (float)valid;
  return (float)((double)a / (double)valid);
}

// Natural logarithm of a normal double 0 < z < 1 in FP64 arithmetic, through a 128-entry table
// (Tang's table-driven method): z = 2^k m with m in [0.5, 1); i = the mantissa's 7 bits below the
// leading one picks m's interval [0.5 + i/256, 0.5 + (i+1)/256) and its right end
// c = 0.5 + (i+1)/256 (the last is 1, so near z = 1 the reduced argument is z - 1 itself), and
// r = m/c - 1, with -2^-7 < r <= 0, is formed as fma(m, 1/c, -1): one rounding plus m times the
// rounding of the table's 1/c, below 2^-52 absolute (and exact on the last interval, where
// 1/c = 1). Then
//   ln z = k ln 2 + ln c + ln(1 + r),  ln(1 + r) = r - r^2/2 + r^3/3 - r^4/4 + r^5/5,
// whose truncation is below |r|^6/6 < 4e-14 absolute (and 5e-12 relative on the last interval,
// where ln c = 0). Where ln c != 0 and k = 0, |ln z| >= ln(256/255), so the relative error is
// below 2e-11 everywhere: rounded to float, the result equals the reference's libm value except
// in ties within ~1e-9 of a float rounding boundary. 1/c and ln c come rounded to double from the
// table (log_table.inc, written by gen_log_table.py); k ln 2 uses ln 2 split into a 32-bit head
// (k * head is exact) and a tail. z = 1 is not special-cased: reduced as 2^1 * 0.5 it gives a
// value within 1e-13 of 0: the DNA corrections select their exact result wherever z can be 1,
// and the protein ones use log_table_one.
// Any other argument (0, negative) reads an in-range table entry and returns a value that the
// callers discard. 15 VALU instructions and one LDS read: the kernel stages the table in LDS
// (stage_log_table), so that every lane reads its own entry without a divergent global load.
struct LogEntry {
  double inv_c, ln_c;
};

__device__ const LogEntry kLogTable[128] = {
#include "log_table.inc"
};

// The work-group's `nthreads` threads copy the table into `dst` (LDS); the caller synchronises
// before the first log_table call.
__device__ __forceinline__ void stage_log_table(LogEntry* dst, unsigned tid, unsigned nthreads) {
  for (unsigned t = tid; t < 128; t += nthreads) dst[t] = kLogTable[t];
}

// The FP64 constants of log_table as values held in SGPRs for a whole epilogue: a VOP3
// instruction takes no 64-bit literal, so constants written as literals were rebuilt with two
// v_mov for every pair an epilogue corrects; an empty asm with an SGPR operand makes each one an
// opaque uniform value that the FP64 instructions read straight from its SGPR pair.
struct LogConsts {
  double ln2_hi, ln2_lo, p5, p4, p3;
  const LogEntry* tab;  // the table in LDS
};

__device__ __forceinline__ LogConsts log_consts(const LogEntry* tab) {
  LogConsts K = {0x1.62e42ff000000p-1,    // ln 2, 32-bit head: k * head is exact
                 -0x1.718432a1b0e26p-35,  // ln 2 - head
                 1.0 / 5.0, -1.0 / 4.0, 1.0 / 3.0, tab};
  asm volatile("" : "+s"(K.ln2_hi));
  asm volatile("" : "+s"(K.ln2_lo));
  asm volatile("" : "+s"(K.p5));
  asm volatile("" : "+s"(K.p4));
  asm volatile("" : "+s"(K.p3));
  return K;
}

__device__ __forceinline__ double log_table(double z, const LogConsts& K) {
  int k;
  const double m = frexp(z, &k);  // m in [0.5, 1)
  const unsigned i = ((unsigned)__double2hiint(m) >> 13) & 127u;
  const LogEntry e = K.tab[i];
  const double r = fma(m, e.inv_c, -1.0);  // m/c - 1
  double p = fma(r, K.p5, K.p4);
  p = fma(p, r, K.p3);
  p = fma(p, r, -0.5);
  const double l1 = fma(r * r, p, r);
  const double kd = (double)k;
  return fma(kd, K.ln2_hi, fma(kd, K.ln2_lo, e.ln_c + l1));
}

// log_table with ln 1 = 0 exactly, for the protein corrections: their 64-bit valid counts (which
// the reference lets wrap) can make a proportion so small that the log argument rounds to 1.
__device__ __forceinline__ double log_table_one(double z, const LogConsts& K) {
  const double lz = log_table(z, K);
  return z == 1.0 ? 0.0 : lz;
}

// Two proportions a/v and b/v as correctly rounded float quotients, from one FP64 reciprocal
// of v (hardware estimate + two Newton steps: within about 2^-52 relative, and a * r adds one
// more rounding). For integers below 2^24 that is enough: a quotient a/v of such integers is
// never a float rounding midpoint (that would need a 25-bit odd numerator a/v' with v' | a), and
// its distance to the nearest midpoint is at least 1 / (v * 2^25) >= 2^-49 relative, so the
// rounded product lands on the same float as the exact quotient. Above 2^24 the double
// division of proportion() is used. kSmall: v < 2^24 is known, and no branch is taken.
__device__ __forceinline__ double reciprocal(uint32_t v) {
  const double dv = (double)v;
  double r = __builtin_amdgcn_rcp(dv);
  r = fma(r, fma(-dv, r, 1.0), r);
  return fma(r, fma(-dv, r, 1.0), r);
}

template <bool kSmall>
__device__ __forceinline__ void proportions2(uint32_t a, uint32_t b, uint32_t v, float& pa,
                                             float& pb) {
  if (kSmall || v < (1u << 24)) {
    const double r = reciprocal(v);
    pa = (float)((double)a * r);
    pb = (float)((double)b * r);
  } else {
    pa = (float)((double)a / (double)v);
    pb = (float)((double)b / (double)v);
  }
}

// One proportion a/v the same way (JC).
template <bool kSmall>
__device__ __forceinline__ float proportion1(uint32_t a, uint32_t v) {
  if (kSmall || v < (1u << 24)) return (float)((double)a * reciprocal(v));
  return (float)((double)a / (double)v);
}

// The corrections as straight-line code (every case computed, the result selected), so that a
// caller can interleave several pairs. kSmall: valid < 2^24 is known (fewer than 2^24 sites).
// The proportions and the saturation tests are always the reference's. For the logarithms,
// kExact evaluates the reference's expression with the library log; otherwise
//  * K2P takes one log: -0.5 ln x - 0.25 ln y = -0.25 ln(x^2 y), with x^2 y rounded twice
//    (2^-52 relative, i.e. 2^-54 absolute in the distance: below 1e-9 relative of any nonzero
//    distance, which is at least about 2^-24), and
//  * JC forms 1 - p/0.75 as fma(-p, 4/3, 1) (2^-52 absolute).
// Either way the float result equals the reference's unless the exact value lies within about
// 1e-9 relative of a float rounding boundary; it is always well inside 1e-6.
// With kSmall the reference's x = 1 - 2P - Q and y = 1 - 2Q are formed in one fma each, to the
// same doubles: P and Q are floats in [0, 1] quotients of integers below 2^24, so each is zero
// or at least 2^-24 with an ulp of at least 2^-47, and 1 - Q, x and y are multiples of 2^-47 in
// [-1, 1], exact in double whatever the order of the operations. For the same reason a nonzero
// x or y has magnitude at least 2^-47, which float holds, so the reference's saturation test on
// (float)x and (float)y is the test on x and y themselves.
// The corrections of the default path take no select on their log argument: where the pair is
// saturated or has no differences the argument is whatever x^2 y is, and the select at the end
// replaces the log's value; every other argument is below 1.
template <bool kExact, bool kSmall>
__device__ __forceinline__ float k2p_value(uint32_t ts, uint32_t tv, uint32_t valid,
                                           const LogConsts& K) {
  const uint32_t v = valid ? valid : 1u;
  float P, Q;
  if (kExact) {
    P = kSmall ? (float)ts / (float)v : proportion(ts, v);
    Q = kSmall ? (float)tv / (float)v : proportion(tv, v);
  } else {
    proportions2<kSmall>(ts, tv, v, P, Q);
  }
  double x, y;
  bool sat;
  if (kSmall && !kExact) {
    x = fma(-2.0, (double)P, 1.0 - (double)Q);
    y = fma(-2.0, (double)Q, 1.0);
    sat = valid == 0 || x <= 0.0 || y <= 0.0;
  } else {
    x = 1.0 - 2.0 * (double)P - (double)Q;
    y = 1.0 - 2.0 * (double)Q;
    sat = valid == 0 || (float)x <= 0.0f || (float)y <= 0.0f;
  }
  float d;
  if (kExact) {
    const double xs = sat ? 1.0 : x, ys = sat ? 1.0 : y;
    d = (float)(-0.5 * log(xs) - 0.25 * log(ys));
  } else {
    d = (float)(-0.25 * log_table(x * x * y, K));
  }
  // valid == 0 implies ts = tv = 0 (and sets sat); otherwise ts = tv = 0 gives x = y = 1
  return sat ? -1.0f : ((ts | tv) == 0 ? -0.0f : d);
}

template <bool kExact, bool kSmall>
__device__ __forceinline__ float jc_value(uint32_t mismatches, uint32_t valid,
                                          const LogConsts& K) {
  const double b = 0.75;  // (alpha - 1) / alpha
  const uint32_t v = valid ? valid : 1u;
  const float p = kExact ? (kSmall ? (float)mismatches / (float)v : proportion(mismatches, v))
                         : proportion1<kSmall>(mismatches, v);
  float d;
  bool sat;
  if (kExact) {
    const double q = (double)p / b;
    sat = q >= 1.0;
    d = (float)(-b * log(sat ? 1.0 : 1.0 - q));
  } else {
    sat = p >= 0.75f;  // the same test: a float p < 0.75 gives p / 0.75 < 1 - 2^-25 in double
    d = (float)(-b * log_table(fma(-(double)p, 4.0 / 3.0, 1.0), K));  // < 1 unless discarded
  }
  return (valid == 0 || mismatches == 0) ? -0.0f : (sat ? -1.0f : d);
}

#pragma clang fp contract(on)

// Row r of the block triangle holds tiles (r, r..nb-1); it starts at r*nb - r*(r-1)/2.
__device__ __forceinline__ void tile_coords(unsigned t, unsigned nb, unsigned& bi, unsigned& bj) {
  const double f = 2.0 * nb + 1.0;
  long long r = (long long)((f - sqrt(f * f - 8.0 * (double)t)) * 0.5);
  if (r < 0) r = 0;
  if (r > (long long)nb - 1) r = nb - 1;
  auto start = [nb](long long q) { return q * (long long)nb - q * (q - 1) / 2; };
  while (r + 1 < (long long)nb && start(r + 1) <= (long long)t) ++r;
  while (r > 0 && start(r) > (long long)t) --r;
  bi = (unsigned)r;
  bj = (unsigned)(r + ((long long)t - start(r)));
}

// Patch order of the pair tiles (bi <= bj) of an nb x nb block grid: block rows in bands of 8,
// each band walked column by column (the band's diagonal triangle first), so that any 32
// consecutive tiles touch about 8 row blocks and 4 column blocks. Row-major order would give
// 32 tiles over 33 blocks; here a group of work-groups resident on one XCD re-reads the same
// blocks from its L2. Bijective onto the upper triangle (checked on the host).
__host__ __device__ __forceinline__ void patch_coords(unsigned t, unsigned nb, unsigned& bi,
                                                      unsigned& bj) {
  unsigned R0 = 0, r = 0;
  for (;;) {
    r = nb - R0 < 8u ? nb - R0 : 8u;
    const unsigned cnt = r * (r + 1) / 2 + (nb - R0 - r) * r;
    if (t < cnt || R0 + r >= nb) break;
    t -= cnt;
    R0 += 8;
  }
  const unsigned tri = r * (r + 1) / 2;
  if (t < tri) {
    unsigned c = 0;
    while ((c + 1) * (c + 2) / 2 <= t) ++c;
    bi = R0 + (t - c * (c + 1) / 2);
    bj = R0 + c;
  } else {
    t -= tri;
    bi = R0 + t % r;
    bj = R0 + r + t / r;
  }
}

// Block coordinates of tile `tile` of job J: tiles 0 .. tri-1 cover the upper triangle of the
// block grid (bi <= bj) in patch order; a kRows job has tri more tiles, the same blocks
// transposed (bi >= bj), whose pairs (i, j) have i > j and are counted with taxon i as the row's
// taxon. Returns whether the tile is such a transposed one.
__device__ __forceinline__ bool job_tile(const JobDesc& J, unsigned tile, unsigned& bi,
                                         unsigned& bj) {
  const unsigned long long tri = (unsigned long long)J.nblk * (J.nblk + 1) / 2;
  const bool lower = tile >= tri;  // only kRows jobs number tiles past tri
  patch_coords(lower ? (unsigned)(tile - tri) : tile, J.nblk, bi, bj);
  if (lower) {
    const unsigned t = bi;
    bi = bj;
    bj = t;
  }
  return lower;
}

// In-register transpose of a 32 x 32 bit matrix, bit c of a[r] <-> bit r of a[c] (recursive
// swap of the off-diagonal blocks: 16 x 16, then 8 x 8 inside each, ... down to 1 x 1).
__device__ __forceinline__ void transpose32(uint32_t (&a)[32]) {
#pragma unroll
  for (int j = 16, m = 0x0000FFFF; j != 0; j >>= 1, m ^= m << j) {
#pragma unroll
    for (int k = 0; k < 32; ++k) {
      if (k & j) continue;
      const uint32_t t = ((a[k] >> j) ^ a[k + j]) & (uint32_t)m;
      a[k] ^= t << j;
      a[k + j] ^= t;
    }
  }
}

// Work-groups are dispatched round-robin over the 8 XCDs; give each XCD a contiguous range of
// items so that tiles sharing a block row also share that XCD's L2.
__device__ __forceinline__ unsigned xcd_item(unsigned bid, unsigned grid) {
  const unsigned q = grid / 8, r = grid % 8, x = bid % 8, k = bid / 8;
  return x < r ? x * (q + 1) + k : r * (q + 1) + (x - r) * q + k;
}

__device__ __forceinline__ unsigned find_job(const JobDesc* jobs, unsigned njobs, unsigned item) {
  unsigned lo = 0, hi = njobs - 1;
  while (lo < hi) {
    const unsigned mid = (lo + hi + 1) / 2;
    if (jobs[mid].item_base <= item) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// Saturation record of one wave of an epilogue, one 32-bit word per wave unit (unit = tile *
// waves-per-work-group + wave of the work-group): bit 31 set when a lane of the wave holds a
// saturated pair, bits 0-30 the wave's largest distance (distances are >= 0, so the bits of the
// float compare as integers; a -0.0 folds into +0.0). When bit 31 is set, every lane also stores
// its micro-tile mask (bit a * MJ + b: pair (i0 + a, j0 + b) saturated) at 64 * unit + lane, and
// lane 0 the tile's block coordinates bij = bi << 16 | bj.
// Only plain stores: job_max_kernel folds the words into the job's maximum and patch_kernel
// rewrites the marked pairs. (One atomic per wave on the job's counters, which this replaces,
// serialised at about 11 ns per atomic across the XCDs: 4.4 ms per launch at 20000 taxa.)
// Every lane of the wave must call it.
__device__ __forceinline__ void wave_record(const JobDesc& J, unsigned unit, unsigned bij,
                                            uint32_t mask, float lmax) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) lmax = fmaxf(lmax, __shfl_xor(lmax, off));
  const bool any = __ballot(mask != 0u) != 0ull;  // uniform over the wave
  const unsigned lane = threadIdx.x & 63;
  if (any) J.sat_mask[(size_t)unit * 64 + lane] = mask;
  if (lane == 0) {
    if (any) J.sat_tile[unit] = bij;
    J.sat_word[unit] = (any ? kSatFlag : 0u) | (__float_as_uint(lmax) & ~kSatFlag);
  }
}

// The counts of one finished pair (i < j < n) of a counts job, to its three upper triangles.
__device__ __forceinline__ void finish_counts(const JobDesc& J, unsigned i, unsigned j, uint32_t ts,
                                              uint32_t tv, uint32_t valid) {
  const unsigned long long p = tri_index(i, j, J.n);
  J.out_tv[p] = tv;
  J.out_ts[p] = ts;
  J.out_valid[p] = valid;
}

// Zero diagonal of a square output (the reference's post-process writes 0 on i == i).
__device__ __forceinline__ void finish_diagonal(const JobDesc& J, unsigned i) {
  if (J.model != kCounts && J.layout != kTriangle) J.out[(size_t)i * J.ld + i] = 0.0f;
}

// The distance of pair (i, j) into its output: the upper triangle, both halves of a square, or
// (kRows) entry (i, j) only.
__device__ __forceinline__ void store_pair(const JobDesc& J, unsigned i, unsigned j, float d) {
  if (J.layout == kTriangle) {
    J.out[tri_index(i, j, J.n)] = d;
  } else {
    J.out[(size_t)i * J.ld + j] = d;
    if (J.layout == kSquare) J.out[(size_t)j * J.ld + i] = d;
  }
}

}  // namespace
}  // namespace rnj
