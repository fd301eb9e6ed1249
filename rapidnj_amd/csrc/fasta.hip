// FASTA ingest on the device: a file image in HBM -> record names + packed H/L/V planes.
//
// The reference reads FASTA with a byte-at-a-time state machine (dataloaderFasta.hpp:78-133,
// 201-217, 253-276), bytes taken as signed chars:
//   * the main loop reads one byte: ';' skips to the end of the line, '>' ends the current
//     record and starts a name (the next byte is its first candidate), anything else is a
//     sequence line;
//   * a name takes bytes > 31 and stops at the first other byte, which is consumed; if that was
//     not the line's '\n', the main loop resumes right after it, on the same line;
//   * a sequence line appends every byte > 32 up to its '\n' (so ';' and '>' inside a line are
//     data, and spaces, tabs, '\r' and bytes >= 128 are dropped);
//   * end of input acts as a final '\n';
//   * a record is stored only if it holds at least one byte; the first stored record fixes the
//     sequence length and every later one must match it; every '>' contributes a name, stored
//     or not, so names follow headers and not stored sequences.
//
// The device restates that per line, in parallel: each physical line is a chain of zero or more
// headers (every one but the last ended by a control byte other than '\n') followed by at most
// one data span (to the end of the line) or a comment. Data bytes go to the record opened by the
// last header before them. Passes (all over HBM, 16 KB per work-group):
//   1. count and list the '\n' positions (per-block counts, scan, ordered write);
//   2. one wave per line: its header count, data span and kept-byte count;
//   3. scans: header index and record of each line, byte offset of each line in the data stream;
//   4. records: length and first byte offset (atomics), stored-record numbering, length check;
//   5. one wave per data line: kept bytes scattered into an n x L character matrix;
//   6. the character matrix is packed into planes by the alignment packer; names are gathered.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "runtime.hpp"

namespace rnj {
namespace {

constexpr unsigned kThreads = 256;
constexpr unsigned kRounds = 4;                       // 16-byte loads per thread and block
constexpr size_t kChunk = (size_t)kThreads * kRounds * 16;  // bytes per work-group in pass 1

__device__ __forceinline__ int sbyte(const uint8_t* b, size_t p) { return (int)(int8_t)b[p]; }

// Bytes of x equal to 0x0A ('\n') as set high bits (exact: no false positives).
__device__ __forceinline__ uint32_t nl_bits(uint32_t x) {
  const uint32_t v = x ^ 0x0A0A0A0Au;
  return ~(((v & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | v | 0x7F7F7F7Fu);
}

// 16 bytes at p (p < N); bytes at or past N read as 0 (not '\n').
__device__ __forceinline__ uint4 load16(const uint8_t* b, size_t N, size_t p) {
  if (p + 16 <= N) return *reinterpret_cast<const uint4*>(b + p);
  uint32_t w[4] = {0, 0, 0, 0};
  for (size_t k = p; k < N; ++k) w[(k - p) / 4] |= (uint32_t)b[k] << (8 * ((k - p) % 4));
  return make_uint4(w[0], w[1], w[2], w[3]);
}

__device__ __forceinline__ unsigned count16(const uint4& q) {
  return __popc(nl_bits(q.x)) + __popc(nl_bits(q.y)) + __popc(nl_bits(q.z)) + __popc(nl_bits(q.w));
}

// Bytes of x in 33..127 (signed chars > 32: what a sequence line keeps) as set high bits.
__device__ __forceinline__ uint32_t kept_bits(uint32_t x) {
  return ((x & 0x7F7F7F7Fu) + 0x5F5F5F5Fu) & ~x & 0x80808080u;
}

__device__ __forceinline__ unsigned kept16(const uint4& q) {
  return __popc(kept_bits(q.x)) + __popc(kept_bits(q.y)) + __popc(kept_bits(q.z)) +
         __popc(kept_bits(q.w));
}

// Pass 1a: '\n' count of each 16 KB block, and the kept-byte count of each 16-byte chunk.
__global__ __launch_bounds__(kThreads) void nl_count_kernel(const uint8_t* __restrict__ b, size_t N,
                                                            uint32_t* __restrict__ blk,
                                                            uint8_t* __restrict__ kc) {
  const size_t base = (size_t)blockIdx.x * kChunk;
  unsigned c = 0;
#pragma unroll
  for (unsigned r = 0; r < kRounds; ++r) {
    const size_t p = base + ((size_t)r * kThreads + threadIdx.x) * 16;
    if (p < N) {
      const uint4 q = load16(b, N, p);
      c += count16(q);
      kc[p / 16] = (uint8_t)kept16(q);
    }
  }
  __shared__ unsigned s[kThreads / 64];
  for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off);
  if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) blk[blockIdx.x] = s[0] + s[1] + s[2] + s[3];
}

// Exclusive prefix of v over the work-group (256 threads), and the total.
__device__ __forceinline__ unsigned block_exclusive(unsigned v, unsigned& total) {
  __shared__ unsigned s[kThreads / 64];
  const unsigned lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  unsigned x = v;
  for (int off = 1; off < 64; off <<= 1) {
    const unsigned y = __shfl_up(x, off);
    if (lane >= (unsigned)off) x += y;
  }
  if (lane == 63) s[wave] = x;
  __syncthreads();
  unsigned before = 0;
  for (unsigned w = 0; w < wave; ++w) before += s[w];
  total = s[0] + s[1] + s[2] + s[3];
  __syncthreads();
  return before + x - v;
}

// Pass 1b: positions of the '\n' bytes, in file order.
__global__ __launch_bounds__(kThreads) void nl_write_kernel(const uint8_t* __restrict__ b, size_t N,
                                                            const unsigned long long* __restrict__ blk_base,
                                                            unsigned long long* __restrict__ pos) {
  const size_t base = (size_t)blockIdx.x * kChunk;
  unsigned long long out = blk_base[blockIdx.x];
  for (unsigned r = 0; r < kRounds; ++r) {
    const size_t p = base + ((size_t)r * kThreads + threadIdx.x) * 16;
    uint4 q = make_uint4(0, 0, 0, 0);
    if (p < N) q = load16(b, N, p);
    const uint32_t w[4] = {nl_bits(q.x), nl_bits(q.y), nl_bits(q.z), nl_bits(q.w)};
    const unsigned c = __popc(w[0]) + __popc(w[1]) + __popc(w[2]) + __popc(w[3]);
    unsigned total;
    const unsigned ex = block_exclusive(c, total);
    unsigned k = 0;
    for (int d = 0; d < 4; ++d) {
      uint32_t m = w[d];
      while (m) {
        const int bit = __ffs(m) - 1;
        pos[out + ex + k++] = p + 4 * d + bit / 8;
        m &= m - 1;
      }
    }
    out += total;
  }
}

struct Line {
  unsigned long long start, end;  // [start, end): the line without its '\n'
};

__device__ __forceinline__ Line line_at(const unsigned long long* nl, unsigned long long n_nl,
                                        size_t N, unsigned long long l) {
  Line L;
  L.start = l == 0 ? 0 : nl[l - 1] + 1;
  L.end = l < n_nl ? nl[l] : N;
  return L;
}

// Walks the header chain of a line (every lane the same, so control flow stays uniform).
// Returns the header count; *ds/*de receive the data span (empty if none). If spans is not
// null the name spans [s, e) are written to spans[0..count).
__device__ __forceinline__ unsigned walk_line(const uint8_t* b, Line L, unsigned long long* ds,
                                              unsigned long long* de,
                                              unsigned long long (*spans)[2]) {
  unsigned h = 0;
  unsigned long long p = L.start;
  *ds = *de = L.start;
  while (p < L.end) {
    const int c = sbyte(b, p);
    if (c == ';') break;                       // comment: the rest of the line is skipped
    if (c != '>') {                            // data to the end of the line
      *ds = p;
      *de = L.end;
      break;
    }
    unsigned long long q = p + 1;              // name: bytes > 31 after '>'
    while (q < L.end && sbyte(b, q) > 31) ++q;
    if (spans) {
      spans[h][0] = p + 1;
      spans[h][1] = q;
    }
    ++h;
    if (q >= L.end) break;                     // the name ended at '\n' (or the end of input)
    p = q + 1;                                 // another control byte: resume after it
  }
  return h;
}

// Kept bytes (33..127) in [0, x): chunk prefix Kc (Kc[N/16 rounded up] = total) plus the
// chunk's bytes before x.
__device__ __forceinline__ unsigned long long kept_before(const uint8_t* b,
                                                          const unsigned long long* Kc,
                                                          unsigned long long x) {
  const unsigned long long c = x / 16;
  unsigned long long k = Kc[c];
  for (unsigned long long p = c * 16; p < x; ++p) k += (unsigned)(sbyte(b, p) > 32);
  return k;
}

// Pass 2: one thread per line: header count, data span, kept bytes of the span, and the kept
// count before the span (for the scatter's in-line rank).
__global__ void line_kernel(const uint8_t* __restrict__ b, size_t N,
                            const unsigned long long* __restrict__ nl, unsigned long long n_nl,
                            unsigned long long M, const unsigned long long* __restrict__ Kc,
                            uint32_t* __restrict__ hcount, unsigned long long* __restrict__ dspan,
                            uint32_t* __restrict__ dcount, unsigned long long* __restrict__ kds) {
  const unsigned long long l = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= M) return;
  const Line L = line_at(nl, n_nl, N, l);
  unsigned long long ds, de;
  const unsigned h = walk_line(b, L, &ds, &de, nullptr);
  const unsigned long long k0 = kept_before(b, Kc, ds);
  const unsigned long long k1 = ds == de ? k0 : kept_before(b, Kc, de);
  hcount[l] = h;
  dspan[2 * l] = ds;
  dspan[2 * l + 1] = de;
  dcount[l] = (uint32_t)(k1 - k0);
  kds[l] = k0;
}

// Name spans of every header, at their header index.
__global__ void header_kernel(const uint8_t* __restrict__ b, size_t N,
                              const unsigned long long* __restrict__ nl, unsigned long long n_nl,
                              unsigned long long M, const uint32_t* __restrict__ hcount,
                              const uint32_t* __restrict__ hbase,
                              unsigned long long* __restrict__ spans) {
  const unsigned long long l = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= M || hcount[l] == 0) return;
  const Line L = line_at(nl, n_nl, N, l);
  unsigned h = 0;
  unsigned long long p = L.start;
  while (p < L.end && h < hcount[l]) {
    unsigned long long q = p + 1;
    while (q < L.end && sbyte(b, q) > 31) ++q;
    spans[2 * (hbase[l] + h)] = p + 1;
    spans[2 * (hbase[l] + h) + 1] = q;
    ++h;
    p = q + 1;
  }
}

// Pass 4a: first byte offset of every record. Records are contiguous runs of lines (the record
// of a line, hbase + hcount, never decreases); the records opened on line l start at its
// offset, which also makes records opened and closed on one line empty.
__global__ void record_base_kernel(unsigned long long M, const uint32_t* __restrict__ hcount,
                                   const uint32_t* __restrict__ hbase,
                                   const unsigned long long* __restrict__ off,
                                   unsigned long long* __restrict__ rec_base) {
  const unsigned long long l = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= M) return;
  const unsigned hi = hbase[l] + hcount[l];
  const unsigned lo = l == 0 ? 0u : hbase[l - 1] + hcount[l - 1] + 1u;
  for (unsigned r = lo; r <= hi; ++r) rec_base[r] = off[l];
}

__global__ void record_len_kernel(unsigned R, const unsigned long long* __restrict__ rec_base,
                                  const unsigned long long* __restrict__ total,
                                  unsigned long long* __restrict__ rec_len) {
  const unsigned r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < R) rec_len[r] = (r + 1 < R ? rec_base[r + 1] : *total) - rec_base[r];
}

__global__ void stored_flag_kernel(unsigned R, const unsigned long long* __restrict__ rec_len,
                                   uint32_t* __restrict__ stored, unsigned* __restrict__ first) {
  const unsigned r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= R) return;
  stored[r] = rec_len[r] > 0;
  if (rec_len[r] > 0) atomicMin(first, r);
}

// Pass 4b: the first stored record whose length differs from the first one (as a sequence index).
__global__ void length_check_kernel(unsigned R, const unsigned long long* __restrict__ rec_len,
                                    const uint32_t* __restrict__ seq, const unsigned* __restrict__ first,
                                    unsigned* __restrict__ bad) {
  const unsigned r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= R || rec_len[r] == 0 || *first >= R) return;
  if (rec_len[r] != rec_len[*first]) atomicMin(bad, seq[r]);
}

// Pass 5a: destination of each data line's first kept byte in the n x L matrix, less the kept
// count before the line's span (so that a byte's destination is dbase + its kept rank in the file).
__global__ void line_dest_kernel(unsigned long long M, const uint32_t* __restrict__ hcount,
                                 const uint32_t* __restrict__ hbase,
                                 const unsigned long long* __restrict__ dspan,
                                 const unsigned long long* __restrict__ kds,
                                 const unsigned long long* __restrict__ off,
                                 const unsigned long long* __restrict__ rec_base,
                                 const uint32_t* __restrict__ seq, unsigned long long L,
                                 unsigned long long* __restrict__ dbase) {
  const unsigned long long l = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= M) return;
  const unsigned r = hbase[l] + hcount[l];
  dbase[l] = dspan[2 * l] < dspan[2 * l + 1]
                 ? (unsigned long long)seq[r] * L + (off[l] - rec_base[r]) - kds[l]
                 : 0ull;
}

// Pass 5b: one wave per 4 KB of the file, lanes on consecutive bytes (coalesced loads, and
// consecutive destinations within a line). The wave finds the line of its first byte (binary
// search over the '\n' positions); per 64-byte window, ballots give every lane its line (line
// at the window start + '\n' bytes below it) and its kept rank (kept bytes before the window +
// kept bytes below it). A kept byte inside its line's data span goes to dbase[line] + rank.
constexpr unsigned kScatterBytes = 4096;

__global__ __launch_bounds__(256) void scatter_kernel(const uint8_t* __restrict__ b, size_t N,
                                                      const unsigned long long* __restrict__ nl,
                                                      unsigned long long n_nl,
                                                      const unsigned long long* __restrict__ Kc,
                                                      const unsigned long long* __restrict__ dspan,
                                                      const unsigned long long* __restrict__ dbase,
                                                      uint8_t* __restrict__ chars) {
  const unsigned lane = threadIdx.x & 63;
  const unsigned long long wave = ((unsigned long long)blockIdx.x * blockDim.x + threadIdx.x) / 64;
  const unsigned long long p0 = wave * kScatterBytes;
  if (p0 >= N) return;  // uniform over the wave
  unsigned long long lo = 0, hi = n_nl;
  while (lo < hi) {
    const unsigned long long mid = (lo + hi) / 2;
    if (nl[mid] < p0) lo = mid + 1; else hi = mid;
  }
  unsigned long long line = lo, rank = Kc[p0 / 16];
  const unsigned long long pend = p0 + kScatterBytes < N ? p0 + kScatterBytes : N;
  for (unsigned long long w = p0; w < pend; w += 64) {
    const unsigned long long p = w + lane;
    const bool in = p < pend;
    const uint8_t c = in ? b[p] : 0;
    const bool is_nl = in && c == '\n';
    const bool kept = in && (int)(int8_t)c > 32;
    const unsigned long long nlm = __ballot(is_nl), km = __ballot(kept);
    const unsigned nl_below = __builtin_amdgcn_mbcnt_hi((unsigned)(nlm >> 32),
                                                        __builtin_amdgcn_mbcnt_lo((unsigned)nlm, 0u));
    const unsigned k_below = __builtin_amdgcn_mbcnt_hi((unsigned)(km >> 32),
                                                       __builtin_amdgcn_mbcnt_lo((unsigned)km, 0u));
    if (kept) {
      const unsigned long long my = line + nl_below;
      if (p >= dspan[2 * my] && p < dspan[2 * my + 1]) chars[dbase[my] + rank + k_below] = c;
    }
    line += (unsigned)__popcll(nlm);
    rank += (unsigned)__popcll(km);
  }
}

__global__ void name_gather_kernel(const uint8_t* __restrict__ b, unsigned H,
                                   const unsigned long long* __restrict__ spans,
                                   const unsigned long long* __restrict__ noff,
                                   uint8_t* __restrict__ blob) {
  const unsigned h = blockIdx.x;
  if (h >= H) return;
  const unsigned long long s = spans[2 * h], e = spans[2 * h + 1];
  for (unsigned long long p = s + threadIdx.x; p < e; p += blockDim.x) blob[noff[h] + (p - s)] = b[p];
}

__global__ void name_len_kernel(unsigned H, const unsigned long long* __restrict__ spans,
                                unsigned long long* __restrict__ len) {
  const unsigned h = blockIdx.x * blockDim.x + threadIdx.x;
  if (h < H) len[h] = spans[2 * h + 1] - spans[2 * h];
}

// ---- Stockholm ------------------------------------------------------------------------------
// The reference reads (single-line) Stockholm with another byte-at-a-time state machine
// (dataloaderStockholm.hpp:83-111, 118-132, 183-236, 294-314), bytes as signed chars:
//   * the main loop skips bytes below 32 (blank lines included); then '#' is a markup line
//     (skipped to its '\n'), "//" ends the alignment, anything else is a sequence line;
//   * a sequence's name is its first run of bytes above 32; then bytes below 32 are skipped,
//     each read after the first ending the input with "has zero length" if it is a '\n': a name
//     that ends its line takes its data from the next line, one followed only by tabs and a
//     '\n' is an error; the data is every byte above 32 up to the next '\n';
//   * reading past the end of the input before "//" is an error.
// Per line this is a map of two states, S (the main loop at the line's start) and P (a name's
// data pending from the previous line), onto S, P, E (ended) and X (error). The maps are
// composed by a scan over the lines; then every line knows its state and what it holds.
enum SthState : uint8_t { kSthS = 0, kSthP = 1, kSthE = 2, kSthX = 3 };
enum SthError : uint8_t { kSthNoError = 0, kSthZeroLength = 1, kSthEof = 2 };

__device__ __forceinline__ unsigned sth_apply(unsigned map, unsigned st) {
  return st < 2 ? (map >> (2 * st)) & 3u : st;
}

struct SthLine {
  uint8_t out, err;                  // the state after the line; why it is X
  bool named;                        // a sequence's name starts on this line
  unsigned long long ns, ne, ds, de;  // name and data spans (empty when absent)
};

// What line L does when the main loop reaches it in state `st` (kSthS or kSthP); `ended`: the
// line has its '\n' (the last line of an input without a final '\n' has not).
__device__ SthLine sth_walk(const uint8_t* b, Line L, bool ended, unsigned st) {
  SthLine r = {kSthX, kSthEof, false, L.start, L.start, L.start, L.start};
  unsigned long long p = L.start;
  auto data_from = [&](unsigned long long q) {  // data from q to the line's end
    r.ds = q;
    r.de = L.end;
    r.out = ended ? kSthS : kSthX;
    r.err = ended ? kSthNoError : kSthEof;
  };
  if (st == kSthP) {  // the skip after a name that ended the previous line
    while (p < L.end && sbyte(b, p) < 32) ++p;
    if (p == L.end) r.err = ended ? kSthZeroLength : kSthEof;
    else data_from(p);
    return r;
  }
  while (p < L.end && sbyte(b, p) < 32) ++p;
  if (p == L.end) {  // blank: the main loop goes on to the next line
    if (ended) r.out = kSthS, r.err = kSthNoError;
    return r;
  }
  const int c = sbyte(b, p);
  if (c == '#') {
    if (ended) r.out = kSthS, r.err = kSthNoError;
    return r;
  }
  if (c == '/' && p + 1 < L.end && sbyte(b, p + 1) == '/') {
    r.out = kSthE;
    r.err = kSthNoError;
    return r;
  }
  unsigned long long q = p;  // the name: bytes > 32 (a leading '/' is one of them)
  while (q < L.end && sbyte(b, q) > 32) ++q;
  r.named = true;
  r.ns = p;
  r.ne = q;
  if (q == L.end) {  // the name ends the line: its data is on the next one
    if (ended) r.out = kSthP, r.err = kSthNoError;
    return r;
  }
  if (sbyte(b, q) == 32) {  // data from the space on
    data_from(q);
    return r;
  }
  unsigned long long t = q + 1;  // tabs, '\r', other controls, bytes >= 128
  while (t < L.end && sbyte(b, t) < 32) ++t;
  if (t == L.end) r.err = ended ? kSthZeroLength : kSthEof;
  else data_from(t);
  return r;
}

// Map of each line: bits 0-1 the state after it from S, bits 2-3 from P.
__global__ void sth_map_kernel(const uint8_t* __restrict__ b, size_t N,
                               const unsigned long long* __restrict__ nl, unsigned long long n_nl,
                               unsigned long long M, uint8_t* __restrict__ map) {
  const unsigned long long l = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= M) return;
  const Line L = line_at(nl, n_nl, N, l);
  const bool ended = l < n_nl;
  map[l] = (uint8_t)(sth_walk(b, L, ended, kSthS).out | (sth_walk(b, L, ended, kSthP).out << 2));
}

// The state at every line's start (one work-group: each thread composes a run of lines, thread 0
// chains the runs, then each thread walks its run again); *final = the state after the last line.
__global__ __launch_bounds__(256) void sth_state_kernel(const uint8_t* __restrict__ map,
                                                        unsigned long long M,
                                                        uint8_t* __restrict__ state,
                                                        uint8_t* __restrict__ final_state) {
  __shared__ uint8_t agg[256], entry[257];
  const unsigned t = threadIdx.x;
  const unsigned long long run = (M + 255) / 256;
  const unsigned long long lo = t * run, hi = lo + run < M ? lo + run : M;
  unsigned m = kSthS | (kSthP << 2);  // identity
  for (unsigned long long l = lo; l < hi; ++l)
    m = sth_apply(map[l], sth_apply(m, kSthS)) | (sth_apply(map[l], sth_apply(m, kSthP)) << 2);
  agg[t] = (uint8_t)m;
  __syncthreads();
  if (t == 0) {
    unsigned st = kSthS;
    for (unsigned k = 0; k < 256; ++k) {
      entry[k] = (uint8_t)st;
      st = sth_apply(agg[k], st);
    }
    entry[256] = (uint8_t)st;
    *final_state = (uint8_t)st;
  }
  __syncthreads();
  unsigned st = entry[t];
  for (unsigned long long l = lo; l < hi; ++l) {
    state[l] = (uint8_t)st;
    st = sth_apply(map[l], st);
  }
}

// Each line in its state: whether a name starts on it, the name and data spans, the data's kept
// bytes and the kept count before the span; the first line that ends in an error.
__global__ void sth_line_kernel(const uint8_t* __restrict__ b, size_t N,
                                const unsigned long long* __restrict__ nl, unsigned long long n_nl,
                                unsigned long long M, const unsigned long long* __restrict__ Kc,
                                const uint8_t* __restrict__ state, uint32_t* __restrict__ named,
                                unsigned long long* __restrict__ nspan,
                                unsigned long long* __restrict__ dspan,
                                uint32_t* __restrict__ dcount, unsigned long long* __restrict__ kds,
                                unsigned long long* __restrict__ err_at) {
  const unsigned long long l = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= M) return;
  const Line L = line_at(nl, n_nl, N, l);
  SthLine r = {kSthE, kSthNoError, false, L.start, L.start, L.start, L.start};
  if (state[l] < 2) r = sth_walk(b, L, l < n_nl, state[l]);
  named[l] = r.named;
  nspan[2 * l] = r.ns;
  nspan[2 * l + 1] = r.ne;
  dspan[2 * l] = r.ds;
  dspan[2 * l + 1] = r.de;
  const unsigned long long k0 = kept_before(b, Kc, r.ds);
  const unsigned long long k1 = r.ds == r.de ? k0 : kept_before(b, Kc, r.de);
  dcount[l] = (uint32_t)(k1 - k0);
  kds[l] = k0;
  if (r.out == kSthX && r.err != kSthNoError) atomicMin(err_at, (l << 2) | r.err);
}

// Every data line's sequence (the last name before or on it), that sequence's length and data
// line, and the first sequence whose length differs from sequence 0's.
__global__ void sth_record_kernel(unsigned long long M, const uint32_t* __restrict__ named,
                                  const uint32_t* __restrict__ nbase,
                                  const unsigned long long* __restrict__ dspan,
                                  const uint32_t* __restrict__ dcount,
                                  unsigned long long* __restrict__ rec_len,
                                  unsigned long long* __restrict__ rec_line) {
  const unsigned long long l = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= M || dspan[2 * l] == dspan[2 * l + 1]) return;
  const unsigned r = nbase[l] + named[l] - 1;
  rec_len[r] = dcount[l];
  rec_line[r] = l;
}

__global__ void sth_length_check_kernel(unsigned R, const unsigned long long* __restrict__ rec_len,
                                        const unsigned long long* __restrict__ rec_line,
                                        unsigned long long* __restrict__ bad) {
  const unsigned r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < R && rec_line[r] != ~0ull && rec_len[r] != rec_len[0])
    atomicMin(bad, (rec_line[r] << 32) | r);
}

// Destination of each data line's first kept byte (less the kept count before its span).
__global__ void sth_dest_kernel(unsigned long long M, const uint32_t* __restrict__ named,
                                const uint32_t* __restrict__ nbase,
                                const unsigned long long* __restrict__ dspan,
                                const unsigned long long* __restrict__ kds, unsigned long long L,
                                unsigned long long* __restrict__ dbase) {
  const unsigned long long l = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= M) return;
  dbase[l] = dspan[2 * l] < dspan[2 * l + 1]
                 ? (unsigned long long)(nbase[l] + named[l] - 1) * L - kds[l]
                 : 0ull;
}

// Name span of every sequence, at its index.
__global__ void sth_name_kernel(unsigned long long M, const uint32_t* __restrict__ named,
                                const uint32_t* __restrict__ nbase,
                                const unsigned long long* __restrict__ nspan,
                                unsigned long long* __restrict__ spans) {
  const unsigned long long l = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= M || !named[l]) return;
  spans[2 * nbase[l]] = nspan[2 * l];
  spans[2 * nbase[l] + 1] = nspan[2 * l + 1];
}

// ---- exclusive scan (three phases, recursive over block totals) ------------------------------
template <class Tin, class Tout>
__global__ __launch_bounds__(kThreads) void scan_blocks_kernel(const Tin* __restrict__ in,
                                                               unsigned long long n,
                                                               Tout* __restrict__ out,
                                                               Tout* __restrict__ sums) {
  constexpr unsigned kItems = 8;
  const unsigned long long base = (unsigned long long)blockIdx.x * kThreads * kItems;
  const unsigned long long mine = base + (unsigned long long)threadIdx.x * kItems;
  Tout v[kItems];
  Tout acc = 0;
#pragma unroll
  for (unsigned k = 0; k < kItems; ++k) {
    v[k] = mine + k < n ? (Tout)in[mine + k] : (Tout)0;
    acc += v[k];
  }
  __shared__ Tout s[kThreads];
  s[threadIdx.x] = acc;
  __syncthreads();
  for (unsigned d = 1; d < kThreads; d <<= 1) {  // inclusive Hillis-Steele over thread totals
    const Tout y = threadIdx.x >= d ? s[threadIdx.x - d] : (Tout)0;
    __syncthreads();
    s[threadIdx.x] += y;
    __syncthreads();
  }
  Tout run = s[threadIdx.x] - acc;
#pragma unroll
  for (unsigned k = 0; k < kItems; ++k) {
    if (mine + k < n) out[mine + k] = run;
    run += v[k];
  }
  if (threadIdx.x == kThreads - 1) sums[blockIdx.x] = s[kThreads - 1];
}

template <class T>
__global__ __launch_bounds__(kThreads) void scan_add_kernel(T* __restrict__ out,
                                                            unsigned long long n,
                                                            const T* __restrict__ base) {
  constexpr unsigned kItems = 8;
  const unsigned long long i0 = (unsigned long long)blockIdx.x * kThreads * kItems;
  const T add = base[blockIdx.x];
  for (unsigned k = threadIdx.x; k < kThreads * kItems; k += kThreads)
    if (i0 + k < n) out[i0 + k] += add;
}

unsigned grid_for(unsigned long long items, unsigned per_block, unsigned cap = 0x7fffffffu) {
  const unsigned long long g = (items + per_block - 1) / per_block;
  return (unsigned)(g == 0 ? 1 : (g > cap ? cap : g));
}

// Temporary device buffers of one parse. Blocks come from the context's cache of earlier parses
// (power-of-two sizes) or from hipMalloc, and go back to the cache when the parse ends (after
// its stream has drained); the cache keeps at most kScratchCache bytes.
constexpr size_t kScratchCache = (size_t)4 << 30;

struct Temps {
  Context* ctx;
  hipStream_t s;
  std::vector<std::pair<size_t, void*>> held;
  template <class T>
  hipError_t get(T** p, unsigned long long count) {
    size_t bytes = 256;
    while (bytes < (count ? count : 1) * sizeof(T)) bytes <<= 1;
    {
      std::lock_guard<std::mutex> lock(ctx->scratch_mu);
      auto it = ctx->scratch_free.find(bytes);
      if (it != ctx->scratch_free.end()) {
        *p = static_cast<T*>(it->second);
        ctx->scratch_free.erase(it);
        ctx->scratch_bytes -= bytes;
        held.emplace_back(bytes, *p);
        return hipSuccess;
      }
    }
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, bytes);
    if (e != hipSuccess) return e;
    *p = static_cast<T*>(q);
    held.emplace_back(bytes, q);
    return hipSuccess;
  }
  ~Temps() {
    (void)hipStreamSynchronize(s);
    std::lock_guard<std::mutex> lock(ctx->scratch_mu);
    for (auto& b : held) {
      if (ctx->scratch_bytes + b.first <= kScratchCache) {
        ctx->scratch_free.emplace(b.first, b.second);
        ctx->scratch_bytes += b.first;
      } else {
        (void)hipFree(b.second);
      }
    }
  }
};

// out[i] = in[0] + .. + in[i-1], and *total (device, may be null) = the sum of all.
template <class Tin, class Tout>
hipError_t exclusive_scan(const Tin* in, unsigned long long n, Tout* out, Tout* total,
                          hipStream_t s, Temps& t) {
  constexpr unsigned long long kPer = (unsigned long long)kThreads * 8;
  const unsigned long long blocks = n == 0 ? 1 : (n + kPer - 1) / kPer;
  Tout* sums = nullptr;
  hipError_t e = t.get(&sums, blocks + 1);
  if (e != hipSuccess) return e;
  scan_blocks_kernel<Tin, Tout><<<dim3((unsigned)blocks), dim3(kThreads), 0, s>>>(in, n, out, sums);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (blocks > 1) {
    Tout* sums_ex = nullptr;
    e = t.get(&sums_ex, blocks);
    if (e == hipSuccess) e = exclusive_scan<Tout, Tout>(sums, blocks, sums_ex, total, s, t);
    if (e == hipSuccess) {
      scan_add_kernel<Tout><<<dim3((unsigned)blocks), dim3(kThreads), 0, s>>>(out, n, sums_ex);
      e = hipGetLastError();
    }
    return e;
  }
  if (total) e = hipMemcpyAsync(total, sums, sizeof(Tout), hipMemcpyDeviceToDevice, s);
  return e;
}

#define FA_HIP(call)                                 \
  do {                                               \
    hipError_t e_ = (call);                          \
    if (e_ != hipSuccess) return hip_fail(e_, #call); \
  } while (0)

template <class T>
T read_scalar(const T* d, hipStream_t s, hipError_t* e) {
  T v{};
  *e = hipMemcpyAsync(&v, d, sizeof(T), hipMemcpyDeviceToHost, s);
  if (*e == hipSuccess) *e = hipStreamSynchronize(s);
  return v;
}

// The file image in device memory, its '\n' positions and its kept-byte chunk prefix.
struct Image {
  const uint8_t* b = nullptr;
  size_t N = 0;
  unsigned long long* nl = nullptr;  // '\n' positions
  unsigned long long n_nl = 0;
  unsigned long long* Kc = nullptr;  // kept bytes (33..127) before every 16-byte chunk
  unsigned long long M = 0;          // lines (a last line without '\n' counts)
};

int image_lines(Context* ctx, const uint8_t* src, size_t N, hipMemcpyKind stage, hipStream_t s,
                Temps& t, Image* im) {
  hipError_t e;
  const uint8_t* b = src;
  if (stage == hipMemcpyHostToDevice || stage == hipMemcpyDeviceToDevice) {
    uint8_t* copy = nullptr;
    FA_HIP(t.get(&copy, N));
    if (N && stage == hipMemcpyHostToDevice)
      FA_HIP(ctx->staging.upload(copy, reinterpret_cast<const char*>(src), N, N, 1, s));
    else if (N)
      FA_HIP(hipMemcpyAsync(copy, src, N, stage, s));
    b = copy;
  }
  const unsigned long long nblk = (N + kChunk - 1) / kChunk;
  const unsigned long long nchunks = (N + 15) / 16;
  uint32_t* blk = nullptr;
  uint8_t* kc = nullptr;
  unsigned long long *blk_base = nullptr, *n_nl_d = nullptr, *Kc = nullptr;
  FA_HIP(t.get(&blk, nblk));
  FA_HIP(t.get(&blk_base, nblk));
  FA_HIP(t.get(&n_nl_d, 1));
  FA_HIP(t.get(&kc, nchunks));
  FA_HIP(t.get(&Kc, nchunks + 1));
  nl_count_kernel<<<dim3((unsigned)nblk), dim3(kThreads), 0, s>>>(b, N, blk, kc);
  FA_HIP(hipGetLastError());
  FA_HIP((exclusive_scan<uint32_t, unsigned long long>(blk, nblk, blk_base, n_nl_d, s, t)));
  // kept bytes before every 16-byte chunk; Kc[nchunks] = all kept bytes
  FA_HIP((exclusive_scan<uint8_t, unsigned long long>(kc, nchunks, Kc, Kc + nchunks, s, t)));
  const unsigned long long n_nl = read_scalar(n_nl_d, s, &e);
  FA_HIP(e);
  unsigned long long* nl = nullptr;
  FA_HIP(t.get(&nl, n_nl));
  if (N > 0 && n_nl > 0) {
    nl_write_kernel<<<dim3((unsigned)nblk), dim3(kThreads), 0, s>>>(b, N, blk_base, nl);
    FA_HIP(hipGetLastError());
  }
  uint8_t last = '\n';
  FA_HIP(hipMemcpyAsync(&last, b + N - 1, 1, hipMemcpyDeviceToHost, s));
  FA_HIP(hipStreamSynchronize(s));
  im->b = b;
  im->N = N;
  im->nl = nl;
  im->n_nl = n_nl;
  im->Kc = Kc;
  im->M = n_nl + (last != '\n' ? 1 : 0);
  return RNJ_OK;
}

// The H names whose spans (device, [2H]) are given, copied to the host.
hipError_t gather_names(Temps& t, hipStream_t s, const uint8_t* b, unsigned H,
                        const unsigned long long* spans, std::vector<std::string>* names) {
  unsigned long long *nlen = nullptr, *noff = nullptr, *ntot = nullptr;
  uint8_t* blob = nullptr;
  hipError_t e = t.get(&nlen, H);
  if (e == hipSuccess) e = t.get(&noff, H);
  if (e == hipSuccess) e = t.get(&ntot, 1);
  if (e == hipSuccess) {
    name_len_kernel<<<dim3(grid_for(H, 256)), dim3(256), 0, s>>>(H, spans, nlen);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = exclusive_scan<unsigned long long, unsigned long long>(nlen, H, noff, ntot, s, t);
  unsigned long long total = 0;
  if (e == hipSuccess) total = read_scalar(ntot, s, &e);
  if (e == hipSuccess) e = t.get(&blob, total);
  if (e == hipSuccess && total > 0) {
    name_gather_kernel<<<dim3(H), dim3(64), 0, s>>>(b, H, spans, noff, blob);
    e = hipGetLastError();
  }
  std::vector<unsigned long long> hoff(H), hlen(H);
  std::vector<char> hblob(total);
  if (e == hipSuccess) e = hipMemcpyAsync(hoff.data(), noff, H * 8, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipMemcpyAsync(hlen.data(), nlen, H * 8, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess && total > 0)
    e = hipMemcpyAsync(hblob.data(), blob, total, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e == hipSuccess) {
    names->reserve(H);
    for (unsigned h = 0; h < H; ++h) names->emplace_back(hblob.data() + hoff[h], hlen[h]);
  }
  return e;
}

}  // namespace

int fasta_to_alignment(Context* ctx, const uint8_t* src, size_t N, hipMemcpyKind stage,
                       hipStream_t s, Alignment** out) {
  Temps t{ctx, s, {}};
  hipError_t e;
  // 1. '\n' positions
  if (N == 0) return fail(RNJ_E_INVALID, "empty FASTA input: no sequences");
  Image im;
  int rc0 = image_lines(ctx, src, N, stage, s, t, &im);
  if (rc0 != RNJ_OK) return rc0;
  const uint8_t* b = im.b;
  unsigned long long *nl = im.nl, *Kc = im.Kc;
  const unsigned long long n_nl = im.n_nl, M = im.M;
  if (M == 0) return fail(RNJ_E_INVALID, "empty FASTA input: no sequences");
  // 2. lines
  uint32_t *hcount = nullptr, *dcount = nullptr, *hbase = nullptr, *H_d = nullptr;
  unsigned long long *dspan = nullptr, *kds = nullptr, *off = nullptr, *total_d = nullptr;
  FA_HIP(t.get(&hcount, M));
  FA_HIP(t.get(&dcount, M));
  FA_HIP(t.get(&dspan, 2 * M));
  FA_HIP(t.get(&kds, M));
  line_kernel<<<dim3(grid_for(M, 256)), dim3(256), 0, s>>>(b, N, nl, n_nl, M, Kc, hcount, dspan,
                                                           dcount, kds);
  FA_HIP(hipGetLastError());

  // 3. scans
  FA_HIP(t.get(&hbase, M));
  FA_HIP(t.get(&H_d, 1));
  FA_HIP(t.get(&off, M));
  FA_HIP(t.get(&total_d, 1));
  FA_HIP((exclusive_scan<uint32_t, uint32_t>(hcount, M, hbase, H_d, s, t)));
  FA_HIP((exclusive_scan<uint32_t, unsigned long long>(dcount, M, off, total_d, s, t)));
  const unsigned H = read_scalar(H_d, s, &e);
  FA_HIP(e);
  const unsigned R = H + 1;  // record 0: data before the first header

  // 4. records
  unsigned long long *rec_len = nullptr, *rec_base = nullptr;
  uint32_t *stored = nullptr, *seq = nullptr, *nseq_d = nullptr;
  unsigned *first_d = nullptr, *bad_d = nullptr;
  FA_HIP(t.get(&rec_len, R));
  FA_HIP(t.get(&rec_base, R));
  FA_HIP(t.get(&stored, R));
  FA_HIP(t.get(&seq, R));
  FA_HIP(t.get(&nseq_d, 1));
  FA_HIP(t.get(&first_d, 1));
  FA_HIP(t.get(&bad_d, 1));
  FA_HIP(hipMemsetAsync(first_d, 0xff, sizeof(unsigned), s));
  FA_HIP(hipMemsetAsync(bad_d, 0xff, sizeof(unsigned), s));
  record_base_kernel<<<dim3(grid_for(M, 256)), dim3(256), 0, s>>>(M, hcount, hbase, off, rec_base);
  FA_HIP(hipGetLastError());
  record_len_kernel<<<dim3(grid_for(R, 256)), dim3(256), 0, s>>>(R, rec_base, total_d, rec_len);
  FA_HIP(hipGetLastError());
  stored_flag_kernel<<<dim3(grid_for(R, 256)), dim3(256), 0, s>>>(R, rec_len, stored,
                                                                                first_d);
  FA_HIP(hipGetLastError());
  FA_HIP((exclusive_scan<uint32_t, uint32_t>(stored, R, seq, nseq_d, s, t)));
  length_check_kernel<<<dim3(grid_for(R, 256)), dim3(256), 0, s>>>(R, rec_len, seq,
                                                                                 first_d, bad_d);
  FA_HIP(hipGetLastError());
  const unsigned nseq = read_scalar(nseq_d, s, &e);
  FA_HIP(e);
  const unsigned first = read_scalar(first_d, s, &e);
  FA_HIP(e);
  const unsigned bad = read_scalar(bad_d, s, &e);
  FA_HIP(e);
  if (nseq < 2)
    return fail(RNJ_E_INVALID, "found " + std::to_string(nseq) +
                                   " sequences in the FASTA input; at least 2 are needed");
  unsigned long long L = 0;
  FA_HIP(hipMemcpyAsync(&L, rec_len + first, sizeof(L), hipMemcpyDeviceToHost, s));
  FA_HIP(hipStreamSynchronize(s));
  if (bad != 0xffffffffu) {
    // the reference reports the first offending sequence with its length
    std::vector<unsigned long long> lens(R);
    std::vector<uint32_t> seqs(R);
    FA_HIP(hipMemcpyAsync(lens.data(), rec_len, R * sizeof(unsigned long long),
                          hipMemcpyDeviceToHost, s));
    FA_HIP(hipMemcpyAsync(seqs.data(), seq, R * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    FA_HIP(hipStreamSynchronize(s));
    unsigned long long blen = 0;
    for (unsigned r = 0; r < R; ++r)
      if (lens[r] > 0 && seqs[r] == bad) blen = lens[r];
    return fail(RNJ_E_INVALID, "sequence " + std::to_string(bad) + " has length " +
                                   std::to_string(blen) + " while the preceding sequences have length " +
                                   std::to_string(L));
  }
  if (L >= (1ull << 32) - 64) return fail(RNJ_E_INVALID, "sequences longer than 2^32 sites");

  // 5. character matrix
  uint8_t* chars = nullptr;
  FA_HIP(t.get(&chars, (unsigned long long)nseq * L));
  unsigned long long* dbase = nullptr;
  FA_HIP(t.get(&dbase, M));
  line_dest_kernel<<<dim3(grid_for(M, 256)), dim3(256), 0, s>>>(M, hcount, hbase, dspan, kds, off,
                                                                rec_base, seq, L, dbase);
  FA_HIP(hipGetLastError());
  const unsigned long long scatter_waves = (N + kScatterBytes - 1) / kScatterBytes;
  scatter_kernel<<<dim3(grid_for(scatter_waves, 4)), dim3(256), 0, s>>>(b, N, nl, n_nl, Kc, dspan,
                                                                        dbase, chars);
  FA_HIP(hipGetLastError());
  std::vector<char> row0(L);
  FA_HIP(hipMemcpyAsync(row0.data(), chars, L, hipMemcpyDeviceToHost, s));
  FA_HIP(hipStreamSynchronize(s));
  const int kind = alignment_kind(ctx, row0.data(), (unsigned)L);

  // 6. pack + names
  Alignment* a = nullptr;
  int rc = pack_alignment(ctx, chars, L, nseq, (unsigned)L, kind, s, &a);
  if (rc != RNJ_OK) return rc;
  e = hipSuccess;
  if (H > 0) {
    unsigned long long* spans = nullptr;
    e = t.get(&spans, 2ull * H);
    if (e == hipSuccess) {
      header_kernel<<<dim3(grid_for(M, 256)), dim3(256), 0, s>>>(b, N, nl, n_nl, M, hcount,
                                                                               hbase, spans);
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = gather_names(t, s, b, H, spans, &a->names);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    rnj_alignment_free(static_cast<rnj_alignment*>(a));
    return hip_fail(e, "FASTA parse");
  }
  *out = a;
  return RNJ_OK;
}

int stockholm_to_alignment(Context* ctx, const uint8_t* src, size_t N, hipMemcpyKind stage,
                           hipStream_t s, Alignment** out) {
  Temps t{ctx, s, {}};
  hipError_t e;
  const char* eof_msg =
      "the Stockholm input ends before the \"//\" line that closes the alignment (the reference "
      "exits: stream ended unexpectedly)";
  if (N == 0) return fail(RNJ_E_INVALID, eof_msg);
  Image im;
  int rc = image_lines(ctx, src, N, stage, s, t, &im);
  if (rc != RNJ_OK) return rc;
  const unsigned long long M = im.M;
  // 1. the state at every line's start
  uint8_t *map = nullptr, *state = nullptr, *final_d = nullptr;
  FA_HIP(t.get(&map, M));
  FA_HIP(t.get(&state, M));
  FA_HIP(t.get(&final_d, 1));
  sth_map_kernel<<<dim3(grid_for(M, 256)), dim3(256), 0, s>>>(im.b, N, im.nl, im.n_nl, M, map);
  FA_HIP(hipGetLastError());
  sth_state_kernel<<<dim3(1), dim3(256), 0, s>>>(map, M, state, final_d);
  FA_HIP(hipGetLastError());
  // 2. what every line holds
  uint32_t *named = nullptr, *dcount = nullptr, *nbase = nullptr, *R_d = nullptr;
  unsigned long long *nspan = nullptr, *dspan = nullptr, *kds = nullptr, *err_d = nullptr;
  FA_HIP(t.get(&named, M));
  FA_HIP(t.get(&dcount, M));
  FA_HIP(t.get(&nspan, 2 * M));
  FA_HIP(t.get(&dspan, 2 * M));
  FA_HIP(t.get(&kds, M));
  FA_HIP(t.get(&err_d, 1));
  FA_HIP(hipMemsetAsync(err_d, 0xff, sizeof(unsigned long long), s));
  sth_line_kernel<<<dim3(grid_for(M, 256)), dim3(256), 0, s>>>(im.b, N, im.nl, im.n_nl, M, im.Kc,
                                                               state, named, nspan, dspan, dcount,
                                                               kds, err_d);
  FA_HIP(hipGetLastError());
  FA_HIP(t.get(&nbase, M));
  FA_HIP(t.get(&R_d, 1));
  FA_HIP((exclusive_scan<uint32_t, uint32_t>(named, M, nbase, R_d, s, t)));
  const unsigned R = read_scalar(R_d, s, &e);
  FA_HIP(e);
  const uint8_t final_state = read_scalar(final_d, s, &e);
  FA_HIP(e);
  const unsigned long long err = read_scalar(err_d, s, &e);
  FA_HIP(e);
  // 3. sequences: length and data line of each, the first whose length differs from the first's
  unsigned long long *rec_len = nullptr, *rec_line = nullptr, *bad_d = nullptr;
  FA_HIP(t.get(&rec_len, R));
  FA_HIP(t.get(&rec_line, R));
  FA_HIP(t.get(&bad_d, 1));
  FA_HIP(hipMemsetAsync(bad_d, 0xff, sizeof(unsigned long long), s));
  if (R > 0) {
    FA_HIP(hipMemsetAsync(rec_line, 0xff, R * sizeof(unsigned long long), s));
    sth_record_kernel<<<dim3(grid_for(M, 256)), dim3(256), 0, s>>>(M, named, nbase, dspan, dcount,
                                                                   rec_len, rec_line);
    FA_HIP(hipGetLastError());
    sth_length_check_kernel<<<dim3(grid_for(R, 256)), dim3(256), 0, s>>>(R, rec_len, rec_line,
                                                                         bad_d);
    FA_HIP(hipGetLastError());
  }
  const unsigned long long bad = read_scalar(bad_d, s, &e);
  FA_HIP(e);
  // The reference stops at the first error in file order: a line that errs (a name with an
  // empty line, the input ending), a sequence of another length (detected at its '\n'), or the
  // end of the input without "//".
  const unsigned long long err_line = err != ~0ull ? err >> 2 : ~0ull;
  const unsigned long long bad_line = bad != ~0ull ? bad >> 32 : ~0ull;
  const unsigned long long end_line = (final_state == kSthS || final_state == kSthP) ? M : ~0ull;
  if (bad_line < err_line && bad_line < end_line) {
    const unsigned r = (unsigned)(bad & 0xffffffffu);
    unsigned long long blen = 0, len0 = 0;
    FA_HIP(hipMemcpyAsync(&blen, rec_len + r, sizeof(blen), hipMemcpyDeviceToHost, s));
    FA_HIP(hipMemcpyAsync(&len0, rec_len, sizeof(len0), hipMemcpyDeviceToHost, s));
    FA_HIP(hipStreamSynchronize(s));
    return fail(RNJ_E_INVALID, "sequence " + std::to_string(r) + " has length " +
                                   std::to_string(blen) + " while the preceding sequences have length " +
                                   std::to_string(len0));
  }
  if (err_line != ~0ull && err_line <= end_line) {
    if ((err & 3u) == kSthEof) return fail(RNJ_E_INVALID, eof_msg);
    // the name of the sequence whose data is missing: the last name on or before the line
    uint32_t nm = 0, nb = 0;
    FA_HIP(hipMemcpyAsync(&nm, named + err_line, sizeof(nm), hipMemcpyDeviceToHost, s));
    FA_HIP(hipMemcpyAsync(&nb, nbase + err_line, sizeof(nb), hipMemcpyDeviceToHost, s));
    FA_HIP(hipStreamSynchronize(s));
    const unsigned long long src_line = nm ? err_line : err_line - 1;  // a P line's name is above
    unsigned long long sp[2] = {0, 0};
    FA_HIP(hipMemcpyAsync(sp, nspan + 2 * src_line, sizeof(sp), hipMemcpyDeviceToHost, s));
    FA_HIP(hipStreamSynchronize(s));
    std::string name(sp[1] - sp[0], '\0');
    if (!name.empty())
      FA_HIP(hipMemcpyAsync(&name[0], im.b + sp[0], name.size(), hipMemcpyDeviceToHost, s));
    FA_HIP(hipStreamSynchronize(s));
    return fail(RNJ_E_INVALID, "sequence " + std::to_string(nb + nm - 1) + " (\"" + name +
                                   "\") has zero length");
  }
  if (end_line != ~0ull) return fail(RNJ_E_INVALID, eof_msg);
  if (R < 2)
    return fail(RNJ_E_INVALID, "found " + std::to_string(R) +
                                   " sequences in the Stockholm input; at least 2 are needed");
  unsigned long long L = 0;
  FA_HIP(hipMemcpyAsync(&L, rec_len, sizeof(L), hipMemcpyDeviceToHost, s));
  FA_HIP(hipStreamSynchronize(s));
  if (L == 0) return fail(RNJ_E_INVALID, "the Stockholm sequences are empty");
  if (L >= (1ull << 32) - 64) return fail(RNJ_E_INVALID, "sequences longer than 2^32 sites");
  // 4. character matrix
  uint8_t* chars = nullptr;
  unsigned long long* dbase = nullptr;
  FA_HIP(t.get(&chars, (unsigned long long)R * L));
  FA_HIP(t.get(&dbase, M));
  sth_dest_kernel<<<dim3(grid_for(M, 256)), dim3(256), 0, s>>>(M, named, nbase, dspan, kds, L,
                                                               dbase);
  FA_HIP(hipGetLastError());
  const unsigned long long scatter_waves = (N + kScatterBytes - 1) / kScatterBytes;
  scatter_kernel<<<dim3(grid_for(scatter_waves, 4)), dim3(256), 0, s>>>(im.b, N, im.nl, im.n_nl,
                                                                        im.Kc, dspan, dbase, chars);
  FA_HIP(hipGetLastError());
  // 5. type: the reference discovers it on the first sequence (dataloaderStockholm.hpp:53-76); on
  // the -f path it does so after resolveChar, while the type is still unknown, i.e. with the
  // protein rule that turns - . X x z Z b B J j ? into '-' (:238-249, :261-274)
  std::vector<char> row0(L);
  FA_HIP(hipMemcpyAsync(row0.data(), chars, L, hipMemcpyDeviceToHost, s));
  FA_HIP(hipStreamSynchronize(s));
  if (ctx->char_rules == RNJ_CHARS_NAIVE)
    for (char& c : row0)
      if (std::strchr("-.XxzZbBJj?", c) != nullptr && c != '\0') c = '-';
  const int kind = alignment_kind(ctx, row0.data(), (unsigned)L);
  // 6. pack + names
  Alignment* a = nullptr;
  rc = pack_alignment(ctx, chars, L, R, (unsigned)L, kind, s, &a);
  if (rc != RNJ_OK) return rc;
  unsigned long long* spans = nullptr;
  e = t.get(&spans, 2ull * R);
  if (e == hipSuccess) {
    sth_name_kernel<<<dim3(grid_for(M, 256)), dim3(256), 0, s>>>(M, named, nbase, nspan, spans);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = gather_names(t, s, im.b, R, spans, &a->names);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    rnj_alignment_free(static_cast<rnj_alignment*>(a));
    return hip_fail(e, "Stockholm parse");
  }
  *out = a;
  return RNJ_OK;
}

}  // namespace rnj
