/* CPU restatement of rapidNJ's DNA pairwise-distance path -- TEST INFRASTRUCTURE ONLY.
 * Loaded only by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg, as the
 * checker. The declarations in rnj_oracle.h document each function and cite the reference. */
#define _POSIX_C_SOURCE 200809L
#include "rnj_oracle.h"

#include <emmintrin.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* dataloaderFasta.hpp:222 -- DNA: L/64 + 6 words of 128 bits. */
unsigned rnj_oracle_bitstrings_count(unsigned L) { return L / 64 + 6; }

/* dataloaderFasta.hpp:54-76. The reference stores chars as (signed) char, so bytes >= 128 are
 * negative and fall under `c < 65`. */
int rnj_oracle_is_dna(const char* seq, unsigned L) {
  for (unsigned i = 0; i < L; i++) {
    signed char c = (signed char)seq[i];
    if (c < 65) continue;
    switch (c) {
      case 'A': case 'a': case 'C': case 'c': case 'G': case 'g': case 'T': case 't':
      case 'U': case 'u': case 'N': case 'n':
        continue;
      default:
        return 0;
    }
  }
  return 1;
}

/* 2-bit code of a base, or -1 when the site is not counted (stdinclude.h:25,
 * dataloaderFasta.hpp:156-180: only A/C/G/T of either case set the gap-filter slot). */
static int base_code(char c) {
  switch (c) {
    case 'A': case 'a': return 0;
    case 'G': case 'g': return 1;
    case 'T': case 't': return 2;
    case 'C': case 'c': return 3;
    default: return -1;
  }
}

void rnj_oracle_pack(const char* aln, unsigned n, unsigned L, unsigned bsc, uint32_t* bits,
                     uint32_t* gaps) {
  const unsigned words = bsc * 4;
  for (unsigned t = 0; t < n; t++) {
    uint32_t* b = bits + (size_t)t * words;
    uint32_t* g = gaps + (size_t)t * words;
    memset(b, 0, words * sizeof(uint32_t));
    memset(g, 0, words * sizeof(uint32_t));
    const char* s = aln + (size_t)t * L;
    for (unsigned k = 0; k < L; k++) { /* sites >= L are '-' pads: zero code, zero filter */
      int c = base_code(s[k]);
      if (c < 0) continue;
      b[k / 16] |= (uint32_t)c << (2 * (k % 16));
      g[k / 16] |= 1u << (2 * (k % 16));
    }
  }
}

void rnj_oracle_count_chars(const char* a, const char* b, unsigned L, uint64_t out[3]) {
  uint64_t ts = 0, tv = 0, valid = 0;
  for (unsigned k = 0; k < L; k++) {
    char x = a[k], y = b[k];
    if (x >= 'a' && x <= 'z') x = (char)(x - 32);
    if (y >= 'a' && y <= 'z') y = (char)(y - 32);
    int xv = (x == 'A' || x == 'C' || x == 'G' || x == 'T');
    int yv = (y == 'A' || y == 'C' || y == 'G' || y == 'T');
    if (!xv || !yv) continue;
    valid++;
    if (x == y) continue;
    int x_purine = (x == 'A' || x == 'G');
    int y_purine = (y == 'A' || y == 'G');
    if (x_purine != y_purine) tv++; else ts++;
  }
  out[0] = ts; out[1] = tv; out[2] = valid;
}

/* dataloaderFasta.hpp:184-191 (DNA branch of resolveChar). */
static char resolve_dna(char c) {
  switch (c) {
    case 'a': case 'A': case 'c': case 'C': case 'g': case 'G': case 't': case 'T':
    case 'u': case 'U':
      return c;
    default:
      return '-';
  }
}

void rnj_oracle_count_naive(const char* a, const char* b, unsigned L, uint64_t out[3]) {
  uint64_t ts = 0, tv = 0, len = 0;
  for (unsigned k = 0; k < L; k++) {
    char x = resolve_dna(a[k]), y = resolve_dna(b[k]);
    if (x == '-' || y == '-') continue;
    if (x != y) {
      /* only uppercase letters enter the ts/tv branches (simpleDistanceCalculator.cpp:66-91) */
      if (x == 'A') { if (y == 'G') ts++; else tv++; }
      else if (x == 'C') { if (y == 'T') ts++; else tv++; }
      else if (x == 'G') { if (y == 'A') ts++; else tv++; }
      else if (x == 'T') { if (y == 'C') ts++; else tv++; }
    }
    len++;
  }
  out[0] = ts; out[1] = tv; out[2] = len;
}

void rnj_oracle_count_words(const uint32_t* bi, const uint32_t* gi, const uint32_t* bj,
                            const uint32_t* gj, unsigned nwords, uint64_t out[3]) {
  const uint32_t lo = 0x55555555u;
  uint64_t ts = 0, tv = 0, valid = 0;
  for (unsigned w = 0; w < nwords; w++) {
    uint32_t r = bi[w] ^ bj[w];
    uint32_t g = gi[w] & gj[w];
    uint32_t hi_diff = (r >> 1) & lo;          /* high bits differ: transversion */
    uint32_t lo_only = (r & lo) & ~hi_diff;    /* only the low bit differs: transition */
    tv += (uint64_t)__builtin_popcount(hi_diff & g);
    ts += (uint64_t)__builtin_popcount(lo_only & g);
    valid += (uint64_t)__builtin_popcount(g);
  }
  out[0] = ts; out[1] = tv; out[2] = valid;
}

/* ---- SSE2 SWAR counter -------------------------------------------------------------------
 * Same schedule as the reference ladder: per group of three 128-bit words the 1-bit indicators
 * (in the low bit of each 2-bit slot) add into 2-bit lanes (<= 3); two groups widen to 4-bit
 * lanes (<= 12); up to 8 such steps accumulate in 8-bit lanes (<= 192); up to 128 of those in
 * 16-bit lanes (<= 49152); then 32-bit lanes and finally 64-bit totals. */
typedef struct { __m128i ts, tv, gf; } tri128;

static inline __m128i widen(__m128i v, __m128i m, int sh) {
  return _mm_add_epi32(_mm_and_si128(v, m), _mm_and_si128(_mm_srli_epi32(v, sh), m));
}

static inline void group3(const __m128i* a, const __m128i* ga, const __m128i* b,
                          const __m128i* gb, tri128* out) {
  const __m128i m1 = _mm_set1_epi32(0x55555555);
  __m128i ts = _mm_setzero_si128(), tv = _mm_setzero_si128(), gf = _mm_setzero_si128();
  for (int k = 0; k < 3; k++) {
    __m128i r = _mm_xor_si128(_mm_load_si128(a + k), _mm_load_si128(b + k));
    __m128i g = _mm_and_si128(_mm_load_si128(ga + k), _mm_load_si128(gb + k));
    __m128i hv = _mm_and_si128(_mm_srli_epi32(r, 1), m1);
    __m128i lv = _mm_andnot_si128(hv, _mm_and_si128(r, m1));
    tv = _mm_add_epi32(tv, _mm_and_si128(hv, g));
    ts = _mm_add_epi32(ts, _mm_and_si128(lv, g));
    gf = _mm_add_epi32(gf, g);
  }
  out->ts = ts; out->tv = tv; out->gf = gf;
}

void rnj_oracle_count_swar(const uint32_t* bi, const uint32_t* gi, const uint32_t* bj,
                           const uint32_t* gj, unsigned bsc, uint64_t out[3]) {
  const __m128i m2 = _mm_set1_epi32(0x33333333), m4 = _mm_set1_epi32(0x0f0f0f0f);
  const __m128i m8 = _mm_set1_epi32(0x00ff00ff), m16 = _mm_set1_epi32(0x0000ffff);
  const __m128i* A = (const __m128i*)bi;
  const __m128i* GA = (const __m128i*)gi;
  const __m128i* B = (const __m128i*)bj;
  const __m128i* GB = (const __m128i*)gj;
  unsigned blocks = bsc / 6; /* 6 words per 4-bit step */
  uint64_t tot[3] = {0, 0, 0};
  unsigned blk = 0;
  while (blk < blocks) {
    __m128i s32[3] = {_mm_setzero_si128(), _mm_setzero_si128(), _mm_setzero_si128()};
    for (unsigned r16 = 0; r16 < 32768u * 128u && blk < blocks; r16 += 128) {
      __m128i s16[3] = {_mm_setzero_si128(), _mm_setzero_si128(), _mm_setzero_si128()};
      for (unsigned r8 = 0; r8 < 128 && blk < blocks; r8++) {
        __m128i s8[3] = {_mm_setzero_si128(), _mm_setzero_si128(), _mm_setzero_si128()};
        for (unsigned r4 = 0; r4 < 8 && blk < blocks; r4++, blk++) {
          size_t w = (size_t)blk * 6;
          tri128 x, y;
          group3(A + w, GA + w, B + w, GB + w, &x);
          group3(A + w + 3, GA + w + 3, B + w + 3, GB + w + 3, &y);
          __m128i c4_ts = _mm_add_epi32(widen(x.ts, m2, 2), widen(y.ts, m2, 2));
          __m128i c4_tv = _mm_add_epi32(widen(x.tv, m2, 2), widen(y.tv, m2, 2));
          __m128i c4_gf = _mm_add_epi32(widen(x.gf, m2, 2), widen(y.gf, m2, 2));
          s8[0] = _mm_add_epi32(s8[0], widen(c4_ts, m4, 4));
          s8[1] = _mm_add_epi32(s8[1], widen(c4_tv, m4, 4));
          s8[2] = _mm_add_epi32(s8[2], widen(c4_gf, m4, 4));
        }
        for (int q = 0; q < 3; q++) s16[q] = _mm_add_epi32(s16[q], widen(s8[q], m8, 8));
      }
      for (int q = 0; q < 3; q++) s32[q] = _mm_add_epi32(s32[q], widen(s16[q], m16, 16));
    }
    for (int q = 0; q < 3; q++) {
      uint32_t lane[4];
      _mm_storeu_si128((__m128i*)lane, s32[q]);
      tot[q] += (uint64_t)lane[0] + lane[1] + lane[2] + lane[3];
    }
  }
  out[0] = tot[0]; out[1] = tot[1]; out[2] = tot[2];
}

/* KimuraDistance.cpp:339-348 (DNA branch of distThread). */
float rnj_oracle_k2p(uint64_t ts, uint64_t tv, uint64_t valid) {
  float d = -1.0f;
  if (valid != 0) {
    float P = (float)((double)ts / (double)valid);
    float Q = (float)((double)tv / (double)valid);
    float t1 = (float)(1.0 - 2.0 * (double)P - (double)Q);
    float t2 = (float)(1.0 - 2.0 * (double)Q);
    if (!(t1 <= 0 || t2 <= 0))
      d = (float)(-0.5 * log(1.0 - 2.0 * (double)P - (double)Q) - 0.25 * log(1.0 - 2.0 * (double)Q));
  }
  return d;
}

/* JCdistance.cpp:155-183 with alpha = 4 (DNA). */
float rnj_oracle_jc(uint64_t ts, uint64_t tv, uint64_t valid) {
  const double alpha = 4.0;
  long total = (long)(ts + tv);
  float p;
  if (valid == 0) p = 0;
  else p = (float)total / (float)valid;
  float d;
  if ((double)p / ((alpha - 1) / alpha) >= 1.0) d = -1;
  else d = (float)(-((alpha - 1) / alpha) * log(1.0 - (double)p / ((alpha - 1) / alpha)));
  return d;
}

/* ---- threaded drivers ----------------------------------------------------------------- */
static inline size_t tri_index(size_t i, size_t j, size_t n) {
  return i * n - i * (i + 1) / 2 + (j - i - 1);
}

typedef struct {
  const uint32_t *bits, *gaps;
  unsigned n, bsc;
  int tid, threads, model;
  uint32_t *ts, *tv, *valid; /* counts mode */
  float* tri;                /* distance mode */
  float maxd;
} job_t;

static void* count_rows(void* p) {
  job_t* J = (job_t*)p;
  const size_t words = (size_t)J->bsc * 4;
  uint64_t c[3];
  float maxd = 0.0f;
  for (unsigned i = (unsigned)J->tid; i < J->n; i += (unsigned)J->threads) {
    for (unsigned j = i + 1; j < J->n; j++) {
      rnj_oracle_count_swar(J->bits + i * words, J->gaps + i * words, J->bits + j * words,
                            J->gaps + j * words, J->bsc, c);
      size_t k = tri_index(i, j, J->n);
      if (J->tri) {
        float d = J->model == RNJ_ORACLE_K2P ? rnj_oracle_k2p(c[0], c[1], c[2])
                                             : rnj_oracle_jc(c[0], c[1], c[2]);
        if (d > maxd) maxd = d;
        J->tri[k] = d;
      } else {
        J->ts[k] = (uint32_t)c[0];
        J->tv[k] = (uint32_t)c[1];
        J->valid[k] = (uint32_t)c[2];
      }
    }
  }
  J->maxd = maxd;
  return NULL;
}

static int run_threads(job_t* proto, int threads, float* maxd) {
  if (threads < 1) threads = 1;
  pthread_t* th = (pthread_t*)calloc((size_t)threads, sizeof(pthread_t));
  job_t* jobs = (job_t*)calloc((size_t)threads, sizeof(job_t));
  if (!th || !jobs) { free(th); free(jobs); return -1; }
  for (int t = 0; t < threads; t++) {
    jobs[t] = *proto;
    jobs[t].tid = t;
    jobs[t].threads = threads;
    if (pthread_create(&th[t], NULL, count_rows, &jobs[t]) != 0) {
      for (int u = 0; u < t; u++) pthread_join(th[u], NULL);
      free(th); free(jobs);
      return -1;
    }
  }
  float m = 0.0f;
  for (int t = 0; t < threads; t++) {
    pthread_join(th[t], NULL);
    if (m < jobs[t].maxd) m = jobs[t].maxd;
  }
  if (maxd) *maxd = m;
  free(th); free(jobs);
  return 0;
}

int rnj_oracle_counts_triangle(const uint32_t* bits, const uint32_t* gaps, unsigned n,
                               unsigned bsc, int threads, uint32_t* ts, uint32_t* tv,
                               uint32_t* valid) {
  job_t J;
  memset(&J, 0, sizeof J);
  J.bits = bits; J.gaps = gaps; J.n = n; J.bsc = bsc;
  J.ts = ts; J.tv = tv; J.valid = valid;
  return run_threads(&J, threads, NULL);
}

/* KimuraDistance.cpp:366-395 / JCdistance.cpp:201-230 on an upper triangle. */
static void post_process(const float* tri, unsigned n, float maxd, float* full) {
  for (size_t i = 0; i < n; i++) {
    full[i * n + i] = 0.0f;
    for (size_t j = i + 1; j < n; j++) {
      float d = tri[tri_index(i, j, n)];
      if (d == -1) d = maxd * 2;
      full[i * n + j] = d;
      full[j * n + i] = d;
    }
  }
}

int rnj_oracle_distance_matrix(const uint32_t* bits, const uint32_t* gaps, unsigned n,
                               unsigned bsc, int model, int threads, float* tri, float* full,
                               float* max_out) {
  size_t npairs = (size_t)n * (n - 1) / 2;
  float* t = tri;
  if (!t) {
    t = (float*)malloc((npairs ? npairs : 1) * sizeof(float));
    if (!t) return -1;
  }
  job_t J;
  memset(&J, 0, sizeof J);
  J.bits = bits; J.gaps = gaps; J.n = n; J.bsc = bsc; J.model = model; J.tri = t;
  float maxd = 0.0f;
  int rc = run_threads(&J, threads, &maxd);
  if (rc == 0 && full) post_process(t, n, maxd, full);
  if (max_out) *max_out = maxd;
  if (!tri) free(t);
  return rc;
}

int rnj_oracle_correct_triangle(const uint32_t* ts, const uint32_t* tv, const uint32_t* valid,
                                unsigned n, int model, float* tri, float* full, float* max_out) {
  size_t npairs = (size_t)n * (n - 1) / 2;
  float* t = tri;
  if (!t) {
    t = (float*)malloc((npairs ? npairs : 1) * sizeof(float));
    if (!t) return -1;
  }
  float maxd = 0.0f;
  for (size_t k = 0; k < npairs; k++) {
    float d = model == RNJ_ORACLE_K2P ? rnj_oracle_k2p(ts[k], tv[k], valid[k])
                                      : rnj_oracle_jc(ts[k], tv[k], valid[k]);
    if (d > maxd) maxd = d;
    t[k] = d;
  }
  if (full) post_process(t, n, maxd, full);
  if (max_out) *max_out = maxd;
  if (!tri) free(t);
  return 0;
}

/* ---- FASTA reader (dataloaderFasta.hpp:78-133, 201-276) ----------------------------------- */
typedef struct {
  const unsigned char* d;
  size_t len, pos;
  int eof;
} fasta_reader;

/* getNextChar (dataloaderFasta.hpp:118-126): the next byte as a signed char; past the end,
 * 10 ('\n'), and the main loop stops (fillBuffer leaves buf NULL). */
static int fasta_get(fasta_reader* r) {
  if (r->pos < r->len) return (int)(signed char)r->d[r->pos++];
  r->eof = 1;
  return 10;
}

int rnj_oracle_parse_fasta(const unsigned char* data, size_t len, unsigned* n, unsigned* L,
                           unsigned* H, size_t* name_bytes, unsigned* bad, unsigned* bad_len,
                           unsigned char* rows, unsigned char* names, unsigned* name_len) {
  fasta_reader r = {data, len, 0, 0};
  unsigned nseq = 0, nh = 0, first_len = 0, have_first = 0;
  size_t nb = 0, cur = 0;
  const unsigned row_len = rows ? *L : 0;
#define FASTA_STORE()                                   \
  do {                                                  \
    if (cur) {                                          \
      if (!have_first) {                                \
        first_len = (unsigned)cur;                      \
        have_first = 1;                                 \
      } else if (cur != first_len) {                    \
        *bad = nseq;                                    \
        *bad_len = (unsigned)cur;                       \
        return 1;                                       \
      }                                                 \
      nseq++;                                           \
      cur = 0;                                          \
    }                                                   \
  } while (0)
  while (!r.eof) {                                      /* parseData */
    int c = fasta_get(&r);
    if (c == ';') {                                     /* parseComment */
      while (c != 10) c = fasta_get(&r);
    } else if (c == '>') {
      FASTA_STORE();
      c = fasta_get(&r);
      size_t s = 0;                                     /* parseSequenceName */
      while (c > 31) {
        if (names) names[nb + s] = (unsigned char)c;
        s++;
        c = fasta_get(&r);
      }
      if (name_len) name_len[nh] = (unsigned)s;
      nh++;
      nb += s;
    } else {                                            /* parseSequenceLine */
      while (c != 10) {
        if (c > 32) {
          if (rows && cur < row_len) rows[(size_t)nseq * row_len + cur] = (unsigned char)c;
          cur++;
        }
        c = fasta_get(&r);
      }
    }
  }
  FASTA_STORE();
#undef FASTA_STORE
  *n = nseq;
  if (!rows) *L = first_len;
  *H = nh;
  *name_bytes = nb;
  return nseq < 2 ? 2 : 0;
}

/* ---- Stockholm reader (dataloaderStockholm.hpp:83-111, 118-132, 183-236, 294-314) --------- */
typedef struct {
  const unsigned char* d;
  size_t len, pos;
} sth_reader;

/* getNextChar (dataloaderStockholm.hpp:118-123): the next byte as a signed char, or -1000 past
 * the end, where the reference's fillBuffer exits ("Stream ended unexpectedly", :109-117). */
static int sth_get(sth_reader* r) {
  if (r->pos < r->len) return (int)(signed char)r->d[r->pos++];
  return -1000;
}

int rnj_oracle_parse_stockholm(const unsigned char* data, size_t len, unsigned* n, unsigned* L,
                               size_t* name_bytes, unsigned* bad, unsigned* bad_len,
                               unsigned char* rows, unsigned char* names, unsigned* name_len) {
  sth_reader r = {data, len, 0};
  unsigned nseq = 0, first_len = 0;
  size_t nb = 0;
  const unsigned row_len = rows ? *L : 0;
#define STH_GET()                        do {                                     c = sth_get(&r);                       if (c == -1000) return 4;            } while (0)
  int c;
  STH_GET();
  for (;;) {                                          /* parseData */
    while (c < 32) STH_GET();
    if (c == '#') {                                   /* parseMarkupLine */
      while (c != 10) STH_GET();
    } else {
      size_t s = 0;
      if (c == '/') {
        STH_GET();
        if (c == '/') break;                          /* "//": the end of the alignment */
        if (names) names[nb] = '/';                   /* tempBuffer = "/" */
        s = 1;
      }
      while (c > 32) {                                /* parseSequenceName */
        if (names) names[nb + s] = (unsigned char)c;
        s++;
        STH_GET();
      }
      if (name_len) name_len[nseq] = (unsigned)s;
      nb += s;
      while (c < 32) {
        STH_GET();
        if (c == 10) {                                /* "has zero length" */
          *bad = nseq;
          return 3;
        }
      }
      size_t cur = 0;                                 /* parseSequenceLineFastDist */
      while (c != 10) {
        if (c > 32) {
          if (rows && cur < row_len) rows[(size_t)nseq * row_len + cur] = (unsigned char)c;
          cur++;
        }
        STH_GET();
      }
      if (nseq == 0) {
        first_len = (unsigned)cur;
      } else if (cur != first_len) {
        *bad = nseq;
        *bad_len = (unsigned)cur;
        return 1;
      }
      nseq++;
    }
    STH_GET();
  }
#undef STH_GET
  *n = nseq;
  if (!rows) *L = first_len;
  *name_bytes = nb;
  return nseq < 2 ? 2 : 0;
}

/* ---- protein (dataloaderFasta.hpp:184-199, bitDistanceProtein.cpp:31-119, KimuraDistance.cpp
 * :30-49,322-338, JCdistance.cpp:155-183) --------------------------------------------------- */
char rnj_oracle_protein_resolve(char c) {
  switch (c) {
    case '-': case '.': case 'X': case 'x': case 'z': case 'Z': case 'b': case 'B':
    case 'J': case 'j': case '?':
      return '-';
    default:
      return c;
  }
}

void rnj_oracle_count_protein(const char* a, const char* b, unsigned L, uint64_t out[3]) {
  const uint64_t walked = 128ull * (L / 128 + 1);     /* (bitStringsCount / 8) * 8 * 16 chars */
  const uint64_t padded = ((uint64_t)L + 127) & ~127ull;
  uint64_t dist = 0, gaps = walked - L;                /* padding is '-' */
  for (unsigned k = 0; k < L; k++) {
    const char x = rnj_oracle_protein_resolve(a[k]), y = rnj_oracle_protein_resolve(b[k]);
    if (x == '-' || y == '-') gaps++;
    else if (x != y) dist++;
  }
  out[0] = dist;
  out[1] = 0;
  out[2] = padded - gaps;
}

/* simpleDistanceCalculator::calculateHammingDistance (simpleDistanceCalculator.cpp:40-55) on the
 * loader's resolved characters (storeSequence with fastdist off, dataloaderFasta.hpp:268-271):
 * a position counts when neither resolved character is '.' or '-', and is a difference when they
 * differ; no padding is walked. */
void rnj_oracle_count_protein_naive(const char* a, const char* b, unsigned L, uint64_t out[3]) {
  uint64_t dist = 0, len = 0;
  for (unsigned k = 0; k < L; k++) {
    const char x = rnj_oracle_protein_resolve(a[k]), y = rnj_oracle_protein_resolve(b[k]);
    if (x != '.' && x != '-' && y != '.' && y != '-') {
      if (x != y) dist++;
      len++;
    }
  }
  out[0] = dist;
  out[1] = 0;
  out[2] = len;
}

/* The reference's dormant protein kernel, computeDistanceProtein
 * (gpu/distanceCalculatorProtein_gpu.cu:52-104), on two staged rows of `padded` bytes (the
 * loader's resolved characters, its '-' padding and the zero fill of dataloader::initGPUData):
 * a byte below 64 is a gap (uchar4 components, unsigned); out[0] = positions where neither is a
 * gap and the bytes differ, out[1] = positions where either is a gap, out[2] = 0. */
void rnj_oracle_protein_gpu_counts(const unsigned char* a, const unsigned char* b,
                                   unsigned padded, uint64_t out[3]) {
  uint64_t r = 0, g = 0;
  for (unsigned k = 0; k < padded; k++) {
    const int gap = a[k] < 64 || b[k] < 64;
    g += gap;
    r += (a[k] != b[k]) && !gap;
  }
  out[0] = r;
  out[1] = g;
  out[2] = 0;
}

/* the Dayhoff PAM distances (x100) for p = 0.750 .. 0.930, KimuraDistance.cpp:30-49 */
static const int kDayhoffPams[] = {
  195, 196, 197, 198, 199, 200, 200, 201, 202, 203, 204, 205, 206, 207, 208, 209, 209, 210,
  211, 212, 213, 214, 215, 216, 217, 218, 219, 220, 221, 222, 223, 224, 226, 227, 228, 229,
  230, 231, 232, 233, 234, 236, 237, 238, 239, 240, 241, 243, 244, 245, 246, 248, 249, 250,
  252, 253, 254, 255, 257, 258, 260, 261, 262, 264, 265, 267, 268, 270, 271, 273, 274, 276,
  277, 279, 281, 282, 284, 285, 287, 289, 291, 292, 294, 296, 298, 299, 301, 303, 305, 307,
  309, 311, 313, 315, 317, 319, 321, 323, 325, 328, 330, 332, 335, 337, 339, 342, 344, 347,
  349, 352, 354, 357, 360, 362, 365, 368, 371, 374, 377, 380, 383, 386, 389, 393, 396, 399,
  403, 407, 410, 414, 418, 422, 426, 430, 434, 438, 442, 447, 451, 456, 461, 466, 471, 476,
  482, 487, 493, 498, 504, 511, 517, 524, 531, 538, 545, 553, 560, 569, 577, 586, 595, 605,
  615, 626, 637, 649, 661, 675, 688, 703, 719, 736, 754, 775, 796, 819, 845, 874, 907, 945, 988};

float rnj_oracle_kimura_protein(uint64_t dist, uint64_t valid) {
  float distance = -1.0f;
  const long total = (long)dist;
  if (valid != 0) distance = total / (float)valid;
  if (distance < 0.75) {
    if (distance > 0.0) distance = (float)(-log(1.0 - distance - (distance * distance * 0.20)));
  } else {
    if (distance > 0.930) {
      distance = 10.0f;
    } else {
      const int idx = (int)((distance * 1000.0) - 750.0);
      distance = (float)(kDayhoffPams[idx] / 100.0);
    }
  }
  return distance;
}

float rnj_oracle_jc_protein(uint64_t dist, uint64_t valid) {
  const double alpha = 20.0;
  const long total = (long)dist;
  float distance;
  if (valid == 0) distance = 0;
  else distance = total / (float)valid;
  if (distance / ((alpha - 1) / alpha) >= 1.0) distance = -1;
  else distance = (float)(-((alpha - 1) / alpha) * log(1.0 - distance / ((alpha - 1) / alpha)));
  return distance;
}

int rnj_oracle_protein_matrix(const char* aln, unsigned n, unsigned L, int model, int naive,
                              float* full, float* max_out) {
  const size_t npairs = (size_t)n * (n - 1) / 2;
  float* t = (float*)malloc((npairs ? npairs : 1) * sizeof(float));
  if (!t) return -1;
  float maxd = 0.0f;
  size_t k = 0;
  for (unsigned i = 0; i < n; i++)
    for (unsigned j = i + 1; j < n; j++, k++) {
      uint64_t c[3];
      if (naive) rnj_oracle_count_protein_naive(aln + (size_t)i * L, aln + (size_t)j * L, L, c);
      else rnj_oracle_count_protein(aln + (size_t)i * L, aln + (size_t)j * L, L, c);
      const float d = model == RNJ_ORACLE_K2P ? rnj_oracle_kimura_protein(c[0], c[2])
                                              : rnj_oracle_jc_protein(c[0], c[2]);
      if (d > maxd) maxd = d;
      t[k] = d;
    }
  post_process(t, n, maxd, full);
  if (max_out) *max_out = maxd;
  free(t);
  return 0;
}
