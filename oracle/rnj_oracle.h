/* CPU restatement of rapidNJ's DNA pairwise-distance path.
 *
 * TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg
 * may load this library, and only as the checker; the product path (rapidnj_amd, include/) never
 * links, loads or falls back to it.
 *
 * Every function cites the reference lines it restates. Layout of packed data follows the
 * reference exactly: per taxon `bsc` 128-bit words (bsc = L/64 + 6), 64 sites per word, site s in
 * 32-bit word s/16 at bits 2*(s%16); codes A=00 G=01 T=10 C=11; the gap filter holds 01 in the
 * slot of every A/C/G/T (either case) and 00 elsewhere.
 *
 * Parity pins (tests/test_oracle.py):
 *   packing        golden words from the compiled reference (tests/golden/) + char rules
 *   counts         golden ts/tv/valid from the compiled reference; char-level brute force with
 *                  an independent purine/pyrimidine formulation; invariants
 *   corrections    golden float bits from the compiled reference; closed-form special cases
 *   post-process   golden matrices (symmetry, diagonal, sentinel = 2*max)
 */
#ifndef RNJ_ORACLE_H
#define RNJ_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { RNJ_ORACLE_JC = 0, RNJ_ORACLE_K2P = 1 };

/* bitStringsCount for a DNA alignment of length L (dataloaderFasta.hpp:222). */
/* FASTA reader, dataloaderFasta::parseData / storeSequence restated (dataloaderFasta.hpp:78-133,
 * 201-276), bytes as signed chars. Two calls: with rows == NULL it returns the shape (*n
 * sequences of *L sites, *H names holding *name_bytes bytes); with buffers sized from that it
 * fills rows (n * L bytes, the kept bytes of every stored record), names (the name bytes
 * back to back) and name_len (H lengths). Returns 0, 1 when a stored record's length differs
 * from the first one's (*bad = its sequence index, *bad_len = its length; the reference exits
 * there) or 2 with fewer than 2 sequences. */
int rnj_oracle_parse_fasta(const unsigned char* data, size_t len, unsigned* n, unsigned* L,
                           unsigned* H, size_t* name_bytes, unsigned* bad, unsigned* bad_len,
                           unsigned char* rows, unsigned char* names, unsigned* name_len);

/* Stockholm reader, dataloaderStockholm::parseData / parseSequenceName /
 * parseSequenceLineFastDist restated (dataloaderStockholm.hpp:83-111, 118-132, 183-236, 294-314),
 * bytes as signed chars: lines starting (after bytes below 32) with '#' are markup, "//" ends the
 * alignment, any other line is a sequence whose name is its first run of bytes above 32 (a
 * leading '/' included) and whose data is every byte above 32 after it up to the '\n' -- on
 * the next line when the name ends the line. Two calls as for rnj_oracle_parse_fasta (one name
 * per sequence). Returns 0; 1 when a sequence's length differs from the first one's (*bad =
 * its index, *bad_len = its length); 2 with fewer than 2 sequences; 3 when a name is followed
 * by an empty line (*bad = its index: "has zero length"); 4 when the input ends before "//"
 * ("Stream ended unexpectedly"). */
int rnj_oracle_parse_stockholm(const unsigned char* data, size_t len, unsigned* n, unsigned* L,
                               size_t* name_bytes, unsigned* bad, unsigned* bad_len,
                               unsigned char* rows, unsigned char* names, unsigned* name_len);

/* Protein (PROTEIN type): resolveChar (dataloaderFasta.hpp:184-199) maps - . X x z Z b B J j ?
 * to '-', and the byte compare of bitDistanceProtein::computeDistance (bitDistanceProtein.cpp
 * :31-119) over the resolved residues: out[0] = positions where both are residues and differ,
 * out[1] = 0, out[2] = roundup(L, 128) - gap positions among the 128 * (L / 128 + 1) positions
 * the estimator walks (padding counts as gap), in unsigned 64-bit arithmetic. */
char rnj_oracle_protein_resolve(char c);
void rnj_oracle_count_protein(const char* a, const char* b, unsigned L, uint64_t out[3]);
/* The -f path's protein counter, simpleDistanceCalculator::calculateHammingDistance
 * (simpleDistanceCalculator.cpp:40-55), on the resolved characters: out[0] = positions where
 * both are residues and differ, out[1] = 0, out[2] = positions where both are residues. */
void rnj_oracle_count_protein_naive(const char* a, const char* b, unsigned L, uint64_t out[3]);
/* The reference's dormant protein kernel (gpu/distanceCalculatorProtein_gpu.cu:52-104) on two
 * staged rows of `padded` bytes: bytes below 64 are gaps; out[0] = both residues and differ,
 * out[1] = either a gap, out[2] = 0. */
void rnj_oracle_protein_gpu_counts(const unsigned char* a, const unsigned char* b,
                                   unsigned padded, uint64_t out[3]);
/* Kimura protein with the Dayhoff PAM table (KimuraDistance.cpp:30-49,322-338) and JC with
 * alpha = 20 (JCdistance.cpp:155-183) on those counts. */
float rnj_oracle_kimura_protein(uint64_t dist, uint64_t valid);
float rnj_oracle_jc_protein(uint64_t dist, uint64_t valid);
/* Post-processed n x n protein matrix (model RNJ_ORACLE_K2P = Kimura protein, RNJ_ORACLE_JC) of
 * an (n, L) character alignment (raw characters; resolved here), counted by bitDistanceProtein
 * (naive = 0) or by the -f path's Hamming counter (naive = 1). */
int rnj_oracle_protein_matrix(const char* aln, unsigned n, unsigned L, int model, int naive,
                              float* full, float* max_out);

/* rapidNJ's neighbour joining (rapidNJ.cpp, restated in rnj_oracle_nj.c) on an n x n
 * row-major matrix (copied): joins k = 0..n-3 as polytree::addInternalNode receives them
 * (left[k], right[k] = the joined clusters' ids, llen/rlen their branch lengths; the new cluster
 * is id n + k), then the last two cluster ids and their distance
 * (polytree::set_serialization_indices). Returns 0. */
int rnj_oracle_nj(const float* matrix, unsigned n, int negative_branches, int* left, int* right,
                  double* llen, double* rlen, int* last, float* last_dist);
/* The device's row-skipping rule audited along rapidNJ's run (rnj_oracle_nj.c); out[4] = bound
 * violations, record violations, rows skipped, rows evaluated; flags bit 0 plants a mistake. */
int rnj_oracle_nj_skip_audit(const float* matrix, unsigned n, int negative_branches, int flags,
                             long long* out);
/* polytree::serialize_tree (polytree.cpp:402-458, no bootstrap counts) of such a tree into out
 * (NUL-terminated when it fits); returns the length the text needs. */
size_t rnj_oracle_newick(unsigned n, const int* left, const int* right, const double* llen,
                         const double* rlen, const int* last, float last_dist,
                         const char* const* names, char* out, size_t cap);

unsigned rnj_oracle_bitstrings_count(unsigned L);

/* 1 if the first sequence classifies the alignment as DNA (dataloaderFasta.hpp:54-76). */
int rnj_oracle_is_dna(const char* seq, unsigned L);

/* Pack n sequences (row-major n x L chars) into bits/gaps, each n x (bsc*4) u32
 * (dataloaderFasta.hpp:147-182 encoding, :219-251 padding with '-'). */
void rnj_oracle_pack(const char* aln, unsigned n, unsigned L, unsigned bsc, uint32_t* bits,
                     uint32_t* gaps);

/* Char-level brute force with the packed path's semantics: a site counts when both chars are
 * one of ACGTacgt; tv when one is a purine (A,G) and the other a pyrimidine (C,T); ts when both
 * are in the same class and differ. out = {ts, tv, valid}. */
void rnj_oracle_count_chars(const char* a, const char* b, unsigned L, uint64_t out[3]);

/* The reference's naive (-f) counter on raw chars, including its resolveChar step
 * (simpleDistanceCalculator.cpp:58-98, dataloaderFasta.hpp:184-199,268-271). */
void rnj_oracle_count_naive(const char* a, const char* b, unsigned L, uint64_t out[3]);

/* Word-level restatement of bitDistanceGap's per-word logic (bitDistanceGap.cpp:138-161) with
 * scalar popcounts over `nwords` 32-bit words. */
void rnj_oracle_count_words(const uint32_t* bi, const uint32_t* gi, const uint32_t* bj,
                            const uint32_t* gj, unsigned nwords, uint64_t out[3]);

/* SSE2 SWAR counter with the reference's lane-widening schedule (bitDistanceGap.cpp:60-98,
 * 134-336): processes 6*floor(bsc/6) 128-bit words. */
void rnj_oracle_count_swar(const uint32_t* bi, const uint32_t* gi, const uint32_t* bj,
                           const uint32_t* gj, unsigned bsc, uint64_t out[3]);

/* Corrections (KimuraDistance.cpp:339-348, JCdistance.cpp:172-183); -1 marks saturation. */
float rnj_oracle_k2p(uint64_t ts, uint64_t tv, uint64_t valid);
float rnj_oracle_jc(uint64_t ts, uint64_t tv, uint64_t valid);

/* Counts of every pair i<j into packed upper-triangle arrays (row-major, row i holds j=i+1..n-1),
 * with `threads` pthreads walking rows i = tid (mod threads) (KimuraDistance.cpp:300-320). */
int rnj_oracle_counts_triangle(const uint32_t* bits, const uint32_t* gaps, unsigned n,
                               unsigned bsc, int threads, uint32_t* ts, uint32_t* tv,
                               uint32_t* valid);

/* The whole driver: threaded counts + correction + global max + post-process
 * (KimuraDistance.cpp:75-99,212-297,300-364,366-395; JCdistance.cpp:24-44,92-145,148-230).
 * tri (n(n-1)/2, may be NULL) receives the per-pair values before post-processing (-1 for
 * saturated pairs); full (n*n, may be NULL) the post-processed symmetric matrix; *max_out the
 * global maximum used for the sentinel (sentinel = 2*max). Returns 0 on success. */
int rnj_oracle_distance_matrix(const uint32_t* bits, const uint32_t* gaps, unsigned n,
                               unsigned bsc, int model, int threads, float* tri, float* full,
                               float* max_out);

/* Correction + post-process only, from given upper-triangle counts. */
int rnj_oracle_correct_triangle(const uint32_t* ts, const uint32_t* tv, const uint32_t* valid,
                                unsigned n, int model, float* tri, float* full, float* max_out);

#ifdef __cplusplus
}
#endif

#endif /* RNJ_ORACLE_H */
