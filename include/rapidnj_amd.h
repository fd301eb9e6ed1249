/* rapidnj_amd: MI355X (gfx950) pairwise evolutionary-distance engine for DNA and protein
 * alignments.
 *
 * C ABI of librapidnj_amd.so. Plain pointers and sizes only. Two groups of entry points:
 *
 *  1. The reference's own device boundary, same names, arguments and meaning
 *     (src/distanceCalculation/gpu/kimuraDNA_gpu.cu:199-276,
 *      src/distanceCalculation/gpu/distanceCalculatorDNA_gpu.cu:190-287, declared by the caller at
 *      src/distanceCalculation/KimuraDistance.cpp:17-28). One implicit module-wide state on device
 *      0; errors print a message and exit(1), as the reference's cutilSafeCall does.
 *
 *  2. A handle-based API (rnj_*) that returns a status code (RNJ_OK = 0) and keeps the reason in
 *     a per-thread message (rnj_last_error). Used by the C++ host classes in
 *     include/rapidnj_amd_distance.hpp, which mirror the reference's KimuraDistance / JCdistance.
 *
 * Counts are exact integers: ts (transitions), tv (transversions) and valid (sites where both
 * taxa hold one of ACGTacgt). Distances follow the reference's CPU recipes in FP64
 * (KimuraDistance.cpp:339-348, JCdistance.cpp:170-183); saturated pairs carry -1 until the
 * reference post-process replaces them by 2 * max (KimuraDistance.cpp:366-395).
 * Protein alignments (type discovered as the reference's readers do, or set with
 * rnj_set_input_type) count mismatched residues and the estimator's valid sites
 * (bitDistanceProtein.cpp:31-61); RNJ_K2P is then the Kimura protein distance with the Dayhoff
 * table (KimuraDistance.cpp:322-338) and RNJ_JC the Jukes-Cantor distance with alpha = 20
 * (JCdistance.cpp:155-158).
 */
#ifndef RAPIDNJ_AMD_H
#define RAPIDNJ_AMD_H

#include <stddef.h>
#include <stdint.h>

#if defined(__GNUC__)
#define RNJ_EXPORT __attribute__((visibility("default")))
#else
#define RNJ_EXPORT
#endif

#ifdef __cplusplus
extern "C" {
#endif

#define RNJ_API_VERSION 2

/* status codes */
enum {
  RNJ_OK = 0,
  RNJ_E_INVALID = 1,    /* bad argument (null pointer, n < 2, unknown model, ...) */
  RNJ_E_HIP = 2,        /* a HIP runtime call failed */
  RNJ_E_NOMEM = 3,      /* device or pinned host allocation failed */
  RNJ_E_NOT_DNA = 4,    /* a DNA-only operation (served jobs, character updates) on protein */
  RNJ_E_ADMISSION = 5,  /* a served job can never fit the server's HBM budget */
  RNJ_E_CANCELLED = 6,  /* a served job was cancelled */
  RNJ_E_NOT_READY = 7,  /* a served job has not finished (rnj_server_poll) */
  RNJ_E_NO_DEVICE = 8   /* no HIP device */
};

/* models: rnj_*_distances */
enum { RNJ_JC = 0, RNJ_K2P = 1 };

/* alignment types (the reference's InputType and its -i d|p option, src/main.cpp:397-399) */
enum { RNJ_TYPE_AUTO = 0, RNJ_TYPE_DNA = 1, RNJ_TYPE_PROTEIN = 2 };

/* output layouts for device distance matrices */
enum {
  RNJ_TRIANGLE = 0, /* n(n-1)/2 floats, upper triangle row-major: row i holds j = i+1..n-1 */
  RNJ_SQUARE = 1    /* n rows of `ld` floats: symmetric, diagonal 0 */
};

/* flags for rnj_distances_device / rnj_bootstrap_device */
enum {
  RNJ_FLAG_RAW = 1,  /* keep -1 for saturated pairs (no 2*max post-process) */
  RNJ_FLAG_EXACT = 2 /* the library's compensated FP64 log (<= 1 ulp). Default: a leaner FP64
                        log (~2 ulp of double), which rounds to the same float except in rare
                        ties; counts, saturated pairs and zeros are exact either way */
};

typedef struct rnj_context rnj_context;
typedef struct rnj_alignment rnj_alignment;
typedef struct rnj_server rnj_server;
typedef struct rnj_plan rnj_plan;
typedef struct rnj_batch_plan rnj_batch_plan;

/* ---- reference device boundary (group 1) ------------------------------------------------ */

/* kimuraDNA_gpu.cu:241-276. bsStride: u32 words per taxon in the staging layout of
 * dataloader::initGPUData (dataloader.hpp:38-70, 16 sites per word); the first
 * min(bsStride, 4*bitStringCount) words of each taxon are counted. */
RNJ_EXPORT void initialiseDNA2_gpu(unsigned int sequenceCount, unsigned int bitStringCount,
                        unsigned int bsStride);
/* kimuraDNA_gpu.cu:235-239: host arrays [sequenceCount][bsStride] of codes and gap filters. */
RNJ_EXPORT void storeDataDNA2_gpu(unsigned int* bitStrings, unsigned int* gapFilters);
/* kimuraDNA_gpu.cu:199-233: K2P distances into host results[sequenceCount*sequenceCount],
 * row-major. Entries j > i hold the distance with -1 for saturated pairs (the caller,
 * KimuraDistance::computeDistancesDNAGPU2 at KimuraDistance.cpp:164-193, post-processes);
 * this implementation also fills j < i symmetrically and the diagonal with 0. */
RNJ_EXPORT void computeDistancesDNA2_gpu(float* results);

/* distanceCalculatorDNA_gpu.cu:244-287, :238-242: same arguments for the raw-count path. */
RNJ_EXPORT void initialiseDNA_gpu(unsigned int sequenceCount, unsigned int bitStringCount,
                       unsigned int bsStride);
RNJ_EXPORT void storeDataDNA_gpu(unsigned int* bitStrings, unsigned int* gapFilters);
/* distanceCalculatorDNA_gpu.cu:190-226: ts, tv and gaps (= valid sites) into host arrays of
 * sequenceCount*sequenceCount u32, row-major, entries j > i (j < i mirrored, diagonal 0). */
RNJ_EXPORT void computeDistancesDNA_gpu(unsigned int* ts, unsigned int* tv, unsigned int* gaps);
/* distanceCalculatorDNA_gpu.cu:228-236: waits for the device; results were already copied. */
RNJ_EXPORT void getResultsDNA_gpu(unsigned int* ts, unsigned int* tv, unsigned int* gaps);

/* The protein half (gpu/distanceCalculatorProtein_gpu.cu:139-229, declared by the caller at
 * KimuraDistance.cpp:20-22 and called by KimuraDistance::computeDistancesProteinGPU at
 * :113-162). initialiseProtein_gpu: sequenceCount taxa staged as dataloader::initGPUData does
 * for protein (dataloader.hpp:38-70: bsStride u32 words per taxon, 4 characters each), of which
 * the first paddedLength characters (the caller passes L rounded up to 128) are compared. */
RNJ_EXPORT void initialiseProtein_gpu(unsigned int sequenceCount, unsigned int bsStride,
                                      unsigned int paddedLength);
/* distanceCalculatorProtein_gpu.cu:139-176: bitStrings is the staged host array
 * [sequenceCount][bsStride]. For every pair, over the paddedLength positions, a character below
 * 64 is a gap (the kernel's test, :76-90; the loader's '-' padding and the zero fill beyond it
 * are gaps); dist[i*n+j] receives the positions where both are residues and differ, gaps[i*n+j]
 * the positions where either is a gap, for j > i (this implementation also mirrors j < i and
 * writes 0 on the diagonal). Host arrays of sequenceCount^2 u32. */
RNJ_EXPORT void computeDistancesProtein_gpu(unsigned int* bitStrings, unsigned int* dist,
                                            unsigned int* gaps);
/* distanceCalculatorProtein_gpu.cu:178-187: waits for the device; computeDistancesProtein_gpu
 * has copied the results (the reference's version prints that it is not implemented and exits). */
RNJ_EXPORT void getResultsProtein_gpu(unsigned int* dist, unsigned int* gaps);

/* ---- handle API (group 2) --------------------------------------------------------------- */

RNJ_EXPORT const char* rnj_last_error(void);           /* message of this thread's last failure */
RNJ_EXPORT const char* rnj_status_string(int status);  /* static name of a status code */
RNJ_EXPORT int rnj_api_version(void);

/* Context on HIP device `device`: streams, kernel configuration and scratch buffers.
 * No counterpart in the reference (its GPU state is module-static, kimuraDNA_gpu.cu:178-191). */
RNJ_EXPORT int rnj_create(int device, rnj_context** out);
RNJ_EXPORT int rnj_destroy(rnj_context* ctx);
RNJ_EXPORT int rnj_synchronize(rnj_context* ctx); /* waits for all work the context enqueued */

/* Tile-kernel configurations: -1 = choose per shape with the built-in autotuner (default). */
RNJ_EXPORT int rnj_tile_config_count(void);
RNJ_EXPORT const char* rnj_tile_config_name(int config);
RNJ_EXPORT int rnj_set_tile_config(rnj_context* ctx, int config);
RNJ_EXPORT int rnj_get_tile_config(rnj_context* ctx, unsigned n, unsigned L, int model, int* config);
/* The configuration the next launch of this alignment and model uses (tuned per shape class and
 * alignment type: DNA, or protein of a code width). */
RNJ_EXPORT int rnj_get_tile_config_for(rnj_context* ctx, const rnj_alignment* aln, int model,
                                       int* config);

/* Measured issue rates (lane-operations per second) of the count loop's two instruction classes
 * on this context's device: the bitwise class (v_bitop3_b32, as v_and / v_xor / v_or) and
 * v_bcnt_u32_b32. Each is timed on 16 independent dependency chains per lane at 8 waves per SIMD.
 * The ceiling of a count loop that issues b bitwise and c bcnt instructions per 32 site-pairs is
 * (b + c) / (b / bitwise + c / bcnt) lane-ops per second: K2P issues 3 + 3, JC 3 + 2. Takes a
 * few milliseconds. No counterpart in the reference; it is the device's int32 VALU peak for this
 * path, which no data sheet states. */
RNJ_EXPORT int rnj_valu_issue_rates(rnj_context* ctx, double* bitwise, double* bcnt);

/* Count-loop instruction mixes (rnj_valu_mix_rate). */
enum { RNJ_MIX_K2P = 0, RNJ_MIX_JC = 1, RNJ_MIX_NAIVE = 2, RNJ_MIX_PROTEIN = 3 };
/* Measured rate (lane-operations per second) at which this context's device executes one count
 * loop's own per-pair instruction sequence: K2P (v_and, 2 v_bitop3, 3 v_bcnt), JC (v_and,
 * 2 v_bitop3, 2 v_bcnt), the naive rules (v_and, v_xor, 4 v_bitop3, 3 v_bcnt) or protein with
 * `planes` = 5, 6 or 8 code planes (v_xor, planes - 1 v_bitop3, v_and, 2 v_bcnt), from registers
 * at 8 waves per SIMD, best of three timed launches. Mixed, the two classes do not reach the
 * rates rnj_valu_issue_rates measures for each alone, so this is the VALU ceiling of a count
 * loop: no schedule of that instruction sequence runs faster. Takes a few milliseconds. No
 * counterpart in the reference. */
RNJ_EXPORT int rnj_valu_mix_rate(rnj_context* ctx, int mix, int planes, double* rate);

/* Type of the alignments this context reads from characters, rows and FASTA from now on:
 * RNJ_TYPE_AUTO (default) classifies the first sequence as the reference's readers do
 * (dataloaderFasta.hpp:54-76: DNA unless a byte >= 'A' is not one of ACGTUN of either case, else
 * protein); RNJ_TYPE_DNA / RNJ_TYPE_PROTEIN force the type, like the reference's `-i d|p`. */
RNJ_EXPORT int rnj_set_input_type(rnj_context* ctx, int type);
/* Character rules of alignments read from characters, rows or FASTA afterwards
 * (rnj_alignment_from_chars / _rows / _fasta / _fasta_file). RNJ_CHARS_BITS (default): the counts
 * of the bit-string engines, bitDistanceGap for DNA (bitDistanceGap.cpp:60-336), which counts
 * ACGT of either case, and bitDistanceProtein for protein (bitDistanceProtein.cpp:31-61).
 * RNJ_CHARS_NAIVE: the counts of the estimator the reference's `-f` option (fastdist off)
 * selects, simpleDistanceCalculator (src/main.cpp:589,597; KimuraDistance.cpp:195-208), after the
 * loader's resolveChar (dataloaderFasta.hpp:184-199):
 *  - DNA, calculateTsTv (simpleDistanceCalculator.cpp:58-98): a site counts when neither
 *    character is a gap (any byte but ACGTU of either case), and a mismatch is a transition or
 *    transversion only when the lower-index taxon's character is uppercase A, C, G or T. Such
 *    alignments are packed with a fourth bit plane;
 *  - protein, calculateHammingDistance (simpleDistanceCalculator.cpp:40-55): the valid count is
 *    the exact number of residue pairs (no 128-position shortfall when L is a multiple of 128).
 * Naive-rule alignments are accepted by the distance, count, hand-off, bootstrap and batch-plan
 * calls; served jobs and character updates take bit-rule alignments only (served jobs fail with
 * RNJ_E_INVALID while the server's context has naive rules set). */
enum { RNJ_CHARS_BITS = 0, RNJ_CHARS_NAIVE = 1 };
RNJ_EXPORT int rnj_set_char_rules(rnj_context* ctx, int rules);

/* Alignments packed once into HBM as bit planes.
 * from_chars: n rows of L bytes, row t at chars + t*pitch, on the host (chars_on_device = 0) or
 *   already in device memory (chars_on_device = 1). DNA: the character rules of the reference's
 *   FASTA packer, ACGT of either case are counted, everything else ('-', N, U, IUPAC, '.', '?')
 *   is not (dataloaderFasta.hpp:147-182). Protein: every byte is a residue except the gap
 *   characters - . X x z Z b B J j ? (dataloader::resolveChar, as bitDistanceProtein uses it).
 *   The type is the context's input type; `skip_type_check` reads the rows as DNA regardless.
 *   n >= 2 and L >= 1 (dataloaderFasta.hpp:32-35).
 * from_rows: one pointer per taxon, like dataloader::getSequences (dataloader.hpp:26).
 * from_bitstrings: the reference's packed strings, per taxon bitStringsCount x 128 bits of codes
 *   and gap filters, like dataloader::getBitStrings / getGapFilters (dataloader.hpp:20-24,
 *   dataloaderFasta.hpp:219-251); the sites bitDistanceGap counts (6*floor(bsc/6) words of 64,
 *   bitDistanceGap.cpp:19) are taken. */
RNJ_EXPORT int rnj_alignment_from_chars(rnj_context* ctx, const char* chars, size_t pitch, unsigned n,
                             unsigned L, int chars_on_device, int skip_type_check,
                             rnj_alignment** out);
RNJ_EXPORT int rnj_alignment_from_rows(rnj_context* ctx, const char* const* rows, unsigned n, unsigned L,
                            rnj_alignment** out);
/* New characters for a DNA alignment of the same n x L (host rows at `pitch`, or device memory),
 * packed into its planes in place on `stream`: later work on that stream, a captured batch plan's
 * replay included, sees them. Host characters have been read when the call returns. Updates of
 * one alignment are not to overlap. No counterpart in the reference (a new alignment there is a
 * new dataloader). */
RNJ_EXPORT int rnj_alignment_update_chars(rnj_context* ctx, rnj_alignment* aln, const char* chars,
                                          size_t pitch, int chars_on_device, void* stream);
RNJ_EXPORT int rnj_alignment_from_bitstrings(rnj_context* ctx, const unsigned int* const* bitStrings,
                                  const unsigned int* const* gapFilters, unsigned n,
                                  unsigned bitStringsCount, rnj_alignment** out);
/* FASTA parsed on the device: the reader dataloaderFasta::parseData / storeSequence
 * (dataloaderFasta.hpp:78-133, 201-276) restated as parallel passes over the file image in HBM,
 * followed by the packer. `data` holds the file's `len` bytes, in host memory or, with
 * data_on_device, in device memory. The same input gives the same sequences and names as the
 * reference: ';' lines are comments, '>' opens a record whose name ends at the first byte below
 * 32, sequence lines keep bytes 33..127, empty records are not stored but keep their name, data
 * before the first header forms a nameless record. Errors as the reference's exits: fewer than
 * 2 sequences, or a sequence whose length differs from the first (RNJ_E_INVALID, message names
 * the sequence). The type as for from_chars. Synchronous. */
RNJ_EXPORT int rnj_alignment_from_fasta(rnj_context* ctx, const char* data, size_t len,
                                        int data_on_device, rnj_alignment** out);
/* The same from a file: read through pinned buffers and uploaded while reading
 * (dataloaderFasta::load, dataloaderFasta.hpp:24-52). */
RNJ_EXPORT int rnj_alignment_from_fasta_file(rnj_context* ctx, const char* path, rnj_alignment** out);
/* Single-line Stockholm parsed on the device, replacing dataloaderStockholm::load
 * (dataloaderStockholm.hpp:22-111, 183-236, 294-314, the reader behind `-i sth`, src/main.cpp
 * :362-363,400-401): the reference's byte-at-a-time reader restated as a scan of a two-state
 * line automaton over the file image in HBM, then the packer. Lines starting (after bytes below
 * 32) with '#' are markup; "//" ends the alignment and nothing after it is read; any other line
 * is a sequence: its name is the first run of bytes above 32 (a name ending its line takes its
 * data from the next line), its data every byte above 32 after that up to the '
'. Interleaved
 * blocks are read as the reference reads them, every line a sequence. Errors as the reference's
 * exits (RNJ_E_INVALID, first in file order): a sequence whose length differs from the first
 * (message names it), a name followed by an empty line ("has zero length"), the input ending
 * before "//", fewer than 2 sequences; and, where the reference would go on with empty
 * sequences, zero-length sequences. The type as for from_chars, except that under the naive
 * character rules it is discovered after the reference's resolveChar (- . X x z Z b B J j ? as
 * gaps), as the -f path does. Synchronous. */
RNJ_EXPORT int rnj_alignment_from_stockholm(rnj_context* ctx, const char* data, size_t len,
                                            int data_on_device, rnj_alignment** out);
/* The same from a file (dataloaderStockholm::parseData's input, read through pinned buffers and
 * uploaded while reading). */
RNJ_EXPORT int rnj_alignment_from_stockholm_file(rnj_context* ctx, const char* path,
                                                 rnj_alignment** out);
/* Names of an alignment read from FASTA, one per '>' record in file order
 * (dataloaderFasta::getSequenceNames, dataloaderFasta.hpp:298-300); 0 / NULL otherwise. */
RNJ_EXPORT unsigned rnj_alignment_name_count(const rnj_alignment* aln);
RNJ_EXPORT const char* rnj_alignment_name(const rnj_alignment* aln, unsigned k);
RNJ_EXPORT int rnj_alignment_free(rnj_alignment* aln);
/* n taxa, L sites, W 32-site words counted per taxon, bytes of HBM held. */
RNJ_EXPORT int rnj_alignment_info(const rnj_alignment* aln, unsigned* n, unsigned* L, unsigned* W,
                       size_t* bytes);
/* RNJ_TYPE_DNA or RNJ_TYPE_PROTEIN, and the number of bit planes: 3 for DNA (H, L, V), 1 + B for
 * protein (gap bits, then the B bits of each site's residue code; B = 5, 6 or 8 for up to 31,
 * 63 or 255 distinct residue bytes, which are numbered 1.. in byte order, gaps 0). */
RNJ_EXPORT int rnj_alignment_type(const rnj_alignment* aln, int* type, unsigned* planes);
/* Copy the packed planes [planes][W][n] (site s of word w at bit s%32) to the host. */
RNJ_EXPORT int rnj_alignment_planes(const rnj_alignment* aln, uint32_t* host_planes);

/* One distance matrix into device memory `out` (layout RNJ_TRIANGLE or RNJ_SQUARE with row
 * stride `ld` floats), enqueued on `stream` (a hipStream_t; NULL = HIP's null stream).
 * Saturated pairs become 2 * max unless RNJ_FLAG_RAW. If max_out (device pointer) is non-null,
 * the matrix maximum (KimuraDistance.cpp:282-287) is written there as float.
 * Replaces KimuraDistance::computeDistances / JCdistance::computeDistanceMatrix
 * (KimuraDistance.cpp:75-99, JCdistance.cpp:24-44) for one job. */
RNJ_EXPORT int rnj_distances_device(rnj_context* ctx, const rnj_alignment* aln, int model, int layout,
                         float* out, size_t ld, int flags, float* max_out, void* stream);

/* rnj_distances_device, same arguments and results, for callers that cannot allocate device
 * memory: code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc. No
 * counterpart in the reference. It calls no hipMalloc / hipFree and never tunes: the launch's
 * scratch (job table, maximum, saturation records, split partial counts) is a fixed 64 MiB
 * device array per GPU, held from the first rnj_create on that GPU until the process ends.
 * Matrices whose scratch does not fit (above about 22000 taxa) fail with RNJ_E_INVALID before
 * anything is enqueued. Every call on one GPU shares that array: calls that overlap on one GPU
 * (other streams, other threads, other contexts) race, so the caller orders them. */
RNJ_EXPORT int rnj_distances_device_noalloc(rnj_context* ctx, const rnj_alignment* aln, int model,
                                            int layout, float* out, size_t ld, int flags,
                                            float* max_out, void* stream);

/* Exact ts / tv / valid of every pair into device arrays of n(n-1)/2 u32, upper triangle
 * row-major (DistanceEstimate::computeDistance, DistanceEstimate.hpp:14). Protein: ts holds the
 * mismatches (data[0]), tv 0 (data[1]) and valid the low 32 bits of the 64-bit data[2], which
 * is 128 short of the residue pairs when L is a multiple of 128 (bitDistanceProtein.cpp:31-61)
 * unless the alignment was read under the naive character rules (then the residue pairs). */
RNJ_EXPORT int rnj_counts_device(rnj_context* ctx, const rnj_alignment* aln, uint32_t* ts, uint32_t* tv,
                      uint32_t* valid, void* stream);

/* Copy `bytes` of device memory (e.g. a triangle from rnj_distances_device) into caller-owned
 * pageable host memory, after the work already enqueued on `stream` (null: the context's
 * stream). The bytes pass through the context's pinned staging chunks, emptied by library
 * threads while the next chunk's DMA runs. Synchronous. The reference copies its GPU results
 * with one synchronous cudaMemcpy per launch into malloc'ed memory (kimuraDNA_gpu.cu:224);
 * this is that copy for any device result. */
RNJ_EXPORT int rnj_copy_to_host(rnj_context* ctx, void* dst, const void* src, size_t bytes,
                                void* stream);

/* Host hand-off: rows[i] are n caller-owned arrays of n floats (e.g. new distType[n]), filled
 * with the post-processed symmetric matrix, diagonal 0, sentinels patched: the in-memory matrix
 * handed to distMatrixReader (src/main.cpp:419-437, distMatrixReader.cpp:14-23). Synchronous.
 * max_out (host, may be null) receives the maximum used for the sentinel. The host hand-offs
 * (this, rnj_distances_stream_rows and the group-1 entry points) use the exact FP64 logs. */
RNJ_EXPORT int rnj_distances_rows(rnj_context* ctx, const rnj_alignment* aln, int model,
                                  float** rows, float* max_out);
/* The same hand-off into the rows rapidNJMem keeps: rows[i] holds i + 1 floats, row i of the
 * post-processed matrix up to and including the diagonal (the lower triangle that
 * distMatrixReader::initializeData cuts out for halfMatrix, distMatrixReader.cpp:26-35). Half
 * the bytes of rnj_distances_rows. */
RNJ_EXPORT int rnj_distances_rows_lower(rnj_context* ctx, const rnj_alignment* aln, int model,
                                        float** rows, float* max_out);

/* Row streaming for the disk-backed path: the matrix is computed once in HBM, then every row
 * (post-processed, full length n) is passed to `sink` in row order, as
 * KimuraDistance::computeDistanceMatrixMT hands rows to diskMatrix::writeArray
 * (KimuraDistance.cpp:267-281, diskMatrix.cpp:96-101). A non-zero return from the sink stops
 * the stream with RNJ_E_CANCELLED. `values` is valid only during the call. The sink is called
 * from one library thread, in row order, never concurrently. */
typedef int (*rnj_row_sink)(void* user, unsigned row, const float* values, unsigned n);
RNJ_EXPORT int rnj_distances_stream_rows(rnj_context* ctx, const rnj_alignment* aln, int model,
                                         rnj_row_sink sink, void* user, float* max_out);

/* Batched bootstrap: R replicates given as column-weight vectors weights[r*L + c] (host u16;
 * replicate r repeats column c weights[r][c] times, the expanded alignment of the reference's
 * resampling: dataloader::sample_fast_dist for DNA with 2 bits and protein with 8 bits per site,
 * sample_char for the `-f` character path, dataloader.cpp:16-69). Counts do not depend on the
 * column order, so replicate r equals the reference's result on its resampled alignment. DNA
 * under either character rules and protein alignments; a protein replicate of Lr = sum of its
 * weights sites counts as an alignment of Lr sites (for the multinomial case Lr = L). All
 * replicates run in one tile launch. out holds R matrices, matrix r at out + r*matrix_stride
 * floats (layout as above). */
RNJ_EXPORT int rnj_bootstrap_device(rnj_context* ctx, const rnj_alignment* aln, int model,
                         const uint16_t* weights, unsigned R, int layout, float* out,
                         size_t matrix_stride, size_t ld, int flags, void* stream);

/* Captured multi-alignment launch: the matrices of `count` alignments of any taxa and site counts
 * (models[k] RNJ_JC or RNJ_K2P; all DNA, or all protein of one code width) in ONE tile launch,
 * captured once into a HIP graph. outs[k] (device) receives alignment k's matrix in `layout`
 * (square rows of lds[k] floats). rnj_batch_plan_run replays it: the alignments' current planes
 * (see rnj_alignment_update_chars) into the same outputs, after the previous replay, allocating
 * nothing. Alignments and outputs must outlive the plan. No counterpart in the reference (its
 * distance calculators handle one alignment per call, KimuraDistance.cpp:75-99). */
RNJ_EXPORT int rnj_batch_plan_create(rnj_context* ctx, const rnj_alignment* const* alns,
                                     const int* models, unsigned count, int layout,
                                     float* const* outs, const size_t* lds, int flags,
                                     rnj_batch_plan** plan);
RNJ_EXPORT int rnj_batch_plan_run(rnj_batch_plan* plan, void* stream);
RNJ_EXPORT int rnj_batch_plan_tile_config(const rnj_batch_plan* plan);
RNJ_EXPORT int rnj_batch_plan_destroy(rnj_batch_plan* plan);

/* Captured batched bootstrap: the launch of rnj_bootstrap_device for fixed (alignment, model, R,
 * layout, out) captured once into a HIP graph and replayed per call with new weights. All
 * device memory is allocated at creation; a run allocates nothing. max_length bounds the sum
 * of every replicate's weights (0 = L, the multinomial case); a run whose weights exceed it is
 * rejected. Runs of one plan are serialised by the plan. No counterpart in the reference (its
 * bootstrap recomputes replicates one by one, src/main.cpp:519-532). */
RNJ_EXPORT int rnj_bootstrap_plan_create(rnj_context* ctx, const rnj_alignment* aln, int model,
                                         unsigned R, unsigned max_length, int layout, float* out,
                                         size_t matrix_stride, size_t ld, int flags,
                                         rnj_plan** plan);
RNJ_EXPORT int rnj_bootstrap_plan_run(rnj_plan* plan, const uint16_t* weights, void* stream);
RNJ_EXPORT int rnj_plan_tile_config(const rnj_plan* plan); /* configuration the graph uses */
RNJ_EXPORT int rnj_plan_destroy(rnj_plan* plan);

/* ---- neighbour joining on the device ----------------------------------------------------------
 * rapidNJ::run (src/rapidNJ.cpp:45-88) with the matrix kept in HBM: the same joins in the same
 * order (the criterion's float arithmetic, rapidNJ's visiting order and row caches, redundant
 * copies, negative_branches), giving polytree's Newick text (src/polytree.cpp:402-458). Every
 * live pair is evaluated on the device; the iterations where rapidNJ's sorted row lists decide
 * the result -- several pairs of a winning row with equal criterion and distance (taken in
 * std::sort's order of the row list), or a row whose scan rapidNJ ends by its rounded bound
 * (src/rapidNJ.cpp:340) before a pair below the running minimum -- are detected on the device
 * and searched on the host over the rows' lists as rapidNJ built and sorted them. Replaces
 * `new rapidNJ(reader, n, negative_branches, pb)->run()` (src/main.cpp) for in-memory matrices.
 * Device memory: two slot matrices of (1.5 n)^2 floats and a log of every join's new row
 * (about 4.5 n^2 bytes). */
typedef struct rnj_tree rnj_tree;
/* matrix: n rows of ld floats in device memory (post-processed, symmetric, diagonal 0), read
 * only; enqueued on `stream` after the work already there, synchronous. n >= 3. */
RNJ_EXPORT int rnj_nj_device(rnj_context* ctx, const float* matrix, unsigned n, size_t ld,
                             int negative_branches, void* stream, rnj_tree** out);
/* The same from host rows: rows[i] holds row i of the n x n matrix (distType** as
 * distMatrixReader::getMatrix returns it, distMatrixReader.cpp:155-157); uploaded through the
 * context's pinned buffers. */
RNJ_EXPORT int rnj_nj_rows(rnj_context* ctx, const float* const* rows, unsigned n,
                           int negative_branches, rnj_tree** out);
/* Distances (model RNJ_K2P / RNJ_JC, with the exact FP64 logs the host hand-off uses) and
 * neighbour joining without the matrix leaving HBM; leaf names are the alignment's FASTA names. */
RNJ_EXPORT int rnj_nj_alignment(rnj_context* ctx, const rnj_alignment* aln, int model,
                                int negative_branches, rnj_tree** out);
RNJ_EXPORT unsigned rnj_tree_size(const rnj_tree* tree); /* leaves */
/* Iterations whose search the host settled (ties or the scan bound, above). No counterpart in
 * the reference. */
RNJ_EXPORT unsigned rnj_tree_host_searches(const rnj_tree* tree);
/* Join k (0 <= k < n-2) joins clusters left[k] and right[k] (leaves 0..n-1, join k creates id
 * n+k) with branch lengths left_len[k], right_len[k] (polytree::addInternalNode); last[0..1] and
 * last_dist are the final two clusters and their distance. Any pointer may be NULL. */
RNJ_EXPORT int rnj_tree_joins(const rnj_tree* tree, int* left, int* right, double* left_len,
                              double* right_len, int* last, float* last_dist);
/* Newick text as polytree::serialize_tree writes it (branch lengths "%.5g", names quoted, ";\n"):
 * names[i] (or, when names is NULL, the alignment's names, else "t<i>"). Writes up to cap-1
 * bytes and a NUL; *length receives the full length. */
RNJ_EXPORT int rnj_tree_newick(const rnj_tree* tree, const char* const* names, char* out,
                               size_t cap, size_t* length);
RNJ_EXPORT int rnj_tree_free(rnj_tree* tree);

/* ---- served jobs ----------------------------------------------------------------------------
 * A request loop on its own thread admits submitted jobs while their HBM need fits the budget,
 * batches the admitted jobs into one multi-job tile launch per batch, keeps up to `slots` batches
 * on the device at once (each on its own stream, so uploads, counting and downloads of
 * different batches overlap, and a job admitted while a batch runs is launched beside it) and
 * hands each job's post-processed upper triangle back. A job is the reference's distance
 * computation of one alignment (src/distanceCalculation/DistanceEstimate.cpp, the -o m / NJ
 * matrix); the loop itself has no counterpart in the reference. */
typedef struct {
  unsigned slots;           /* batches in flight, each on its own stream (default and most: 4) */
  unsigned max_batch_jobs;  /* jobs per launch (default 64) */
  size_t hbm_budget;        /* bytes the server may hold; 0 = 80% of free HBM at creation */
  unsigned batch_window_us; /* hold a batch up to this long after its first job is ready, to
                               gather more jobs into the launch (default 0: launch at once) */
  unsigned long long max_batch_work; /* counting work per launch, in pair-words (n(n-1)/2 x
                               32-site words, summed over its jobs; a larger job runs alone);
                               0 = 2^28, about 50 us: only small jobs share launches */
} rnj_server_config;

typedef struct {
  double queued_s;   /* submit -> admitted */
  double device_s;   /* admitted -> results on the host */
  double total_s;    /* submit -> results on the host */
  unsigned batch_id; /* launch the job ran in */
  unsigned batch_jobs;
  /* stages of device_s */
  double upload_s;   /* admitted -> rows read into the staging chunks, copies and packing enqueued */
  double launch_s;   /* -> the job's batch enqueued on a slot */
  double run_s;      /* -> the batch seen finished by the request loop */
  double download_s; /* the copy of the triangle into tri_out, inside rnj_server_wait */
  double launched_at; /* seconds from rnj_server_create to the job's batch enqueued (-1: never) */
  double finished_at; /* ... to the batch seen finished by the request loop (-1: never ran) */
} rnj_job_stats;

RNJ_EXPORT int rnj_server_create(rnj_context* ctx, const rnj_server_config* cfg, rnj_server** out);
/* chars: n rows of L bytes (host), row t at chars + t*pitch; copied before returning. DNA only:
 * rows whose type (rnj_set_input_type of the server's context) is protein give RNJ_E_NOT_DNA. */
RNJ_EXPORT int rnj_server_submit(rnj_server* srv, const char* chars, size_t pitch, unsigned n, unsigned L,
                      int model, uint64_t* ticket);
/* A job on an alignment already in HBM (of the server's context; any kind: DNA under either
 * character rules, protein): nothing is uploaded. The alignment must stay alive and unchanged
 * until the job ends. tri_dev: the caller's device buffer of n(n-1)/2 floats for the
 * post-processed upper triangle, valid once rnj_server_wait (or a successful poll) returns; NULL
 * = a server buffer, copied to rnj_server_wait's tri_out. Jobs of one kind (and protein code
 * width) share a launch. */
RNJ_EXPORT int rnj_server_submit_resident(rnj_server* srv, const rnj_alignment* aln, int model,
                                          float* tri_dev, uint64_t* ticket);
/* Blocks until the job ends; tri_out (host, n(n-1)/2 floats, may be null) receives the
 * post-processed upper triangle. Releases the ticket. */
RNJ_EXPORT int rnj_server_wait(rnj_server* srv, uint64_t ticket, float* tri_out, rnj_job_stats* stats);
RNJ_EXPORT int rnj_server_poll(rnj_server* srv, uint64_t ticket); /* RNJ_OK when done, RNJ_E_NOT_READY */
RNJ_EXPORT int rnj_server_cancel(rnj_server* srv, uint64_t ticket); /* only before admission */
RNJ_EXPORT int rnj_server_destroy(rnj_server* srv);                  /* finishes admitted work first */

#ifdef __cplusplus
}
#endif

#endif /* RAPIDNJ_AMD_H */
