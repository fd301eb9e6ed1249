// C++ host classes over the C ABI (rapidnj_amd.h) with the shape of the reference's distance
// calculators, so the existing call site keeps working:
//
//   reference (src/main.cpp:255-274)             this header
//   KimuraDistance(verbose, fastdist, dl, dm)    rapidnj_amd::KimuraDistance<dataloader, diskMatrix>
//   ->computeDistances(numCores)                 ->computeDistances(numCores)
//   ->getDistanceMatrix()                        ->getDistanceMatrix()
//   JCdistance(verbose, fastdist, dl, dm)        rapidnj_amd::JCdistance<dataloader, diskMatrix>
//   ->computeDistanceMatrix(numCores)            ->computeDistanceMatrix(numCores)
//
// KimuraDistance.hpp:12-17 and JCdistance.hpp:14-18 declare the originals. The templates take the
// loader and disk-matrix types so the header needs none of the reference's headers: a Loader
// provides the getters of dataloader (dataloader.hpp:20-29) and, like dataloader, may have a
// `type` member holding the InputType (dataloader.hpp:84); a DiskMatrix provides
// writeArray(distType*, int row, int size) (diskMatrix.h:13).
//
// Protein loaders (type == PROTEIN) get the reference's protein distances: Kimura protein with
// the Dayhoff table (KimuraDistance.cpp:322-338) and JC with alpha = 20 (JCdistance.cpp:155-158),
// counted as bitDistanceProtein counts (bitDistanceProtein.cpp:31-61) with fastdist, and as
// simpleDistanceCalculator::calculateHammingDistance counts (simpleDistanceCalculator.cpp:40-55)
// without it, the estimator KimuraDistance::getDistanceEstimateInstance picks
// (KimuraDistance.cpp:195-208).
//
// Behaviour kept from the reference:
//  * without a disk matrix the result is n rows allocated with new distType[n] (ownership passes
//    to the NJ builder, which deletes them: rapidNJ.cpp:432-437), symmetric, diagonal 0, every
//    saturated pair replaced by 2 * the maximum distance (KimuraDistance.cpp:75-99,366-395);
//  * with the 5-argument constructor (bootstrap) the caller's rows are reused;
//  * with a disk matrix every post-processed row is written with writeArray in row order;
//  * bootstrap rows together with a disk matrix is an error that exits (KimuraDistance.cpp:77-80);
//  * other errors print a message and exit(1), as the reference does.
// The thread count is accepted and ignored: the device does the work.
#pragma once

#include <cstdio>
#include <cstdlib>
#include <iomanip>
#include <iostream>
#include <string>
#include <vector>

#include "rapidnj_amd.h"

namespace rapidnj_amd {

typedef float distType;  // stdinclude.h:37

// One library context per process, on device 0 or the device named by RAPIDNJ_AMD_DEVICE.
inline rnj_context* shared_context() {
  static rnj_context* ctx = [] {
    rnj_context* c = nullptr;
    const char* env = std::getenv("RAPIDNJ_AMD_DEVICE");
    const int dev = env ? std::atoi(env) : 0;
    if (rnj_create(dev, &c) != RNJ_OK) {
      std::cerr << "ERROR: rapidnj_amd: " << rnj_last_error() << std::endl;
      std::exit(1);
    }
    return c;
  }();
  return ctx;
}

inline void fail_exit(const char* where) {
  std::cerr << "ERROR: rapidnj_amd " << where << ": " << rnj_last_error() << std::endl;
  std::exit(1);
}

// The loader's InputType (stdinclude.h:26: DNA, PROTEIN, UNKNOWN are 0, 1, 2) as an RNJ_TYPE_*.
// A loader without a `type` member is classified by the library from its first sequence.
template <class Loader>
auto loader_type(Loader* loader, int) -> decltype((int)loader->type) {
  const int t = (int)loader->type;
  return t == 0 ? RNJ_TYPE_DNA : (t == 1 ? RNJ_TYPE_PROTEIN : RNJ_TYPE_AUTO);
}
template <class Loader>
int loader_type(Loader*, long) {
  return RNJ_TYPE_AUTO;
}

// Packs the loader's alignment into HBM. DNA with fastdist: the loader's bit strings, which are
// exactly what the reference's bitDistanceGap counts. Protein with fastdist: the same bit
// strings hold the resolved characters four to a word (dataloaderFasta.hpp
// encodeProteinSequence), so their first L bytes are the row. Otherwise the character rows,
// packed under the naive counter's rules (RNJ_CHARS_NAIVE), as `-f` does in the reference
// (simpleDistanceCalculator): calculateTsTv for DNA, calculateHammingDistance for protein.
// The shared context's input type is set for the upload and reset after it (the reference's
// callers are single-threaded).
template <class Loader>
rnj_alignment* upload_alignment(Loader* loader, bool fastdist) {
  rnj_alignment* aln = nullptr;
  const unsigned n = loader->getSequenceCount();
  const int type = loader_type(loader, 0);
  int rc;
  if (type == RNJ_TYPE_PROTEIN || !fastdist) {
    const char* const* rows =
        fastdist ? reinterpret_cast<const char* const*>(loader->getBitStrings())
                 : const_cast<const char* const*>(loader->getSequences()->data());
    rnj_set_input_type(shared_context(), type);
    rnj_set_char_rules(shared_context(), fastdist ? RNJ_CHARS_BITS : RNJ_CHARS_NAIVE);
    rc = rnj_alignment_from_rows(shared_context(), rows, n, loader->getSequenceLength(), &aln);
    rnj_set_char_rules(shared_context(), RNJ_CHARS_BITS);
    rnj_set_input_type(shared_context(), RNJ_TYPE_AUTO);
  } else {
    rc = rnj_alignment_from_bitstrings(shared_context(),
                                       (const unsigned int* const*)loader->getBitStrings(),
                                       (const unsigned int* const*)loader->getGapFilters(), n,
                                       loader->getBitStringsCount(), &aln);
  }
  if (rc != RNJ_OK) fail_exit("upload");
  return aln;
}

// The reference's matrix output (`-o m`), printDistanceMatrix (src/main.cpp:212-221): a tab and
// the taxon count, then per row the taxon's name, a tab, and every entry in fixed notation with
// 6 decimals, each followed by a space. For the rows a calculator hands over (names: the
// loader's getSequenceNames()); the stream keeps std::fixed and the precision afterwards, as the
// reference's does.
template <class Names>
void printDistanceMatrix(std::ostream& out, const Names& names, distType* const* matrix, int n) {
  out << "\t" << n << std::endl;
  for (int i = 0; i < n; i++) {
    out << names[i] << "\t";
    for (int j = 0; j < n; j++) out << std::setprecision(6) << std::fixed << matrix[i][j] << " ";
    out << std::endl;
  }
}

template <class DiskMatrix>
int write_row_to_disk(void* user, unsigned row, const float* values, unsigned n) {
  static_cast<DiskMatrix*>(user)->writeArray(const_cast<float*>(values), (int)row, (int)n);
  return 0;
}

// Shared implementation of both calculators; `model` is RNJ_JC or RNJ_K2P.
template <class Loader, class DiskMatrix>
class DistanceCalculator {
 public:
  DistanceCalculator(int model, bool verbose, bool fastdist, Loader* loader, distType** rows,
                     DiskMatrix* dm)
      : model_(model), verbose_(verbose), fastdist_(fastdist), loader_(loader), matrix_(rows),
        dm_(dm), max_(0.0f) {}

  // Rows the calculator allocates itself become the lower triangle with the diagonal
  // (row i: i + 1 values), which is all rapidNJMem reads; distMatrixReader::initializeData
  // with halfMatrix copies exactly those (distMatrixReader.cpp:26-35). Half the transfer.
  void set_lower_rows(bool lower) { lower_rows_ = lower; }

  void compute() {
    if (matrix_ != nullptr && dm_ != nullptr) {
      std::cerr << "ERROR: computation of bootstrap values is not supported for trees which use "
                   "I/O efficient computation."
                << std::endl;
      std::exit(1);
    }
    const unsigned n = loader_->getSequenceCount();
    if (verbose_)
      std::cerr << "rapidnj_amd: " << (model_ == RNJ_K2P ? "Kimura" : "JC") << " distances for "
                << n << " sequences on the GPU" << std::endl;
    rnj_alignment* aln = upload_alignment(loader_, fastdist_);
    int rc;
    if (dm_ == nullptr) {
      if (matrix_ == nullptr && lower_rows_) {
        matrix_ = new distType*[n];
        for (unsigned i = 0; i < n; ++i) matrix_[i] = new distType[i + 1];
        rc = rnj_distances_rows_lower(shared_context(), aln, model_, matrix_, &max_);
      } else {
        if (matrix_ == nullptr) {
          matrix_ = new distType*[n];
          for (unsigned i = 0; i < n; ++i) matrix_[i] = new distType[n];
        }
        rc = rnj_distances_rows(shared_context(), aln, model_, matrix_, &max_);
      }
    } else {
      rc = rnj_distances_stream_rows(shared_context(), aln, model_, &write_row_to_disk<DiskMatrix>,
                                     dm_, &max_);
    }
    rnj_alignment_free(aln);
    if (rc != RNJ_OK) fail_exit("distance matrix");
  }

  distType** matrix() { return matrix_; }
  distType max_distance() const { return max_; }

 private:
  int model_;
  bool verbose_, fastdist_;
  bool lower_rows_ = false;
  Loader* loader_;
  distType** matrix_;
  DiskMatrix* dm_;
  distType max_;
};

template <class Loader, class DiskMatrix>
class KimuraDistance {
 public:
  KimuraDistance(bool verbose, bool fastdist, Loader* loader, DiskMatrix* dm)
      : impl_(RNJ_K2P, verbose, fastdist, loader, nullptr, dm) {}
  KimuraDistance(bool verbose, bool fastdist, Loader* loader, distType** distMatrix, DiskMatrix* dm)
      : impl_(RNJ_K2P, verbose, fastdist, loader, distMatrix, dm) {}
  void computeDistances(int numThreads) {
    (void)numThreads;
    impl_.compute();
  }
  void computeDistancesGPU() { impl_.compute(); }
  // For callers that build rapidNJMem (halfMatrix): rows hold the lower triangle only.
  void useLowerTriangleRows(bool lower) { impl_.set_lower_rows(lower); }
  distType** getDistanceMatrix() { return impl_.matrix(); }
  distType getMaxDistance() const { return impl_.max_distance(); }

 private:
  DistanceCalculator<Loader, DiskMatrix> impl_;
};

template <class Loader, class DiskMatrix>
class JCdistance {
 public:
  JCdistance(bool verbose, bool fastdist, Loader* loader, DiskMatrix* dm)
      : impl_(RNJ_JC, verbose, fastdist, loader, nullptr, dm) {}
  void computeDistanceMatrix(int numThreads) {
    (void)numThreads;
    impl_.compute();
  }
  void useLowerTriangleRows(bool lower) { impl_.set_lower_rows(lower); }
  distType** getDistanceMatrix() { return impl_.matrix(); }
  distType getMaxDistance() const { return impl_.max_distance(); }

 private:
  DistanceCalculator<Loader, DiskMatrix> impl_;
};

}  // namespace rapidnj_amd
