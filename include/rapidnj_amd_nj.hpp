// C++ host classes over the C ABI (rapidnj_amd.h) with the shape of the reference's rapidNJ and
// polytree, so the tree-building call site keeps working:
//
//   reference (src/main.cpp)                       this header
//   rapidNJ* nj = new rapidNJ(reader, n, neg, pb)  rapidnj_amd::rapidNJ<distMatrixReader>(reader, n, neg, pb)
//   polytree* tree = nj->run()                     rapidnj_amd::Tree* tree = nj->run()
//   tree->serialize_tree(out)                      tree->serialize_tree(out)
//
// rapidNJ.h / polytree.h declare the originals. The reader provides getMatrix() (n rows of
// distType, the full square matrix: distMatrixReader with halfMatrix == false) and
// getSequenceNames() (distMatrixReader.cpp:155-161). The rows are uploaded to HBM once
// (rnj_nj_rows) and the joins run on the device; the reader's rows are not modified or freed
// (the reference's rapidNJ destructor deletes them: the caller keeps that ownership here).
//
// rapidnj_amd::NJFromAlignment builds the tree from a loader directly: distances and joins on the
// device, the matrix never in host memory.
#pragma once

#include <cstdlib>
#include <iostream>
#include <string>
#include <utility>
#include <vector>

#include "rapidnj_amd.h"
#include "rapidnj_amd_distance.hpp"

namespace rapidnj_amd {

// polytree's interface for the finished tree.
class Tree {
 public:
  Tree(rnj_tree* t, std::vector<std::string> names) : t_(t), names_(std::move(names)) {}
  ~Tree() { rnj_tree_free(t_); }
  Tree(const Tree&) = delete;
  Tree& operator=(const Tree&) = delete;

  // polytree::serialize_tree (polytree.cpp:402-458): Newick with "%.5g" branch lengths.
  void serialize_tree(std::ostream& out) const {
    std::vector<const char*> nm;
    for (const std::string& s : names_) nm.push_back(s.c_str());
    size_t len = 0;
    const char* const* names = nm.size() == rnj_tree_size(t_) ? nm.data() : nullptr;
    if (rnj_tree_newick(t_, names, nullptr, 0, &len) != RNJ_OK) fail_exit("newick");
    std::string text(len + 1, '\0');
    rnj_tree_newick(t_, names, &text[0], text.size(), &len);
    text.resize(len);
    out << text;
  }
  rnj_tree* handle() const { return t_; }

 private:
  rnj_tree* t_;
  std::vector<std::string> names_;
};

template <class Reader>
class rapidNJ {
 public:
  rapidNJ(Reader* reader, int matrixSize, bool negative_branches, void* pb = nullptr)
      : reader_(reader), n_(matrixSize), negative_branches_(negative_branches) {
    (void)pb;
  }

  Tree* run() {
    rnj_tree* t = nullptr;
    const int rc = rnj_nj_rows(shared_context(), reader_->getMatrix(), (unsigned)n_,
                               negative_branches_ ? 1 : 0, &t);
    if (rc != RNJ_OK) fail_exit("neighbour joining");
    std::vector<std::string> names;
    if (std::vector<std::string>* nm = reader_->getSequenceNames()) names = *nm;
    return new Tree(t, names);
  }

 private:
  Reader* reader_;
  int n_;
  bool negative_branches_;
};

// Loader -> tree without a host matrix: the loader's alignment packed into HBM (as the distance
// classes do), K2P (RNJ_K2P) or JC distances with the exact FP64 logs, then the joins.
template <class Loader>
class NJFromAlignment {
 public:
  NJFromAlignment(bool fastdist, Loader* loader, int model, bool negative_branches)
      : fastdist_(fastdist), loader_(loader), model_(model), negative_branches_(negative_branches) {}

  Tree* run() {
    rnj_alignment* aln = upload_alignment(loader_, fastdist_);
    rnj_tree* t = nullptr;
    const int rc = rnj_nj_alignment(shared_context(), aln, model_, negative_branches_ ? 1 : 0, &t);
    rnj_alignment_free(aln);
    if (rc != RNJ_OK) fail_exit("neighbour joining");
    std::vector<std::string> names;
    if (std::vector<std::string>* nm = loader_->getSequenceNames()) names = *nm;
    return new Tree(t, names);
  }

 private:
  bool fastdist_;
  Loader* loader_;
  int model_;
  bool negative_branches_;
};

}  // namespace rapidnj_amd
