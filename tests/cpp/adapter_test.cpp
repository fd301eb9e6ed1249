// Drives the C++ host classes of include/rapidnj_amd_distance.hpp the way rapidNJ's main does
// (src/main.cpp:237-286): a loader with the reference dataloader's getters, an optional disk
// matrix, KimuraDistance / JCdistance, getDistanceMatrix. Test infrastructure for
// tests/test_gpu_adapter.py.
//
//   adapter_test IN OUT kim|jc mem|disk|reuse|half|tree|print fast|chars [protein]
// (tree: OUT is the Newick text of rapidnj_amd::NJFromAlignment on the loader; print: OUT is the
// text rapidnj_amd::printDistanceMatrix writes for the in-memory rows, names t0, t1, ..)
// IN:  u32 n, L, bitStringsCount; n*L chars; n*bsc*4 u32 bit strings; n*bsc*4 u32 gap filters
//      (protein: the bit strings hold the resolved characters, 16 per 4 words, as the reference's
//      loader stores them; the gap filters are unused)
// OUT: u32 n; n*n f32 matrix (row-major); f32 maximum distance
#include <cstdint>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "rapidnj_amd_distance.hpp"
#include "rapidnj_amd_nj.hpp"

using rapidnj_amd::distType;

// The getters of dataloader (dataloader.hpp:20-29), over arrays held in memory like
// dataloaderMemory (dataloaderMemory.hpp:8-76).
class MemoryLoader {
 public:
  unsigned n = 0, L = 0, bsc = 0;
  std::vector<char> chars;
  std::vector<uint32_t> bits, gaps;
  std::vector<char*> seqs;
  std::vector<unsigned int*> bit_rows, gap_rows;
  std::vector<std::string> names;
  int type = 0;  // InputType (stdinclude.h:26): DNA = 0, PROTEIN = 1

  unsigned int** getBitStrings() { return bit_rows.data(); }
  unsigned int** getGapFilters() { return gap_rows.data(); }
  unsigned int getSequenceCount() { return n; }
  unsigned int getSequenceLength() { return L; }
  unsigned int getBitStringsCount() { return bsc; }
  std::vector<std::string>* getSequenceNames() { return &names; }
  std::vector<char*>* getSequences() { return &seqs; }
};

// Row store with diskMatrix's write interface (diskMatrix.h:13).
class MemoryDiskMatrix {
 public:
  explicit MemoryDiskMatrix(int size) : size(size), data((size_t)size * size, -7.0f) {}
  void writeArray(distType* values, int row, int n) {
    std::memcpy(&data[(size_t)row * size], values, (size_t)n * sizeof(distType));
    ++writes;
  }
  int size;
  std::vector<distType> data;
  int writes = 0;
};

static bool read_input(const char* path, MemoryLoader& ld) {
  std::ifstream in(path, std::ios::binary);
  uint32_t hdr[3];
  if (!in.read(reinterpret_cast<char*>(hdr), sizeof hdr)) return false;
  ld.n = hdr[0];
  ld.L = hdr[1];
  ld.bsc = hdr[2];
  ld.chars.resize((size_t)ld.n * ld.L);
  ld.bits.resize((size_t)ld.n * ld.bsc * 4);
  ld.gaps.resize(ld.bits.size());
  in.read(ld.chars.data(), ld.chars.size());
  in.read(reinterpret_cast<char*>(ld.bits.data()), ld.bits.size() * 4);
  in.read(reinterpret_cast<char*>(ld.gaps.data()), ld.gaps.size() * 4);
  if (!in) return false;
  for (unsigned t = 0; t < ld.n; ++t) {
    ld.seqs.push_back(&ld.chars[(size_t)t * ld.L]);
    ld.bit_rows.push_back(&ld.bits[(size_t)t * ld.bsc * 4]);
    ld.gap_rows.push_back(&ld.gaps[(size_t)t * ld.bsc * 4]);
    ld.names.push_back("t" + std::to_string(t));
  }
  return true;
}

int main(int argc, char** argv) {
  if (argc < 6) {
    std::cerr << "usage: adapter_test IN OUT kim|jc mem|disk|reuse|half fast|chars [protein]"
              << std::endl;
    return 2;
  }
  MemoryLoader ld;
  if (!read_input(argv[1], ld)) {
    std::cerr << "bad input " << argv[1] << std::endl;
    return 2;
  }
  const std::string model = argv[3], mode = argv[4];
  const bool fast = std::string(argv[5]) == "fast";
  if (argc > 6 && std::string(argv[6]) == "protein") ld.type = 1;
  const unsigned n = ld.n;
  if (mode == "tree") {
    rapidnj_amd::NJFromAlignment<MemoryLoader> nj(fast, &ld, model == "kim" ? RNJ_K2P : RNJ_JC,
                                                  false);
    rapidnj_amd::Tree* tree = nj.run();
    std::ofstream o(argv[2]);
    tree->serialize_tree(o);
    delete tree;
    return o ? 0 : 4;
  }
  if (mode == "print") {
    distType** m = nullptr;
    if (model == "kim") {
      rapidnj_amd::KimuraDistance<MemoryLoader, MemoryDiskMatrix> alg(false, fast, &ld, nullptr);
      alg.computeDistances(8);
      m = alg.getDistanceMatrix();
    } else {
      rapidnj_amd::JCdistance<MemoryLoader, MemoryDiskMatrix> alg(false, fast, &ld, nullptr);
      alg.computeDistanceMatrix(8);
      m = alg.getDistanceMatrix();
    }
    std::ofstream o(argv[2]);
    rapidnj_amd::printDistanceMatrix(o, *ld.getSequenceNames(), m, (int)n);
    for (unsigned i = 0; i < n; ++i) delete[] m[i];
    delete[] m;
    return o ? 0 : 4;
  }
  std::vector<distType> out((size_t)n * n);
  distType maxd = 0;
  if (mode == "disk") {
    MemoryDiskMatrix dm((int)n);
    if (model == "kim") {
      rapidnj_amd::KimuraDistance<MemoryLoader, MemoryDiskMatrix> alg(false, fast, &ld, &dm);
      alg.computeDistances(8);
      maxd = alg.getMaxDistance();
    } else {
      rapidnj_amd::JCdistance<MemoryLoader, MemoryDiskMatrix> alg(false, fast, &ld, &dm);
      alg.computeDistanceMatrix(8);
      maxd = alg.getMaxDistance();
    }
    if (dm.writes != (int)n) {
      std::cerr << "expected " << n << " row writes, got " << dm.writes << std::endl;
      return 3;
    }
    out = dm.data;
  } else {
    distType** rows = nullptr;
    if (mode == "reuse") {  // caller-owned rows, as bootstrapTree passes them (src/main.cpp:519-532)
      rows = new distType*[n];
      for (unsigned i = 0; i < n; ++i) {
        rows[i] = new distType[n];
        for (unsigned j = 0; j < n; ++j) rows[i][j] = -9.0f;
      }
    }
    distType** m = nullptr;
    const bool half = mode == "half";  // lower-triangle rows, as rapidNJMem keeps them
    if (model == "kim") {
      auto* alg = rows ? new rapidnj_amd::KimuraDistance<MemoryLoader, MemoryDiskMatrix>(
                             false, fast, &ld, rows, nullptr)
                       : new rapidnj_amd::KimuraDistance<MemoryLoader, MemoryDiskMatrix>(
                             false, fast, &ld, nullptr);
      alg->useLowerTriangleRows(half);
      alg->computeDistances(8);
      m = alg->getDistanceMatrix();
      maxd = alg->getMaxDistance();
      delete alg;
    } else {
      auto* alg = new rapidnj_amd::JCdistance<MemoryLoader, MemoryDiskMatrix>(false, fast, &ld,
                                                                              nullptr);
      alg->useLowerTriangleRows(half);
      alg->computeDistanceMatrix(8);
      m = alg->getDistanceMatrix();
      maxd = alg->getMaxDistance();
      delete alg;
    }
    if (rows && m != rows) {
      std::cerr << "caller rows were not reused" << std::endl;
      return 3;
    }
    for (unsigned i = 0; i < n; ++i) {
      if (half) {  // rapidNJMem reads matrix[max(i,j)][min(i,j)] (rapidNJMem.hpp:62-68)
        for (unsigned j = 0; j < n; ++j) out[(size_t)i * n + j] = j <= i ? m[i][j] : m[j][i];
      } else {
        std::memcpy(&out[(size_t)i * n], m[i], n * sizeof(distType));
      }
    }
    for (unsigned i = 0; i < n; ++i) delete[] m[i];  // the NJ builders own the rows (rapidNJ.cpp:432-437)
    delete[] m;
  }
  std::ofstream o(argv[2], std::ios::binary);
  const uint32_t nn = n;
  o.write(reinterpret_cast<const char*>(&nn), 4);
  o.write(reinterpret_cast<const char*>(out.data()), out.size() * sizeof(distType));
  o.write(reinterpret_cast<const char*>(&maxd), sizeof maxd);
  return o ? 0 : 4;
}
