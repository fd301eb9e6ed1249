// Drives the C++ tree classes of include/rapidnj_amd_nj.hpp the way rapidNJ's main does:
// a matrix reader with distMatrixReader's getters, rapidNJ(reader, n, negative_branches, pb),
// run(), serialize_tree. Test infrastructure for tests/test_gpu_nj.py.
//
//   nj_test MATRIX OUT [neg]        MATRIX = u32 n + n*n f32 (row-major); OUT = Newick text
#include <cstdint>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "rapidnj_amd_nj.hpp"

using rapidnj_amd::distType;

// distMatrixReader's getters over rows held in memory (distMatrixReader.cpp:14-23,155-161).
class MemoryReader {
 public:
  std::vector<std::vector<distType>> data;
  std::vector<distType*> rows;
  std::vector<std::string> names;
  distType** getMatrix() { return rows.data(); }
  std::vector<std::string>* getSequenceNames() { return &names; }
};

int main(int argc, char** argv) {
  if (argc < 3) {
    std::cerr << "usage: nj_test MATRIX OUT [neg]" << std::endl;
    return 2;
  }
  std::ifstream in(argv[1], std::ios::binary);
  uint32_t n = 0;
  in.read(reinterpret_cast<char*>(&n), 4);
  MemoryReader r;
  r.data.assign(n, std::vector<distType>(n));
  for (uint32_t i = 0; i < n; ++i) {
    in.read(reinterpret_cast<char*>(r.data[i].data()), n * sizeof(distType));
    r.rows.push_back(r.data[i].data());
    r.names.push_back("t" + std::to_string(i));
  }
  if (!in) {
    std::cerr << "bad matrix " << argv[1] << std::endl;
    return 2;
  }
  const bool neg = argc > 3 && std::string(argv[3]) == "neg";
  rapidnj_amd::rapidNJ<MemoryReader> nj(&r, (int)n, neg, nullptr);
  rapidnj_amd::Tree* tree = nj.run();
  std::ofstream o(argv[2]);
  tree->serialize_tree(o);
  delete tree;
  return o ? 0 : 4;
}
