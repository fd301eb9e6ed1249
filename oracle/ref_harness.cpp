// Command-line driver around the reference distance engine, built by `make -C oracle ref`
// from the reference's own sources (compiled in place, never copied). It is test
// infrastructure: it writes golden vectors for tests/golden/ and serves as the measured
// CPU baseline ("kind": "reference") of bench.py. It is never part of the product path.
//
// Entry points it drives, all in the reference tree:
//   dataloaderFasta::readData / getBitStrings / getGapFilters   dataloaderFasta.hpp:24-99,219-251
//   dataloaderStockholm::load (fastdist)                        dataloaderStockholm.hpp:22-111,183-236
//   bitDistanceGap::computeDistance                             bitDistanceGap.cpp:60-98
//   bitDistanceProtein::computeDistance                         bitDistanceProtein.cpp:31-61
//   simpleDistanceCalculator::computeDistance (the -f path)     simpleDistanceCalculator.cpp:8-16
//   KimuraDistance::computeDistances / getDistanceMatrix        KimuraDistance.cpp:75-99,397-399
//   JCdistance::computeDistanceMatrix / getDistanceMatrix       JCdistance.cpp:24-44,232-234
//   distMatrixReader + rapidNJ::run + polytree::serialize_tree  src/main.cpp:419-437,288-296
//
// Modes (all binary outputs are little-endian):
//   pack   FASTA OUT                  u32 n, L, bitStringsCount; then n*bsc*4 u32 bit strings,
//                                     then n*bsc*4 u32 gap filters
//   fasta  FASTA OUT                  u32 n, L, bitStringsCount, H; H names (u32 length + bytes,
//                                     every '>' record in file order); then bit strings and gap
//                                     filters as in `pack` (the reference exits 1 on bad input)
//   sth    STOCKHOLM OUT              the same for a (single-line) Stockholm alignment, with the
//                                     name of every sequence line
//   load   FASTA REPS                 prints {"best_s":..} of parse + pack (readData, fastdist)
//   sload  STOCKHOLM REPS             the same for the Stockholm reader
//   pcounts FASTA OUT [naive]         protein (bitDistanceProtein, or with naive the -f path's
//                                     simpleDistanceCalculator): u32 n, L; then for i<j
//                                     u64 data[0], data[1], data[2] as the estimator returns them
//   pmatrix FASTA jc|kim OUT [naive]  protein distances: u32 n; then n*n f32 (post-processed)
//   counts FASTA OUT [naive]          u32 n, L; then for i<j (row-major) u32 ts, tv, valid
//   matrix FASTA jc|kim T OUT [naive] u32 n; then n*n f32 (post-processed matrix); prints seconds
//   dmatrix FASTA jc|kim T DIR OUT [naive]
//                                     the disk-matrix path (KimuraDistance / JCdistance with a
//                                     diskMatrix in DIR, KimuraDistance.cpp:75-99,256-281): every
//                                     row computed with j from 0 and written with writeArray, then
//                                     read back with readArray; u32 n; then n*n f32
//   time   FASTA jc|kim T REPS        prints {"best_s":..,"mean_s":..,"n":..,"L":..,"threads":..}
//   ptime  FASTA jc|kim T REPS        the same for a protein alignment
//   nj     MATRIX OUT [neg]           MATRIX = u32 n + n*n f32; writes Newick (rapidNJ, negative_branches with neg) with
//                                     leaf names "t<i>"
//   tree   FASTA jc|kim T OUT         distance matrix + rapidNJ; writes Newick

#include "stdinclude.h"
#include "dataloaderFasta.hpp"
#include "dataloaderStockholm.hpp"
#include "bitDistanceGap.hpp"
#include "bitDistanceProtein.hpp"
#include "simpleDistanceCalculator.hpp"
#include "KimuraDistance.hpp"
#include "JCdistance.hpp"
#include "distMatrixReader.hpp"
#include "rapidNJ.h"
#include "polytree.h"
#include "ProgressBar.hpp"
#include "diskMatrix.h"

#include <chrono>
#include <cstdint>
#include <fstream>
#include <sstream>

namespace {

void die(const std::string& msg) {
  std::cerr << "ref_harness: " << msg << std::endl;
  std::exit(2);
}

dataloaderFasta* load(const std::string& path, bool fastdist) {
  dataloaderFasta* dl = new dataloaderFasta();
  dl->readData(path, fastdist, false, UNKNOWN, false);
  if (dl->type != DNA) die("input is not detected as DNA");
  return dl;
}

void put_u32(std::ofstream& o, uint32_t v) { o.write(reinterpret_cast<const char*>(&v), 4); }

std::ofstream open_out(const std::string& path) {
  std::ofstream o(path.c_str(), std::ios::binary);
  if (!o) die("cannot write " + path);
  return o;
}

distType** compute_matrix(dataloader* dl, const std::string& model, int threads, bool fastdist,
                          double* seconds) {
  auto t0 = std::chrono::steady_clock::now();
  distType** m = NULL;
  if (model == "kim") {
    KimuraDistance* alg = new KimuraDistance(false, fastdist, dl, NULL);
    alg->computeDistances(threads);
    m = alg->getDistanceMatrix();
    delete alg;
  } else if (model == "jc") {
    JCdistance* alg = new JCdistance(false, fastdist, dl, NULL);
    alg->computeDistanceMatrix(threads);
    m = alg->getDistanceMatrix();
    delete alg;
  } else {
    die("model must be jc or kim");
  }
  auto t1 = std::chrono::steady_clock::now();
  if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
  return m;
}

void free_matrix(distType** m, int n) {
  for (int i = 0; i < n; i++) delete[] m[i];
  delete[] m;
}

int mode_pack(int argc, char** argv) {
  if (argc < 4) die("pack FASTA OUT");
  dataloaderFasta* dl = load(argv[2], true);
  uint32_t n = dl->getSequenceCount(), L = dl->getSequenceLength(), bsc = dl->getBitStringsCount();
  std::ofstream o = open_out(argv[3]);
  put_u32(o, n); put_u32(o, L); put_u32(o, bsc);
  for (uint32_t i = 0; i < n; i++) o.write(reinterpret_cast<const char*>(dl->getBitStrings()[i]), bsc * 16);
  for (uint32_t i = 0; i < n; i++) o.write(reinterpret_cast<const char*>(dl->getGapFilters()[i]), bsc * 16);
  delete dl;
  return 0;
}

int mode_fasta(int argc, char** argv) {
  if (argc < 4) die("fasta FASTA OUT");
  dataloaderFasta* dl = new dataloaderFasta();
  dl->readData(argv[2], true, false, UNKNOWN, false);  // exits 1 itself on malformed input
  if (dl->type != DNA) die("input is not detected as DNA");
  uint32_t n = dl->getSequenceCount(), L = dl->getSequenceLength(), bsc = dl->getBitStringsCount();
  vector<string>* names = dl->getSequenceNames();
  std::ofstream o = open_out(argv[3]);
  put_u32(o, n); put_u32(o, L); put_u32(o, bsc); put_u32(o, (uint32_t)names->size());
  for (const string& s : *names) {
    put_u32(o, (uint32_t)s.size());
    o.write(s.data(), s.size());
  }
  for (uint32_t i = 0; i < n; i++) o.write(reinterpret_cast<const char*>(dl->getBitStrings()[i]), bsc * 16);
  for (uint32_t i = 0; i < n; i++) o.write(reinterpret_cast<const char*>(dl->getGapFilters()[i]), bsc * 16);
  delete dl;
  return 0;
}

int mode_sth(int argc, char** argv) {
  if (argc < 4) die("sth STOCKHOLM OUT");
  dataloaderStockholm* dl = new dataloaderStockholm();
  dl->readData(argv[2], true, false, UNKNOWN, false);  // exits 1 itself on malformed input
  if (dl->type != DNA) die("input is not detected as DNA");
  uint32_t n = dl->getSequenceCount(), L = dl->getSequenceLength(), bsc = dl->getBitStringsCount();
  vector<string>* names = dl->getSequenceNames();
  std::ofstream o = open_out(argv[3]);
  put_u32(o, n); put_u32(o, L); put_u32(o, bsc); put_u32(o, (uint32_t)names->size());
  for (const string& s : *names) {
    put_u32(o, (uint32_t)s.size());
    o.write(s.data(), s.size());
  }
  for (uint32_t i = 0; i < n; i++) o.write(reinterpret_cast<const char*>(dl->getBitStrings()[i]), bsc * 16);
  for (uint32_t i = 0; i < n; i++) o.write(reinterpret_cast<const char*>(dl->getGapFilters()[i]), bsc * 16);
  return 0;
}

// load / sload: best wall time of the reference reader (parse + pack, fastdist) over REPS runs.
template <class Loader>
int mode_load(int argc, char** argv) {
  if (argc < 4) die("load|sload FILE REPS");
  int reps = atoi(argv[3]);
  double best = 1e30;
  uint32_t n = 0, L = 0;
  for (int r = 0; r < reps; r++) {
    auto t0 = std::chrono::steady_clock::now();
    Loader* dl = new Loader();
    dl->readData(argv[2], true, false, UNKNOWN, false);
    auto t1 = std::chrono::steady_clock::now();
    best = std::min(best, std::chrono::duration<double>(t1 - t0).count());
    n = dl->getSequenceCount();
    L = dl->getSequenceLength();
    delete dl;
  }
  std::cout << "{\"best_s\": " << best << ", \"n\": " << n << ", \"L\": " << L << "}" << std::endl;
  return 0;
}

dataloaderFasta* load_protein(const std::string& path, bool fastdist = true) {
  dataloaderFasta* dl = new dataloaderFasta();
  dl->readData(path, fastdist, false, UNKNOWN, false);
  if (dl->type != PROTEIN) die("input is not detected as protein");
  return dl;
}

int mode_pcounts(int argc, char** argv) {
  if (argc < 4) die("pcounts FASTA OUT [naive]");
  bool naive = argc > 4 && std::string(argv[4]) == "naive";
  dataloaderFasta* dl = load_protein(argv[2], !naive);
  DistanceEstimate* est = naive ? static_cast<DistanceEstimate*>(new simpleDistanceCalculator(dl))
                                : static_cast<DistanceEstimate*>(new bitDistanceProtein(dl));
  uint32_t n = dl->getSequenceCount(), L = dl->getSequenceLength();
  std::ofstream o = open_out(argv[3]);
  put_u32(o, n); put_u32(o, L);
  unsigned long long d[3];
  for (uint32_t i = 0; i < n; i++)
    for (uint32_t j = i + 1; j < n; j++) {
      est->computeDistance(i, j, d);
      o.write(reinterpret_cast<const char*>(d), sizeof d);
    }
  delete est;
  delete dl;
  return 0;
}

int mode_pmatrix(int argc, char** argv) {
  if (argc < 5) die("pmatrix FASTA jc|kim OUT [naive]");
  bool naive = argc > 5 && std::string(argv[5]) == "naive";
  dataloaderFasta* dl = load_protein(argv[2], !naive);
  int n = dl->getSequenceCount();
  distType** m = compute_matrix(dl, argv[3], 4, !naive, NULL);
  std::ofstream o = open_out(argv[4]);
  put_u32(o, (uint32_t)n);
  for (int i = 0; i < n; i++) o.write(reinterpret_cast<const char*>(m[i]), sizeof(distType) * n);
  free_matrix(m, n);
  delete dl;
  return 0;
}

int mode_counts(int argc, char** argv) {
  if (argc < 4) die("counts FASTA OUT [naive]");
  bool naive = argc > 4 && std::string(argv[4]) == "naive";
  dataloaderFasta* dl = load(argv[2], !naive);
  DistanceEstimate* est = naive ? static_cast<DistanceEstimate*>(new simpleDistanceCalculator(dl))
                                : static_cast<DistanceEstimate*>(new bitDistanceGap(dl));
  uint32_t n = dl->getSequenceCount(), L = dl->getSequenceLength();
  std::ofstream o = open_out(argv[3]);
  put_u32(o, n); put_u32(o, L);
  unsigned long long d[3];
  for (uint32_t i = 0; i < n; i++)
    for (uint32_t j = i + 1; j < n; j++) {
      est->computeDistance(i, j, d);
      put_u32(o, (uint32_t)d[0]); put_u32(o, (uint32_t)d[1]); put_u32(o, (uint32_t)d[2]);
    }
  delete est;
  delete dl;
  return 0;
}

int mode_matrix(int argc, char** argv) {
  if (argc < 6) die("matrix FASTA jc|kim T OUT [naive]");
  bool naive = argc > 6 && std::string(argv[6]) == "naive";
  dataloaderFasta* dl = load(argv[2], !naive);
  int n = dl->getSequenceCount();
  double secs = 0;
  distType** m = compute_matrix(dl, argv[3], atoi(argv[4]), !naive, &secs);
  std::ofstream o = open_out(argv[5]);
  put_u32(o, (uint32_t)n);
  for (int i = 0; i < n; i++) o.write(reinterpret_cast<const char*>(m[i]), sizeof(distType) * n);
  std::cout << "seconds=" << secs << std::endl;
  free_matrix(m, n);
  delete dl;
  return 0;
}

int mode_dmatrix(int argc, char** argv) {
  if (argc < 7) die("dmatrix FASTA jc|kim T DIR OUT [naive]");
  bool naive = argc > 7 && std::string(argv[7]) == "naive";
  dataloaderFasta* dl = load(argv[2], !naive);
  int n = dl->getSequenceCount(), threads = atoi(argv[4]);
  diskMatrix* dm = new diskMatrix(argv[5], n);
  std::string model = argv[3];
  if (model == "kim") {
    KimuraDistance* alg = new KimuraDistance(false, !naive, dl, dm);
    alg->computeDistances(threads);
    delete alg;
  } else if (model == "jc") {
    JCdistance* alg = new JCdistance(false, !naive, dl, dm);
    alg->computeDistanceMatrix(threads);
    delete alg;
  } else {
    die("model must be jc or kim");
  }
  dm->flush();
  std::ofstream o = open_out(argv[6]);
  put_u32(o, (uint32_t)n);
  std::vector<distType> row(n);
  for (int i = 0; i < n; i++) {
    dm->readArray(row.data(), i, n);
    o.write(reinterpret_cast<const char*>(row.data()), sizeof(distType) * n);
  }
  delete dm;
  delete dl;
  return 0;
}

int mode_time(int argc, char** argv) {
  if (argc < 6) die(std::string(argv[1]) + " FASTA jc|kim T REPS");
  dataloaderFasta* dl = std::string(argv[1]) == "ptime" ? load_protein(argv[2]) : load(argv[2], true);
  int n = dl->getSequenceCount();
  int threads = atoi(argv[4]), reps = atoi(argv[5]);
  double best = 1e30, total = 0;
  for (int r = 0; r < reps; r++) {
    double s = 0;
    distType** m = compute_matrix(dl, argv[3], threads, true, &s);
    free_matrix(m, n);
    best = std::min(best, s);
    total += s;
  }
  std::cout << "{\"best_s\": " << best << ", \"mean_s\": " << total / reps << ", \"n\": " << n
            << ", \"L\": " << dl->getSequenceLength() << ", \"threads\": " << threads << "}" << std::endl;
  delete dl;
  return 0;
}

void run_nj(distType** m, int n, std::ostream& out, bool negative_branches = false) {
  vector<string>* names = new vector<string>;
  for (int i = 0; i < n; i++) {
    std::ostringstream s;
    s << "t" << i;
    names->push_back(s.str());
  }
  ProgressBar* pb = new ProgressBar();
  distMatrixReader* reader = new distMatrixReader(false, n, false, names, m);
  reader->initializeData();
  rapidNJ* nj = new rapidNJ(reader, n, negative_branches, pb);
  polytree* tree = nj->run();
  tree->serialize_tree(out);
  out.flush();
}

int mode_nj(int argc, char** argv) {
  if (argc < 4) die("nj MATRIX OUT [neg]");
  std::ifstream in(argv[2], std::ios::binary);
  if (!in) die(std::string("cannot read ") + argv[2]);
  uint32_t n = 0;
  in.read(reinterpret_cast<char*>(&n), 4);
  distType** m = new distType*[n];
  for (uint32_t i = 0; i < n; i++) {
    m[i] = new distType[n];
    in.read(reinterpret_cast<char*>(m[i]), sizeof(distType) * n);
  }
  if (!in) die("short matrix file");
  std::ofstream o(argv[3]);
  // rapidNJ's destructor path owns the rows; the process exits next
  run_nj(m, (int)n, o, argc > 4 && std::string(argv[4]) == "neg");
  return 0;
}

int mode_tree(int argc, char** argv) {
  if (argc < 6) die("tree FASTA jc|kim T OUT");
  dataloaderFasta* dl = load(argv[2], true);
  int n = dl->getSequenceCount();
  distType** m = compute_matrix(dl, argv[3], atoi(argv[4]), true, NULL);
  std::ofstream o(argv[5]);
  run_nj(m, n, o);
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 2) die("usage: ref_harness pack|fasta|sth|load|sload|pcounts|pmatrix|counts|matrix|dmatrix|time|ptime|nj|tree ...");
  std::string mode = argv[1];
  if (mode == "pack") return mode_pack(argc, argv);
  if (mode == "fasta") return mode_fasta(argc, argv);
  if (mode == "sth") return mode_sth(argc, argv);
  if (mode == "load") return mode_load<dataloaderFasta>(argc, argv);
  if (mode == "sload") return mode_load<dataloaderStockholm>(argc, argv);
  if (mode == "pcounts") return mode_pcounts(argc, argv);
  if (mode == "pmatrix") return mode_pmatrix(argc, argv);
  if (mode == "counts") return mode_counts(argc, argv);
  if (mode == "matrix") return mode_matrix(argc, argv);
  if (mode == "dmatrix") return mode_dmatrix(argc, argv);
  if (mode == "time" || mode == "ptime") return mode_time(argc, argv);
  if (mode == "nj") return mode_nj(argc, argv);
  if (mode == "tree") return mode_tree(argc, argv);
  die("unknown mode " + mode);
  return 2;
}
