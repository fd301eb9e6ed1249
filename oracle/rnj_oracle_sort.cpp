// std::sort's order of one rapidNJ row list -- TEST INFRASTRUCTURE ONLY, part of the oracle
// library (oracle/rnj_oracle_nj.c calls it).
//
// rapidNJ sorts each row's (partner id, distance) list with std::sort on the distance alone
// (cluster_pair.h:8-10; rapidNJ.cpp:153 for initial rows, :234 for new rows). std::sort is not
// stable, so among equal distances the order it leaves depends on the whole list in its
// construction order; it is reproduced here by calling std::sort itself on the same sequence.
#include <algorithm>
#include <vector>

namespace {
struct Entry {
  int id;          // partner's cluster id
  float distance;  // compared alone, as cluster_pair::operator<
  bool operator<(const Entry& o) const { return distance < o.distance; }
};
}  // namespace

// Sorts one row list (ids and distances in construction order) in place, as std::sort leaves it.
extern "C" void rnj_oracle_sort_row(int* id, float* dist, int len) {
  std::vector<Entry> row((size_t)len);
  for (int k = 0; k < len; ++k) row[(size_t)k] = Entry{id[k], dist[k]};
  std::sort(row.begin(), row.end());
  for (int k = 0; k < len; ++k) {
    id[k] = row[(size_t)k].id;
    dist[k] = row[(size_t)k].distance;
  }
}
