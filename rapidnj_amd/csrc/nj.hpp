// Device neighbour joining (nj.hip) and its host driver (nj.cpp). Not part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace rnj {

// Scalar state of one run, in device memory (rapidNJ's members, rapidNJ.h).
struct NjState {
  int n;
  int used;           // slots in use (slot order = cluster id order)
  int alive;          // live slots after a compaction
  int cluster_count;  // clusterCount (copies included)
  int current_id;     // next cluster id
  int min1, min2, cache1, cache2;  // findMin's result and row caches (rapidNJ.cpp:290-345)
  int new_row, obs_row;
  int step;           // joins made
  int negative_branches;
  int error;          // kNjFailed or kNjHalted; every kernel does nothing while it is set
  int forced;         // the next step takes f_* as its search result instead of searching
  int f_min1, f_min2, f_cache1, f_cache2;
  // Row skipping (nj.hip): the sum of each step's largest separation rise, and the criterion of
  // one live pair of the next search's start row (an upper bound of that row's answer; +inf when
  // none) with the start row.
  double drift;
  float start_bound;
  int start_row;  // matrix index; -1 when there is no bound
};
// A slot's separation is NaN once its cluster is gone, so that scans skip it without a flag.

// NjState::error values.
constexpr int kNjFailed = 1;  // an iteration found no pair or ran out of slots
constexpr int kNjHalted = 2;  // the step left its search to the host (see nj.hip)

// Device arrays of one run, passed by value to the kernels.
struct NjDevice {
  NjState* st;
  size_t ld;           // row stride of the slot matrices (= slot capacity)
  // by matrix index (n), rapidNJ's bookkeeping
  int* index_to_id;
  int* id_to_index;    // 2n
  int* red_count;      // redundantCount
  int* rtm;            // redundantToMasterMap
  int* red_head;       // redundantMap[i] as a linked list: head, next
  int* red_next;
  int* slot_of_index;  // slot of a live index, -1 otherwise
  // by slot (capacity ld)
  int* slot_index;
  int* alive;
  float* sepsum;       // separationsums
  float* sep;          // separations
  unsigned long long* row_key;  // each row's best: (criterion key << 32) | row index; ~0 = none
  int* row_p;          // its partner index
  unsigned* row_sigma; // its break bound (key): the search is the host's if a record meets a
                       // running minimum at or below it
  int* row_tie;        // 1 when several pairs share its best criterion and distance
  float* row_maxsep;   // its maxRowSeparation, as the scan found it
  float* smin;         // min over the row's live pairs of float(d - sep[partner]) at its last scan
  double* sref;        // NjState::drift at that scan
  int* shave;          // 1 once smin / sref hold a scan's values
  float* init_maxsep;  // maxRowSeparations of the first search (every smaller index)
  float* newcol;       // new cluster's distances (step)
  float* term;         // dist * redundantCount terms of the new separation sum
  int* map;            // compaction: live slots in order
  unsigned* key_by_index;  // each row's criterion key by matrix index (~0: no answer)
  int* newidx;         // matrix index of the cluster made at each step
  float* log_val;      // each step's new row, by slot: distance
  int* log_idx;        //   and partner matrix index (-1: no partner)
  // joins (polytree::addInternalNode arguments), n - 2 of them
  int* left;
  int* right;
  double* llen;
  double* rlen;
};

hipError_t nj_launch_row_sums(const float* M, size_t ldm, int n, float* sums, hipStream_t s);
// Lists pairs (i, j), j < i, with M[i][j] == 0 (up to cap; *count = all of them).
hipError_t nj_launch_zero_pairs(const float* M, size_t ldm, int n, int* pairs, int cap, int* count,
                                hipStream_t s);
// same[k] = 1 when the rows of pair k are identical.
hipError_t nj_launch_same_rows(const float* M, size_t ldm, int n, const int* pairs, int npairs,
                               int* same, hipStream_t s);
hipError_t nj_launch_gather(const float* M, size_t ldm, const int* vis, int v, float* S, size_t ld,
                            hipStream_t s);
// One iteration over slots < used_bound: the scan (after finishing the previous join's new
// cluster when `finish_previous`), then the step, which logs its new row at log_val/log_idx +
// log_off (used_bound entries).
hipError_t nj_launch_iteration(const NjDevice& D, float* S, int used_bound, bool finish_previous,
                               size_t log_off, hipStream_t s);
// The step alone (after the host settled its search).
hipError_t nj_launch_step(const NjDevice& D, float* S, size_t log_off, hipStream_t s);
// Finish the last join's new cluster alone (before a compaction).
hipError_t nj_launch_finish(const NjDevice& D, float* S, hipStream_t s);
// Rebuild the slot matrix into S2 from the live slots (at most alive_bound of them).
hipError_t nj_launch_compact(const NjDevice& D, const float* S, float* S2, int alive_bound,
                             hipStream_t s);

}  // namespace rnj
