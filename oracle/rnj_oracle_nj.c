/* CPU restatement of rapidNJ's neighbour joining (rapidNJ.cpp) and its Newick writer
 * (polytree.cpp). TEST INFRASTRUCTURE ONLY, like rnj_oracle.c: only tests/ and bench.py's CPU leg
 * load it, as the checker.
 *
 * The search is restated with rapidNJ's own data: every row keeps the list of its partners (the
 * live clusters of smaller id when it was built), sorted by distance alone with std::sort
 * (cluster_pair.h:8; rapidNJ.cpp:141-153 for initial rows, :219-234 for new ones; the sort is
 * rnj_oracle_sort.cpp's, std::sort itself, whose order among equal distances is not the
 * construction order). findMin (rapidNJ.cpp:290-361) visits rows by index, cyclically from the
 * row its previous search cached, offers a row's redundant-copy join first (:413-430), and scans
 * the row's live pairs in list order until the first one whose
 * float(distance - (sep[row] + maxRowSeparation[row])) is not below the running minimum (:340).
 * The criterion is the float (d - sep[partner]) - sep[row] (:345). Both are kept to the bit:
 * the break's rounding differs from the criterion's, so it can end a row one pair before a
 * criterion one unit in the last place below the running minimum, and the reference then joins
 * another pair. maxRowSeparations follow rapidNJ.cpp:114-119 (every smaller index, copies
 * included) and :177-214 (the live clusters of smaller id, after each update).
 * Merging (rapidNJ.cpp:240-281), the update of distances and separations (rapidNJ.cpp:166-238, in
 * id order, with float arithmetic and the sequential float sum of the new separation sum),
 * redundancy detection (rapidNJ.cpp:91-138,155-164) and the last join (rapidNJ.cpp:52-85) are
 * restated step for step. Garbage collection of dead pairs (rapidNJ.cpp:363-410) keeps the list
 * order and is left out. Pinned against the reference's own trees (tests/test_nj_oracle.py,
 * tests/golden/nj.npz written by the compiled reference).
 */
#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "rnj_oracle.h"

typedef struct {
  unsigned n;
  float* m;  /* n x n by index, updated in place */
  float* sepsum;
  float* sep;
  float* maxrow;  /* maxRowSeparations by index */
  int *index_to_id, *id_to_index, *red_count, *rtm, *red_head, *red_next;
  int cluster_count, current_id;
  int min1, min2, cache1, cache2;
  float global_min;
  int new_row, obs_row;
  int negative_branches;
  int* list_id;    /* n x n: the sorted list of the row at each index (partner ids) */
  float* list_d;   /* and their distances */
  int* list_len;
} nj_state;

void rnj_oracle_sort_row(int* id, float* dist, int len);

/* ---- audit of the device's row-skipping rule (rapidnj_amd/csrc/nj.hip) ---------------------
 * Not part of rapidNJ: with an audit attached, each search first applies the device's rule to
 * every live row (a row's partners are the live clusters of smaller id) and checks its two
 * claims along rapidNJ's own trajectory: (1) a row's bound (the smallest float(d - sep) of its
 * last evaluation, less the drift of separations since, less its own separation and a 1e-5
 * relative margin) is at most every criterion the row has now; (2) a row the rule skips (bound
 * at least U, the criterion of the start row's pair that was best at its last evaluation, or
 * its copy join) never lowers the running minimum in the search (is never a record). */
typedef struct {
  long long bound_violations, record_violations, skipped, scanned;
  int no_drift; /* a planted mistake for the audit's own test: the drift left out */
  double drift;
  float* smin;
  double* sref;
  int *have, *rowid, *bestid, *skip, *prev_id;
  float* prev_sep;
} nj_audit;

static nj_audit* g_audit = NULL;

static float crit(const nj_state* s, int i, int j) {  /* the search's criterion of pair (i, j) */
  const float d = s->m[(size_t)i * s->n + j];
  float v = d - s->sep[s->rtm[j]];
  return v - s->sep[i];
}

static void audit_before_search(nj_state* s, int start) {
  nj_audit* A = g_audit;
  const unsigned n = s->n;
  /* the drift: the largest rise of a live cluster's separation since the previous search */
  float rise = 0;
  for (unsigned i = 0; i < n; i++) {
    const int id = s->index_to_id[i];
    if (id == -1) continue;
    if (A->prev_id[i] == id) {
      const float r = s->sep[i] - A->prev_sep[i];
      if (r > rise) rise = r;
    }
  }
  A->drift += rise;
  for (unsigned i = 0; i < n; i++) {
    A->prev_id[i] = s->index_to_id[i];
    A->prev_sep[i] = s->sep[i];
  }
  /* U: the start row's pair that was best at its last evaluation, if still a live partner */
  float U = FLT_MAX;
  if (start >= 0 && s->index_to_id[start] != -1 && A->have[start] &&
      A->rowid[start] == s->index_to_id[start]) {
    const int b = A->bestid[start] >= 0 ? s->id_to_index[A->bestid[start]] : -1;
    if (b >= 0 && s->index_to_id[b] != -1 && s->index_to_id[b] < s->index_to_id[start])
      U = crit(s, start, b);
    if (s->red_count[start] > 1 && -2 * s->sep[start] < U) U = -2 * s->sep[start];
  }
  for (unsigned i = 0; i < n; i++) {
    A->skip[i] = 0;
    const int id = s->index_to_id[i];
    if (id == -1) continue;
    const int known = A->have[i] && A->rowid[i] == id;
    float lb = -FLT_MAX;
    if (known) {
      const float m = A->smin[i];
      const float dr = A->no_drift ? 0.0f : (float)(A->drift - A->sref[i]);
      lb = (m - dr) - s->sep[i];
      lb = lb - 1e-5f * (1.0f + fabsf(m) + dr + fabsf(s->sep[i]));
      for (unsigned j = 0; j < n; j++) {
        const int jd = s->index_to_id[j];
        if (j == i || jd == -1 || jd > id) continue;
        if (crit(s, (int)i, (int)j) < lb) A->bound_violations++;
      }
    }
    if (known && (int)i != start && s->red_count[i] <= 1 && U < FLT_MAX && lb >= U) {
      A->skip[i] = 1;
      A->skipped++;
      continue;
    }
    /* evaluated: a fresh bound and best pair */
    A->scanned++;
    float m = INFINITY, bv = INFINITY;
    int bid = -1;
    for (unsigned j = 0; j < n; j++) {
      const int jd = s->index_to_id[j];
      if (j == i || jd == -1 || jd > id) continue;
      const float t = s->m[(size_t)i * n + j] - s->sep[s->rtm[j]];
      if (t < m) m = t;
      const float v = t - s->sep[i];
      if (v < bv) {
        bv = v;
        bid = jd;
      }
    }
    A->smin[i] = m;
    A->sref[i] = A->drift;
    A->have[i] = 1;
    A->rowid[i] = id;
    A->bestid[i] = bid;
  }
}

/* redundantMap[i] as a singly linked list: push_front / front / pop_front */
static void red_push(nj_state* s, int i, int j) {
  s->red_next[j] = s->red_head[i];
  s->red_head[i] = j;
}
static int red_pop(nj_state* s, int i) {
  const int j = s->red_head[i];
  s->red_head[i] = s->red_next[j];
  return j;
}

/* rapidNJ.cpp:155-164 */
static void handle_redundancy(nj_state* s, int i, int j) {
  const unsigned n = s->n;
  for (unsigned k = 0; k < n; k++)
    if (s->m[(size_t)i * n + k] != s->m[(size_t)j * n + k]) return;
  s->index_to_id[j] = -1;
  s->rtm[j] = i;
  red_push(s, i, j);
  s->red_count[i]++;
}

/* rapidNJ.cpp:91-155 */
static void initialize(nj_state* s) {
  const unsigned n = s->n;
  s->cluster_count = (int)n;
  s->current_id = (int)n;
  s->cache1 = s->cache2 = 0;
  s->min1 = s->min2 = 0;
  float max_sep = FLT_MIN;
  for (unsigned i = 0; i < n; i++) {
    s->index_to_id[i] = (int)i;
    s->id_to_index[i] = (int)i;
    s->red_count[i] = 1;
    s->rtm[i] = (int)i;
    s->red_head[i] = -1;
    float sum = 0;
    for (unsigned j = 0; j < n; j++) sum += s->m[(size_t)i * n + j];
    s->sepsum[i] = sum;
    s->sep[i] = sum / (float)(s->cluster_count - 2);
    s->maxrow[i] = max_sep;
    if (s->sep[i] > max_sep) max_sep = s->sep[i];
  }
  for (int i = (int)n - 1; i >= 0; i--) {
    if (s->index_to_id[i] == -1) continue;
    for (int j = 0; j < i; j++)
      if (s->m[(size_t)i * n + j] == 0) handle_redundancy(s, i, j);
  }
  /* buildSortedRow: partners j < i still visible, in index order, then sorted */
  for (unsigned i = 0; i < n; i++) {
    s->list_len[i] = 0;
    if (s->index_to_id[i] == -1) continue;
    int* ids = s->list_id + (size_t)i * n;
    float* ds = s->list_d + (size_t)i * n;
    int len = 0;
    for (unsigned j = 0; j < i; j++) {
      if (s->index_to_id[j] == -1) continue;
      ids[len] = (int)j;
      ds[len++] = s->m[(size_t)i * n + j];
    }
    rnj_oracle_sort_row(ids, ds, len);
    s->list_len[i] = len;
  }
}

/* a new minimum: rapidNJ.cpp:337-345 */
static void improve(nj_state* s, int i, int j, float value) {
  if (g_audit && g_audit->skip[i]) g_audit->record_violations++;
  if (i != s->min1) {
    s->cache2 = s->cache1;
    s->cache1 = s->min1;
  }
  s->global_min = value;
  s->min1 = i;
  s->min2 = j;
}

/* rapidNJ.cpp:290-361, 413-430 */
static void find_min(nj_state* s) {
  const unsigned n = s->n;
  s->global_min = FLT_MAX;
  const int start = s->index_to_id[s->cache1] != -1 ? s->cache1 : s->cache2;
  if (g_audit) audit_before_search(s, start);
  int i = start;
  do {
    if (s->index_to_id[i] != -1) {
      const float row_sep = s->sep[i];
      const float row_max_sep = row_sep + s->maxrow[i];
      if (s->red_count[i] > 1) {
        const float value = -2 * s->sep[i];
        if (value < s->global_min) improve(s, i, s->red_head[i], value);
      }
      const int* ids = s->list_id + (size_t)i * n;
      const float* ds = s->list_d + (size_t)i * n;
      for (int k = 0; k < s->list_len[i]; k++) {
        const int o = s->id_to_index[ids[k]];
        if (o == -1) continue; /* a dead pair */
        const float d = ds[k];
        const float bound = d - row_max_sep;
        if (bound >= s->global_min) break;
        float v = d - s->sep[s->rtm[o]];
        v = v - row_sep;
        if (v < s->global_min) improve(s, i, o, v);
      }
    }
    i = (i + 1) % (int)n;
  } while (i != start);
}

/* rapidNJ.cpp:240-281; the join goes to the caller's arrays */
static void merge_min_nodes(nj_state* s, int* left, int* right, double* llen, double* rlen,
                            int step) {
  const unsigned n = s->n;
  const int r1 = s->rtm[s->min1], r2 = s->rtm[s->min2];
  const double dist = s->m[(size_t)r1 * n + r2];
  const double sep1 = s->sep[r1];
  const double sep2 = s->sep[r2];
  double dist1 = (0.5 * dist) + (0.5 * (sep1 - sep2));
  double dist2 = (0.5 * dist) + (0.5 * (sep2 - sep1));
  if (s->negative_branches) {
    if (dist1 < 0.0) {
      dist2 += dist1;
      dist1 = 0;
    }
    if (dist2 < 0.0) {
      dist1 += dist2;
      if (dist1 < 0) dist1 = 0;
      dist2 = 0;
    }
  }
  s->new_row = r1;
  s->obs_row = r2;
  if (s->red_count[r1] != 1) {
    s->new_row = red_pop(s, r1);
    s->red_count[r1]--;
    s->index_to_id[s->new_row] = s->new_row;
  }
  if (s->red_count[s->rtm[s->min2]] != 1) {
    const int m2 = s->rtm[s->min2];
    s->obs_row = red_pop(s, m2);
    s->red_count[m2]--;
    s->index_to_id[s->obs_row] = s->obs_row;
  }
  left[step] = s->index_to_id[s->new_row];
  right[step] = s->index_to_id[s->obs_row];
  llen[step] = dist1;
  rlen[step] = dist2;
}

/* rapidNJ.cpp:166-238 */
static void update_data(nj_state* s) {
  const unsigned n = s->n;
  float new_sepsum = 0;
  const int r1 = s->rtm[s->min1], r2 = s->rtm[s->min2];
  const float mutual = s->m[(size_t)r1 * n + r2];
  const float* row1 = s->m + (size_t)r1 * n;
  const float* row2 = s->m + (size_t)r2 * n;
  float* new_row = s->m + (size_t)s->new_row * n;
  float max_sep = FLT_MIN;
  for (int cur = 0; cur < s->current_id; cur++) {
    const int i = s->id_to_index[cur];
    if (i == -1 || s->index_to_id[i] == -1) continue;
    if (i == s->new_row || i == s->obs_row) {
      new_row[i] = 0;
    } else {
      const float val1 = row1[i];
      const float val2 = row2[i];
      const float t = val1 + val2 - mutual;
      float dist = (float)((double)t / 2.0);
      if (s->negative_branches && dist < 0) dist = 0;
      new_sepsum += dist * (float)s->red_count[i];
      s->sepsum[i] += (dist - val1 - val2);
      s->sep[i] = s->sepsum[i] / (float)(s->cluster_count - 2);
      s->maxrow[i] = max_sep;
      if (s->sep[i] > max_sep) max_sep = s->sep[i];
      new_row[i] = dist;
      s->m[(size_t)i * n + s->new_row] = dist;
    }
  }
  s->id_to_index[s->index_to_id[s->new_row]] = -1;
  s->id_to_index[s->index_to_id[s->obs_row]] = -1;
  s->index_to_id[s->obs_row] = -1;
  s->sepsum[s->new_row] = new_sepsum;
  s->sepsum[s->obs_row] = 0;
  s->list_len[s->obs_row] = 0;
  s->sep[s->new_row] = new_sepsum / (float)(s->cluster_count - 2);
  s->maxrow[s->new_row] = max_sep;
  s->sep[s->obs_row] = FLT_MIN;
  s->rtm[s->new_row] = s->new_row;
  s->rtm[s->obs_row] = s->obs_row;
  const int next_id = s->current_id++;
  s->index_to_id[s->new_row] = next_id;
  s->id_to_index[next_id] = s->new_row;
  /* the new cluster's list: every other live index, in index order, then sorted */
  int* ids = s->list_id + (size_t)s->new_row * n;
  float* ds = s->list_d + (size_t)s->new_row * n;
  int len = 0;
  for (unsigned i = 0; i < n; i++) {
    if ((int)i == s->new_row || s->index_to_id[i] == -1) continue;
    ids[len] = s->index_to_id[i];
    ds[len++] = new_row[i];
  }
  rnj_oracle_sort_row(ids, ds, len);
  s->list_len[s->new_row] = len;
}

int rnj_oracle_nj(const float* matrix, unsigned n, int negative_branches, int* left, int* right,
                  double* llen, double* rlen, int* last, float* last_dist) {
  if (n < 2) return -1;
  nj_state s;
  memset(&s, 0, sizeof s);
  s.n = n;
  s.negative_branches = negative_branches;
  s.m = (float*)malloc((size_t)n * n * sizeof(float));
  s.sepsum = (float*)malloc(n * sizeof(float));
  s.sep = (float*)malloc(n * sizeof(float));
  s.index_to_id = (int*)malloc(n * sizeof(int));
  s.id_to_index = (int*)malloc(2 * (size_t)n * sizeof(int));
  s.red_count = (int*)malloc(n * sizeof(int));
  s.rtm = (int*)malloc(n * sizeof(int));
  s.red_head = (int*)malloc(n * sizeof(int));
  s.red_next = (int*)malloc(n * sizeof(int));
  s.maxrow = (float*)malloc(n * sizeof(float));
  s.list_id = (int*)malloc((size_t)n * n * sizeof(int));
  s.list_d = (float*)malloc((size_t)n * n * sizeof(float));
  s.list_len = (int*)malloc(n * sizeof(int));
  if (!s.m || !s.sepsum || !s.sep || !s.index_to_id || !s.id_to_index || !s.red_count || !s.rtm ||
      !s.red_head || !s.red_next || !s.maxrow || !s.list_id || !s.list_d || !s.list_len)
    return -2;
  memcpy(s.m, matrix, (size_t)n * n * sizeof(float));
  for (size_t k = 0; k < 2 * (size_t)n; k++) s.id_to_index[k] = -1;
  initialize(&s);
  int step = 0;
  while (s.cluster_count > 2) {
    find_min(&s);
    merge_min_nodes(&s, left, right, llen, rlen, step++);
    s.cluster_count--;
    update_data(&s);
  }
  /* rapidNJ.cpp:52-85 */
  int index1 = -1, index2 = -1;
  for (unsigned i = 0; i < n; i++) {
    if (s.index_to_id[i] != -1) {
      if (index1 == -1) {
        if (s.red_count[i] == 2) {
          index1 = (int)i;
          index2 = s.red_head[i];
          s.index_to_id[index2] = index2;
          break;
        } else {
          index1 = (int)i;
        }
      } else {
        index2 = (int)i;
        break;
      }
    }
  }
  int rc = 0;
  if (index1 == -1 || index2 == -1) {
    rc = -3;
  } else {
    last[0] = s.index_to_id[index1];
    last[1] = s.index_to_id[index2];
    *last_dist = s.m[(size_t)index1 * n + index2];
  }
  free(s.m);
  free(s.sepsum);
  free(s.sep);
  free(s.index_to_id);
  free(s.id_to_index);
  free(s.red_count);
  free(s.rtm);
  free(s.red_head);
  free(s.red_next);
  free(s.maxrow);
  free(s.list_id);
  free(s.list_d);
  free(s.list_len);
  return rc;
}

/* rapidNJ's run with the row-skipping audit attached; out[4] = bound violations, record
 * violations, rows skipped, rows evaluated (O(n^3): small matrices only). flags bit 0 leaves
 * the drift out of the bound (a mistake the audit must catch). */
int rnj_oracle_nj_skip_audit(const float* matrix, unsigned n, int negative_branches, int flags,
                             long long* out) {
  if (n < 3) return -1;
  nj_audit A;
  memset(&A, 0, sizeof A);
  A.no_drift = flags & 1;
  A.smin = (float*)calloc(n, sizeof(float));
  A.sref = (double*)calloc(n, sizeof(double));
  A.have = (int*)calloc(n, sizeof(int));
  A.rowid = (int*)calloc(n, sizeof(int));
  A.bestid = (int*)calloc(n, sizeof(int));
  A.skip = (int*)calloc(n, sizeof(int));
  A.prev_id = (int*)malloc(n * sizeof(int));
  A.prev_sep = (float*)calloc(n, sizeof(float));
  int *left = (int*)malloc(n * sizeof(int)), *right = (int*)malloc(n * sizeof(int)), last[2];
  double *llen = (double*)malloc(n * sizeof(double)), *rlen = (double*)malloc(n * sizeof(double));
  float ld;
  int rc = -2;
  if (A.smin && A.sref && A.have && A.rowid && A.bestid && A.skip && A.prev_id && A.prev_sep &&
      left && right && llen && rlen) {
    for (unsigned i = 0; i < n; i++) A.prev_id[i] = -1;
    g_audit = &A;
    rc = rnj_oracle_nj(matrix, n, negative_branches, left, right, llen, rlen, last, &ld);
    g_audit = NULL;
    out[0] = A.bound_violations;
    out[1] = A.record_violations;
    out[2] = A.skipped;
    out[3] = A.scanned;
  }
  free(A.smin); free(A.sref); free(A.have); free(A.rowid); free(A.bestid); free(A.skip);
  free(A.prev_id); free(A.prev_sep); free(left); free(right); free(llen); free(rlen);
  return rc;
}

/* ---- Newick (polytree.cpp:39-51, 402-458 without bootstrap counts) ------------------------ */
typedef struct {
  char* out;
  size_t cap, len;
} sbuf;

static void put(sbuf* b, const char* s, size_t k) {
  if (b->len + k < b->cap) memcpy(b->out + b->len, s, k);
  b->len += k;
}
static void puts_(sbuf* b, const char* s) { put(b, s, strlen(s)); }
static void put_len(sbuf* b, double x) {
  char tmp[64];
  const int k = snprintf(tmp, sizeof tmp, "%.5g", x);
  put(b, tmp, (size_t)k);
}
static void put_leaf(sbuf* b, const char* name, double len) {
  puts_(b, "'");
  puts_(b, name);
  puts_(b, "':");
  put_len(b, len);
}

typedef struct {
  unsigned n;
  const int *left, *right;
  const double* dist; /* by node id (leaves 0..n-1, internal n..2n-2) */
  const char* const* names;
} tree_view;

static void node(const tree_view* t, sbuf* b, int l, int r) {
  const unsigned n = t->n;
  if ((unsigned)l < n) {
    put_leaf(b, t->names[l], t->dist[l]);
  } else {
    puts_(b, "(");
    node(t, b, t->left[l - n], t->right[l - n]);
    puts_(b, "):");
    put_len(b, t->dist[l]);
  }
  puts_(b, ",");
  if ((unsigned)r < n) {
    put_leaf(b, t->names[r], t->dist[r]);
  } else {
    puts_(b, "(");
    node(t, b, t->left[r - n], t->right[r - n]);
    puts_(b, "):");
    put_len(b, t->dist[r]);
  }
}

size_t rnj_oracle_newick(unsigned n, const int* left, const int* right, const double* llen,
                         const double* rlen, const int* last, float last_dist,
                         const char* const* names, char* out, size_t cap) {
  double* dist = (double*)calloc(2 * (size_t)n, sizeof(double));
  if (!dist) return 0;
  for (unsigned k = 0; k + 2 < n; k++) {  /* addInternalNode: distances by child id */
    dist[left[k]] = llen[k];
    dist[right[k]] = rlen[k];
  }
  tree_view t = {n, left, right, dist, names};
  sbuf b = {out, cap, 0};
  const int sl = last[0], sr = last[1];
  puts_(&b, "(");
  if ((unsigned)sl < n) {
    put_leaf(&b, names[sl], last_dist);
  } else {
    if ((unsigned)sr >= n) puts_(&b, "(");
    node(&t, &b, left[sl - n], right[sl - n]);
    if ((unsigned)sr >= n) {
      puts_(&b, "):");
      put_len(&b, last_dist);
    }
  }
  puts_(&b, ",");
  if ((unsigned)sr < n) put_leaf(&b, names[sr], last_dist);
  else node(&t, &b, left[sr - n], right[sr - n]);
  puts_(&b, ");\n");
  if (b.len < b.cap) b.out[b.len] = 0;
  free(dist);
  return b.len;
}
