/*
 * g5_oracle.c -- CPU oracle (test infrastructure only; see g5_oracle.h).
 *
 * Plain C restatement of the reference Graph500 hot path.  It is written for
 * clarity, not speed: full 5x5 matrices for the MRG, a counting sort plus
 * per-row qsort for the CSR, a queue/scan BFS.  OpenMP is used only where
 * iterations are independent (tuple generation, per-row sorting).
 */
#include "g5_oracle.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* MRG: L'Ecuyer k=5, m=2^31-1, a1=107374182, a5=104480                     */
/* (generator/splittable_mrg.c:17-32).                                      */
/* ------------------------------------------------------------------------ */

#define OR_M 0x7FFFFFFFu
#define OR_X 107374182u
#define OR_Y 104480u

typedef struct { uint32_t m[5][5]; } mat5;

static uint32_t mulmod(uint32_t a, uint32_t b) { return (uint32_t)(((uint64_t)a * b) % OR_M); }
static uint32_t addmod(uint32_t a, uint32_t b) { return (uint32_t)(((uint64_t)a + b) % OR_M); }

static void mat_identity(mat5* r) {
  memset(r, 0, sizeof(*r));
  for (int i = 0; i < 5; ++i) r->m[i][i] = 1;
}

/* Companion matrix: z1' = x z1 + y z5; z(k+1)' = z(k). */
static void mat_A(mat5* r) {
  memset(r, 0, sizeof(*r));
  r->m[0][0] = OR_X;
  r->m[0][4] = OR_Y;
  for (int i = 1; i < 5; ++i) r->m[i][i - 1] = 1;
}

static void mat_mul(const mat5* a, const mat5* b, mat5* out) {
  mat5 r;
  for (int i = 0; i < 5; ++i)
    for (int j = 0; j < 5; ++j) {
      uint32_t acc = 0;
      for (int k = 0; k < 5; ++k) acc = addmod(acc, mulmod(a->m[i][k], b->m[k][j]));
      r.m[i][j] = acc;
    }
  *out = r;
}

static void mat_apply(const mat5* a, uint32_t z[5]) {
  uint32_t r[5];
  for (int i = 0; i < 5; ++i) {
    uint32_t acc = 0;
    for (int k = 0; k < 5; ++k) acc = addmod(acc, mulmod(a->m[i][k], z[k]));
    r[i] = acc;
  }
  memcpy(z, r, sizeof(r));
}

/* table[j][v] = A^(v * 256^j), j = 0..23 (192-bit exponents, byte-wise). */
static mat5 (*g_table)[256] = NULL;

static void build_table(void) {
  if (g_table) return;
  mat5 (*t)[256] = (mat5(*)[256])malloc(sizeof(mat5) * 24 * 256);
  mat5 p; /* A^(256^j) */
  mat_A(&p);
  for (int j = 0; j < 24; ++j) {
    if (j > 0) {
      for (int s = 0; s < 8; ++s) mat_mul(&p, &p, &p); /* p^256 */
    }
    mat_identity(&t[j][0]);
    for (int v = 1; v < 256; ++v) mat_mul(&t[j][v - 1], &p, &t[j][v]);
  }
#pragma omp critical(or_table)
  {
    if (!g_table) g_table = t; else free(t);
  }
}

void or_make_mrg_seed(uint64_t s1, uint64_t s2, uint32_t z[5]) {
  z[0] = (uint32_t)((s1 & 0x3FFFFFFF) + 1);
  z[1] = (uint32_t)(((s1 >> 30) & 0x3FFFFFFF) + 1);
  z[2] = (uint32_t)((s2 & 0x3FFFFFFF) + 1);
  z[3] = (uint32_t)(((s2 >> 30) & 0x3FFFFFFF) + 1);
  z[4] = (uint32_t)(((s2 >> 60) << 4) + (s1 >> 60) + 1);
}

void or_mrg_transition_entry(int byte_index, int value, uint32_t out9[9]) {
  build_table();
  const mat5* a = &g_table[byte_index][value];
  /* Bottom row is (s t u v w); first column top-down is (d c b a s). */
  out9[0] = a->m[4][0];
  out9[1] = a->m[4][1];
  out9[2] = a->m[4][2];
  out9[3] = a->m[4][3];
  out9[4] = a->m[4][4];
  out9[5] = a->m[3][0];
  out9[6] = a->m[2][0];
  out9[7] = a->m[1][0];
  out9[8] = a->m[0][0];
}

void or_mrg_skip(uint32_t z[5], uint64_t hi, uint64_t mid, uint64_t lo) {
  build_table();
  uint64_t parts[3] = {lo, mid, hi};
  for (int p = 0; p < 3; ++p) {
    uint64_t e = parts[p];
    for (int b = 0; e; ++b, e >>= 8) {
      int v = (int)(e & 0xFF);
      if (v) mat_apply(&g_table[p * 8 + b][v], z);
    }
  }
}

uint32_t or_mrg_get_uint(uint32_t z[5]) {
  uint32_t nz = addmod(mulmod(OR_X, z[0]), mulmod(OR_Y, z[4]));
  z[4] = z[3];
  z[3] = z[2];
  z[2] = z[1];
  z[1] = z[0];
  z[0] = nz;
  return nz;
}

double or_mrg_get_double(uint32_t z[5]) {
  /* Two draws scaled by (2^31-1)^-1 and (2^31-1)^-2; first draw first. */
  uint32_t u1 = or_mrg_get_uint(z);
  uint32_t u2 = or_mrg_get_uint(z);
  double a = (double)u1 * .000000000465661287524579692;
  double b = (double)u2 * .0000000000000000002168404346990492787;
  return a + b;
}

void or_make_random_numbers(int64_t n, uint64_t s1, uint64_t s2, int64_t pos, double* out) {
  uint32_t z[5];
  or_make_mrg_seed(s1, s2, z);
  or_mrg_skip(z, 2, 0, 2 * (uint64_t)pos);
  for (int64_t i = 0; i < n; ++i) out[i] = or_mrg_get_double(z);
}

/* ------------------------------------------------------------------------ */
/* Kronecker generator                                                        */
/* ------------------------------------------------------------------------ */

void or_scramble_values(uint64_t s1, uint64_t s2, uint64_t out[2]) {
  uint32_t z[5];
  or_make_mrg_seed(s1, s2, z);
  or_mrg_skip(z, 50, 7, 0);
  uint64_t v0 = or_mrg_get_uint(z);
  v0 *= UINT64_C(0xFFFFFFFF);
  v0 += or_mrg_get_uint(z);
  uint64_t v1 = or_mrg_get_uint(z);
  v1 *= UINT64_C(0xFFFFFFFF);
  v1 += or_mrg_get_uint(z);
  out[0] = v0;
  out[1] = v1;
}

static uint64_t bitrev64(uint64_t x) {
  uint64_t r = 0;
  for (int i = 0; i < 64; ++i) {
    r = (r << 1) | (x & 1);
    x >>= 1;
  }
  return r;
}

int64_t or_scramble(int64_t v0, int scale, uint64_t val0, uint64_t val1) {
  uint64_t v = (uint64_t)v0;
  v += val0 + val1;
  v *= (val0 | UINT64_C(0x4519840211493211));
  v = bitrev64(v) >> (64 - scale);
  v *= (val1 | UINT64_C(0x3050852102C843A5));
  v = bitrev64(v) >> (64 - scale);
  return (int64_t)v;
}

static int bernoulli4(uint32_t z[5], int a_num, int bc_num) {
  const uint32_t limit = UINT32_C(0xFFFFFFFF) % 10000u; /* 7295 */
  uint32_t val = or_mrg_get_uint(z);
  while (val < limit) val = or_mrg_get_uint(z);
  val %= 10000u;
  if (val < (uint32_t)bc_num) return 1;
  val -= bc_num;
  if (val < (uint32_t)bc_num) return 2;
  val -= bc_num;
  if (val < (uint32_t)a_num) return 0;
  return 3;
}

static void set_edge(or_packed_edge* e, int64_t v0, int64_t v1) {
  e->v0_low = (uint32_t)v0;
  e->v1_low = (uint32_t)v1;
  e->high = (uint32_t)(((v0 >> 32) & 0xFFFF) | ((v1 >> 16) & 0xFFFF0000u));
}

/* GraphGeneratorBase::generateInitialEdge, BINARY_TREE
 * (generator/graph_generator.hpp:515-524): tuple i joins vertex i+1 to its
 * parent (i+1)/2 in the heap order, both scrambled; 2^scale - 1 of them come
 * before the R-MAT tuples (:385-393, :612-619). */
void or_initial_tree_edges(int scale, uint64_t s1, uint64_t s2, int64_t start, int64_t end, or_packed_edge* out) {
  uint64_t sv[2];
  or_scramble_values(s1, s2, sv);
  for (int64_t i = start; i < end; ++i)
    set_edge(&out[i - start], or_scramble(i + 1, scale, sv[0], sv[1]), or_scramble((i + 1) / 2, scale, sv[0], sv[1]));
}

/* Edge weights of GraphGeneratorBase::generateWeight
 * (generator/graph_generator.hpp:480-506).  max_weight <= 0: the reference's
 * compiled branch, which stores the constant 0xBEEF in every tuple
 * (:499-500).  max_weight > 0: its per-edge form (:483-488, disabled in the
 * reference by #if 0): the seed state skipped by (30, 46, edge_index), one
 * draw, modulo max_weight, plus 1. */
void or_edge_weights(uint64_t s1, uint64_t s2, int64_t start, int64_t count, int max_weight, int32_t* out) {
  uint32_t seed[5];
  or_make_mrg_seed(s1, s2, seed);
  for (int64_t i = 0; i < count; ++i) {
    if (max_weight <= 0) {
      out[i] = 0xBEEF;
      continue;
    }
    uint32_t z[5];
    memcpy(z, seed, sizeof(z));
    or_mrg_skip(z, 30, 46, (uint64_t)(start + i));
    out[i] = (int32_t)(or_mrg_get_uint(z) % (uint32_t)max_weight) + 1;
  }
}

void or_generate_edges(int scale, int edgefactor, uint64_t s1, uint64_t s2, int a_num, int bc_num,
                       int64_t start, int64_t end, or_packed_edge* out) {
  (void)edgefactor;
  build_table();
  uint32_t seed[5];
  or_make_mrg_seed(s1, s2, seed);
  uint64_t sv[2];
  or_scramble_values(s1, s2, sv);
#pragma omp parallel for schedule(static)
  for (int64_t i = start; i < end; ++i) {
    uint32_t z[5];
    memcpy(z, seed, sizeof(z));
    or_mrg_skip(z, 0, (uint64_t)i, 0);
    int64_t nverts = (int64_t)1 << scale;
    int64_t bs = 0, bt = 0;
    while (nverts > 1) {
      int sq = bernoulli4(z, a_num, bc_num);
      int so = sq / 2, to = sq % 2;
      if (bs == bt && so > to) {
        int t = so;
        so = to;
        to = t;
      }
      nverts /= 2;
      bs += nverts * so;
      bt += nverts * to;
    }
    set_edge(&out[i - start], or_scramble(bs, scale, sv[0], sv[1]),
             or_scramble(bt, scale, sv[0], sv[1]));
  }
}

/* ------------------------------------------------------------------------ */
/* CSR                                                                        */
/* ------------------------------------------------------------------------ */

static int msb_index(uint64_t v) { /* index of highest set bit, -1 for 0 */
  int r = -1;
  while (v) {
    ++r;
    v >>= 1;
  }
  return r;
}

static const uint32_t* g_sort_deg; /* for qsort comparator */
static int cmp_deg_desc_id_asc(const void* a, const void* b) {
  uint32_t x = *(const uint32_t*)a, y = *(const uint32_t*)b;
  if (g_sort_deg[x] != g_sort_deg[y]) return g_sort_deg[x] > g_sort_deg[y] ? -1 : 1;
  return x < y ? -1 : (x > y);
}
static int cmp_u32(const void* a, const void* b) {
  uint32_t x = *(const uint32_t*)a, y = *(const uint32_t*)b;
  return x < y ? -1 : (x > y);
}

or_csr* or_csr_build(const or_packed_edge* e, int64_t m) {
  or_csr* g = (or_csr*)calloc(1, sizeof(or_csr));
  uint64_t orv = 0;
  for (int64_t i = 0; i < m; ++i) orv |= (uint64_t)(or_edge_v0(&e[i]) | or_edge_v1(&e[i]));
  g->scale_estimate = msb_index(orv) + 1;
  const int64_t n = (int64_t)1 << g->scale_estimate;
  g->num_verts = n;
  g->num_tuples = m;

  uint32_t* deg_old = (uint32_t*)calloc((size_t)n, sizeof(uint32_t));
  uint32_t* loops_old = (uint32_t*)calloc((size_t)n, sizeof(uint32_t));
  for (int64_t i = 0; i < m; ++i) {
    int64_t a = or_edge_v0(&e[i]), b = or_edge_v1(&e[i]);
    if (a == b) {
      ++loops_old[a];
      ++g->num_self_loops;
      continue;
    }
    ++deg_old[a];
    ++deg_old[b];
  }
  g->directed_before = 2 * (m - g->num_self_loops);

  /* VERTEX_REORDERING=2: degree descending, ties by original id ascending. */
  g->new2old = (uint32_t*)malloc((size_t)n * sizeof(uint32_t));
  g->old2new = (uint32_t*)malloc((size_t)n * sizeof(uint32_t));
  for (int64_t i = 0; i < n; ++i) g->new2old[i] = (uint32_t)i;
  g_sort_deg = deg_old;
  qsort(g->new2old, (size_t)n, sizeof(uint32_t), cmp_deg_desc_id_asc);
  for (int64_t i = 0; i < n; ++i) g->old2new[g->new2old[i]] = (uint32_t)i;
  int64_t na = 0;
  while (na < n && deg_old[g->new2old[na]] > 0) ++na;
  g->num_active = na;
  g->degree_raw = (uint32_t*)malloc((size_t)n * sizeof(uint32_t));
  g->self_loops = (uint32_t*)malloc((size_t)n * sizeof(uint32_t));
  for (int64_t i = 0; i < n; ++i) {
    g->degree_raw[i] = deg_old[g->new2old[i]];
    g->self_loops[i] = loops_old[g->new2old[i]];
  }

  /* Counting sort of directed pairs by new source, then per-row sort+unique. */
  uint64_t* start = (uint64_t*)malloc((size_t)(na + 1) * sizeof(uint64_t));
  start[0] = 0;
  for (int64_t v = 0; v < na; ++v) start[v + 1] = start[v] + g->degree_raw[v];
  uint32_t* tmp = (uint32_t*)malloc((size_t)(start[na] ? start[na] : 1) * sizeof(uint32_t));
  uint64_t* fill = (uint64_t*)malloc((size_t)(na + 1) * sizeof(uint64_t));
  memcpy(fill, start, (size_t)(na + 1) * sizeof(uint64_t));
  for (int64_t i = 0; i < m; ++i) {
    int64_t a = or_edge_v0(&e[i]), b = or_edge_v1(&e[i]);
    if (a == b) continue;
    uint32_t na_ = g->old2new[a], nb_ = g->old2new[b];
    tmp[fill[na_]++] = nb_;
    tmp[fill[nb_]++] = na_;
  }
  free(fill);
  uint64_t* ulen = (uint64_t*)malloc((size_t)(na + 1) * sizeof(uint64_t));
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t v = 0; v < na; ++v) {
    uint32_t* r = tmp + start[v];
    uint64_t len = start[v + 1] - start[v];
    qsort(r, (size_t)len, sizeof(uint32_t), cmp_u32);
    uint64_t k = 0;
    for (uint64_t j = 0; j < len; ++j)
      if (j == 0 || r[j] != r[j - 1]) r[k++] = r[j];
    ulen[v] = k;
  }
  g->offsets = (uint64_t*)malloc((size_t)(na + 1) * sizeof(uint64_t));
  g->offsets[0] = 0;
  for (int64_t v = 0; v < na; ++v) g->offsets[v + 1] = g->offsets[v] + ulen[v];
  g->nnz = (int64_t)g->offsets[na];
  g->cols = (uint32_t*)malloc((size_t)(g->nnz ? g->nnz : 1) * sizeof(uint32_t));
  for (int64_t v = 0; v < na; ++v)
    memcpy(g->cols + g->offsets[v], tmp + start[v], (size_t)ulen[v] * sizeof(uint32_t));
  free(ulen);
  free(tmp);
  free(start);
  free(deg_old);
  free(loops_old);
  return g;
}

void or_csr_free(or_csr* g) {
  if (!g) return;
  free(g->offsets);
  free(g->cols);
  free(g->new2old);
  free(g->old2new);
  free(g->degree_raw);
  free(g->self_loops);
  free(g);
}

int or_csr_has_edge(const or_csr* g, int64_t v) {
  if (v < 0 || v >= g->num_verts) return 0;
  return g->old2new[v] < (uint64_t)g->num_active;
}

int64_t or_find_max_used_vertex(const or_csr* g) {
  for (int64_t v = g->num_verts - 1; v > 0; --v)
    if (or_csr_has_edge(g, v)) return v;
  return 0;
}

int or_find_roots(const or_csr* g, uint64_t s1, uint64_t s2, int num_roots, int64_t* roots) {
  int64_t counter = 0;
  const int64_t nglobal = g->num_verts;
  int idx;
  for (idx = 0; idx < num_roots; ++idx) {
    int64_t root;
    for (;;) {
      double d[2];
      or_make_random_numbers(2, s1, s2, counter, d);
      root = (int64_t)((d[0] + d[1]) * nglobal) % nglobal;
      counter += 2;
      if (counter > 2 * nglobal) break;
      int dup = 0;
      for (int i = 0; i < idx; ++i)
        if (roots[i] == root) {
          dup = 1;
          break;
        }
      if (dup) continue;
      if (or_csr_has_edge(g, root)) break;
    }
    roots[idx] = root;
  }
  return idx;
}

/* ------------------------------------------------------------------------ */
/* BFS                                                                        */
/* ------------------------------------------------------------------------ */

int or_bfs(const or_csr* g, int64_t root_orig, uint8_t* level_by_orig, int64_t* pred_by_orig,
           or_bfs_trace* tr) {
  const int64_t n = g->num_verts, na = g->num_active;
  if (root_orig < 0 || root_orig >= n) return -1;
  const uint32_t root = g->old2new[root_orig];
  if (root >= na) return -2; /* root without edges */
  uint8_t* lv = (uint8_t*)malloc((size_t)n);
  int64_t* par = (int64_t*)malloc((size_t)n * sizeof(int64_t));
  memset(lv, 0xFF, (size_t)n);
  for (int64_t i = 0; i < n; ++i) par[i] = -1;
  uint32_t* q = (uint32_t*)malloc((size_t)(na + 1) * sizeof(uint32_t));
  uint32_t* nq = (uint32_t*)malloc((size_t)(na + 1) * sizeof(uint32_t));
  int64_t qn = 1;
  q[0] = root;
  lv[root] = 0;
  par[root] = root;

  or_bfs_trace local;
  if (!tr) tr = &local;
  memset(tr, 0, sizeof(*tr));
  double total = 1.0; /* 1 B level write for the root */

  int next_forward = 1;
  int growing = 1;
  int64_t global_nq = 0; /* global_nq_size_ before level 1 */
  int64_t visited = 1;
  int level = 0;
  int rc = 0;
  for (;;) {
    ++level;
    if (level >= 255) {
      rc = -3; /* uint8 level overflow */
      break;
    }
    int64_t prev_nq = global_nq;
    int forward = next_forward;
    double bytes = 0;
    int64_t examined = 0;
    int64_t found = 0;
    if (forward) {
      for (int64_t i = 0; i < qn; ++i) {
        uint32_t u = q[i];
        uint64_t b = g->offsets[u], e = g->offsets[u + 1];
        bytes += 4 + 8 + 4.0 * (double)(e - b);
        examined += (int64_t)(e - b);
        for (uint64_t k = b; k < e; ++k) {
          uint32_t w = g->cols[k];
          if (lv[w] == 0xFF) {
            lv[w] = (uint8_t)level;
            par[w] = u;
            nq[found++] = w;
          }
        }
      }
      bytes += 4.0 * (double)found;
    } else {
      bytes += 3.0 * (double)n / 8.0;
      for (int64_t v = 0; v < na; ++v) {
        if (lv[v] != 0xFF) continue;
        uint64_t b = g->offsets[v], e = g->offsets[v + 1];
        int64_t k_v = 0;
        for (uint64_t k = b; k < e; ++k) {
          uint32_t w = g->cols[k];
          ++k_v;
          if (lv[w] < (uint8_t)level) { /* visited before this level */
            lv[v] = (uint8_t)level;
            par[v] = w;
            nq[found++] = (uint32_t)v;
            break;
          }
        }
        bytes += 8 + 4.0 * (double)k_v;
        examined += k_v;
      }
    }
    bytes += (double)found; /* uint8 level writes */
    tr->direction[level] = forward;
    tr->nq[level] = found;
    tr->bytes[level] = bytes;
    tr->edges_examined[level] = examined;
    total += bytes;
    global_nq = found;
    visited += found;
    tr->num_levels = level + 1;
    if (found == 0) break;
    /* alpha/beta switch, mpi/bfs.hpp:2799-2829 */
    int64_t unvisited = na - visited;
    if (growing && global_nq > prev_nq) {
      if (forward && (double)global_nq > (double)na / 2000.0) next_forward = 0;
    } else {
      if (!forward && unvisited < (int64_t)((double)na / 8.0)) {
        next_forward = 1;
        growing = 0;
      }
    }
    uint32_t* t = q;
    q = nq;
    nq = t;
    qn = found;
  }
  tr->total_bytes = total;
  tr->reached = visited;
  /* e_r from per-vertex tuple weights: sum over reached of (deg_raw/2 + loops). */
  uint64_t dsum = 0, lsum = 0;
  for (int64_t v = 0; v < na; ++v)
    if (lv[v] != 0xFF) {
      dsum += g->degree_raw[v];
      lsum += g->self_loops[v];
    }
  tr->teps_edges = (int64_t)(dsum / 2 + lsum);

  for (int64_t v = 0; v < n; ++v) {
    int64_t o = g->new2old[v];
    if (level_by_orig) level_by_orig[o] = lv[v];
    if (pred_by_orig) {
      if (lv[v] == 0xFF)
        pred_by_orig[o] = -1;
      else if ((uint32_t)v == root)
        pred_by_orig[o] = root_orig;
      else
        pred_by_orig[o] = (int64_t)g->new2old[par[v]] | ((int64_t)lv[v] << 48);
    }
  }
  free(lv);
  free(par);
  free(q);
  free(nq);
  return rc;
}

double or_msbfs_bytes(const or_csr* g, int lanes, const uint8_t* levels_new, int* num_levels,
                      double* per_level, const uint64_t* pull_lanes, const uint64_t* push_lanes) {
  const int64_t n = g->num_verts, na = g->num_active;
  int maxl = 0;
  for (int r = 0; r < lanes; ++r)
    for (int64_t v = 0; v < na; ++v) {
      uint8_t l = levels_new[(size_t)r * n + v];
      if (l != 0xFF && l > maxl) maxl = l;
    }
  uint64_t* front = (uint64_t*)calloc((size_t)na, sizeof(uint64_t));
  double total = 0;
  /* Batch levels 1..maxl+1 (the last one discovers nothing and ends the batch). */
  int L = maxl + 1;
  for (int l = 1; l <= L; ++l) {
    for (int64_t v = 0; v < na; ++v) {
      uint64_t f = 0;
      for (int r = 0; r < lanes; ++r)
        if (levels_new[(size_t)r * n + v] == l - 1) f |= (uint64_t)1 << r;
      front[v] = f;
    }
    double bytes = 24.0 * (double)n;
    const uint64_t pull_mask = pull_lanes ? pull_lanes[l] : ~(uint64_t)0;
    const uint64_t push_mask = push_lanes ? push_lanes[l] : 0;
    for (int64_t v = 0; v < na; ++v) {
      if (front[v] & push_mask) bytes += 8 + 12.0 * (double)(g->offsets[v + 1] - g->offsets[v]);
      uint64_t need = 0;
      for (int r = 0; r < lanes; ++r) {
        uint8_t lv = levels_new[(size_t)r * n + v];
        if (lv == 0xFF || lv >= l) need |= (uint64_t)1 << r;
      }
      need &= pull_mask;
      if (!need) continue;
      uint64_t acc = 0;
      int64_t k = 0;
      for (uint64_t j = g->offsets[v]; j < g->offsets[v + 1]; ++j) {
        ++k;
        acc |= front[g->cols[j]];
        if ((acc & need) == need) break;
      }
      bytes += 8 + 12.0 * (double)k;
    }
    if (per_level) per_level[l] = bytes;
    total += bytes;
  }
  if (num_levels) *num_levels = L + 1;
  free(front);
  return total;
}

/* ------------------------------------------------------------------------ */
/* Validation                                                                 */
/* ------------------------------------------------------------------------ */

static int64_t pred_of(int64_t e) { return (int64_t)((uint64_t)e << 16) >> 16; }
static uint16_t depth_of(int64_t e) { return (uint16_t)(((uint64_t)e >> 48) & 0xFFFF); }

#define FAIL(...)                                             \
  do {                                                        \
    if (msg && msglen > 0) snprintf(msg, msglen, __VA_ARGS__); \
    goto fail;                                                \
  } while (0)

int or_validate(const or_packed_edge* e, int64_t m, int64_t nglobalverts, int64_t nlocalverts,
                int64_t root, const int64_t* pred, int64_t* edge_visit_count, char* msg,
                int msglen) {
  unsigned char* valid = NULL;
  *edge_visit_count = 0;
  if (msg && msglen > 0) msg[0] = 0;
  /* check_value_ranges, validate.hpp:822-838 */
  for (int64_t i = 0; i < nlocalverts; ++i) {
    int64_t p = pred_of(pred[i]);
    if (p < -1 || p >= nglobalverts)
      FAIL("parent of vertex %lld is out-of-range value %lld", (long long)i, (long long)p);
  }
  /* root is its own parent, validate.hpp:530-536 */
  if (root < 0 || root >= nlocalverts) FAIL("root vertex %lld is invalid", (long long)root);
  if (pred_of(pred[root]) != root)
    FAIL("parent of root vertex %lld is %lld", (long long)root, (long long)pred_of(pred[root]));
  /* nothing else is its own parent, validate.hpp:542-569 */
  for (int64_t i = 0; i < nlocalverts; ++i)
    if (i != root && pred_of(pred[i]) != -1 && pred_of(pred[i]) == i)
      FAIL("parent of non-root vertex %lld is itself", (long long)i);
  /* check_bfs_depth_map_using_predecessors, validate.hpp:920-995 */
  if (depth_of(pred[root]) != 0) FAIL("depth of root vertex is %u, not 0", depth_of(pred[root]));
  for (int64_t i = 0; i < nlocalverts; ++i) {
    int64_t p = pred_of(pred[i]);
    uint16_t d = depth_of(pred[i]);
    if (p == -1 && d != UINT16_MAX) FAIL("depth of vertex %lld with no predecessor is %u", (long long)i, d);
    if (p != -1 && d == UINT16_MAX) FAIL("predecessor of unreachable vertex %lld is %lld", (long long)i, (long long)p);
  }
  for (int64_t i = 0; i < nlocalverts; ++i) {
    if (i == root) continue;
    int64_t p = pred_of(pred[i]);
    if (p == -1) continue;
    uint16_t dp = depth_of(pred[p]);
    if (dp == UINT16_MAX) FAIL("predecessor %lld of vertex %lld is marked as unreachable", (long long)p, (long long)i);
    if ((int)depth_of(pred[i]) != (int)dp + 1)
      FAIL("BFS predecessors do not form a tree; see vertices %lld and %lld", (long long)i, (long long)p);
  }
  /* edge checks and TEPS count, validate.hpp:588-792 */
  valid = (unsigned char*)calloc(nlocalverts > 0 ? (size_t)nlocalverts : 1, 1);
  int64_t count = 0;
  for (int64_t i = 0; i < m; ++i) {
    int64_t s = or_edge_v0(&e[i]), t = or_edge_v1(&e[i]);
    uint16_t sd = depth_of(pred[s]), td = depth_of(pred[t]);
    if (sd != UINT16_MAX && td == UINT16_MAX)
      FAIL("edge connects vertex %lld in the BFS tree to vertex %lld outside the tree", (long long)s, (long long)t);
    else if (sd == UINT16_MAX && td != UINT16_MAX)
      FAIL("edge connects vertex %lld in the BFS tree to vertex %lld outside the tree", (long long)t, (long long)s);
    else if ((int)sd - (int)td < -1 || (int)sd - (int)td > 1)
      FAIL("depths of edge endpoints %lld and %lld are too far apart", (long long)s, (long long)t);
    else if (sd != UINT16_MAX)
      ++count;
    if (pred_of(pred[s]) == t) valid[s] = 1;
    if (pred_of(pred[t]) == s) valid[t] = 1;
  }
  for (int64_t i = 0; i < nlocalverts; ++i) {
    if (pred_of(pred[i]) == -1) continue;
    if (i == root) continue;
    if (!valid[i]) FAIL("no graph edge from vertex %lld to its parent %lld", (long long)i, (long long)pred_of(pred[i]));
  }
  free(valid);
  *edge_visit_count = count;
  return 1;
fail:
  free(valid);
  return 0;
}

static int cmp_double(const void* a, const void* b) {
  double x = *(const double*)a, y = *(const double*)b;
  return x < y ? -1 : (x == y ? 0 : 1);
}

void or_statistics(const double* x, int n, double out[7]) {
  double mean = 0;
  for (int i = 0; i < n; ++i) mean += x[i];
  mean /= n;
  double var = 0;
  for (int i = 0; i < n; ++i) var += (x[i] - mean) * (x[i] - mean);
  var /= (n - 1);
  double* xx = (double*)malloc((size_t)n * sizeof(double));
  memcpy(xx, x, (size_t)n * sizeof(double));
  qsort(xx, (size_t)n, sizeof(double), cmp_double);
  out[0] = xx[0];
  out[1] = (xx[(n - 1) / 4] + xx[n / 4]) * .5;
  out[2] = (xx[(n - 1) / 2] + xx[n / 2]) * .5;
  out[3] = (xx[n - 1 - (n - 1) / 4] + xx[n - 1 - n / 4]) * .5;
  out[4] = xx[n - 1];
  out[5] = mean;
  out[6] = sqrt(var);
  free(xx);
}
