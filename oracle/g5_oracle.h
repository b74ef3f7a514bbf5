/*
 * g5_oracle.h -- CPU oracle for the Graph500 BFS hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * algorithm (suzumura/graph500) used as the checker for the HIP product path.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * it; the product library never links or calls it.
 *
 * Every function cites the reference file:line whose behaviour it restates.
 * Pinning status of each function (what known answers check it) is listed in
 * the comment above its declaration.  No function here is "parity unpinned".
 */
#ifndef G5_ORACLE_H
#define G5_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* 12-byte packed edge tuple: mpi/primitives.hpp:51-64 (UnweightedPackedEdge). */
typedef struct or_packed_edge {
  uint32_t v0_low;
  uint32_t v1_low;
  uint32_t high; /* v0 bits 47:32 in 15:0, v1 bits 47:32 in 31:16 */
} or_packed_edge;

/* ---------------- MRG (generator/splittable_mrg.c) ---------------- */

/* make_mrg_seed, generator/graph_generator.hpp:327-333.
 * Pinned: (2,3) -> (3,1,4,1,1) closed form. */
void or_make_mrg_seed(uint64_t userseed1, uint64_t userseed2, uint32_t z[5]);

/* A^(value * 256^byte_index) in the (s,t,u,v,w,a,b,c,d) form of
 * generator/splittable_mrg.c:36-47, computed here from full 5x5 matrix powers
 * (restating dump_mrg_powers, generator/splittable_mrg.c:212-263).
 * Pinned: entries of generator/mrg_transitions.c kept as golden fixtures. */
void or_mrg_transition_entry(int byte_index, int value, uint32_t out9[9]);

/* mrg_skip, generator/splittable_mrg.c:185-200: state <- A^e state with the
 * 192-bit exponent e = (high, middle, low). Pinned: skip(e1) then skip(e2) ==
 * skip(e1+e2) == e1+e2 plain steps (tests), and the generator goldens. */
void or_mrg_skip(uint32_t z[5], uint64_t exp_high, uint64_t exp_middle, uint64_t exp_low);

/* mrg_get_uint_orig, generator/splittable_mrg.c:170-177,266-269. */
uint32_t or_mrg_get_uint(uint32_t z[5]);

/* mrg_get_double_orig, generator/splittable_mrg.c:272-276. */
double or_mrg_get_double(uint32_t z[5]);

/* make_random_numbers, generator/graph_generator.hpp:338-354. */
void or_make_random_numbers(int64_t nvalues, uint64_t userseed1, uint64_t userseed2,
                            int64_t position, double* result);

/* ---------------- Kronecker generator (generator/graph_generator.hpp) ------- */

/* scramble_val0_/val1_ from GraphGeneratorBase ctor, graph_generator.hpp:375-384. */
void or_scramble_values(uint64_t userseed1, uint64_t userseed2, uint64_t out[2]);

/* scramble(), graph_generator.hpp:457-469 (bitreverse :405-451). */
int64_t or_scramble(int64_t v, int scale, uint64_t val0, uint64_t val1);

/* RmatGraphGenerator::generateRange + make_one_edge + generate_4way_bernoulli,
 * graph_generator.hpp:612-627,651-704; spec2010 initiator 5700/1900
 * (mpi/benchmark_helper.hpp:308-314).  Writes tuples [start,end).
 * Pinned: pf_nedge[16] (mpi/utils.hpp:3681), reference tuple-stream SHA-256 at
 * SCALE 16 (golden), directed-entry counts (golden). Uses OpenMP. */
void or_generate_edges(int scale, int edgefactor, uint64_t userseed1, uint64_t userseed2,
                       int a_num, int bc_num, int64_t start, int64_t end, or_packed_edge* out);

/* Binary-tree initial tuples [start, end) of the spec-2012 generator
 * (generateInitialEdge, graph_generator.hpp:515-524); the R-MAT tuples then
 * take indices 2^scale - 1 onward (or_generate_edges with that start).
 * Pinned: the reference spec-2012 tuple stream (golden, SCALE 14 and 16). */
/* Tuple weights of generateWeight, generator/graph_generator.hpp:480-506:
 * 0xBEEF for max_weight <= 0 (the reference's compiled branch), else the
 * per-edge draw (skip (30, 46, edge_index)) % max_weight + 1. */
void or_edge_weights(uint64_t userseed1, uint64_t userseed2, int64_t start, int64_t count, int max_weight,
                     int32_t* out);
void or_initial_tree_edges(int scale, uint64_t userseed1, uint64_t userseed2, int64_t start, int64_t end,
                           or_packed_edge* out);

static inline int64_t or_edge_v0(const or_packed_edge* e) {
  return (int64_t)(e->v0_low | ((uint64_t)(e->high & 0xFFFF) << 32));
}
static inline int64_t or_edge_v1(const or_packed_edge* e) {
  return (int64_t)(e->v1_low | ((uint64_t)(e->high >> 16) << 32));
}

/* ---------------- CSR (mpi/graph_constructor.hpp) ---------------- */

/* Canonical single-GPU CSR: rows and columns in VERTEX_REORDERING=2 order
 * (degree counting duplicates but not self-loops, descending; ties by original
 * id ascending), each row ascending by new id, duplicates removed.
 * Restates searchMaxVertex/initializeParameters (:735-763,:620-635),
 * DegreeCalculation::add/calc_degree (:249-352), sortEdges (:1363-1442),
 * constructFromWideCSR/isolateFirstEdge (:881-990,:845-879),
 * computeNumVertices (:1444-1464).  Pinned: directed-entry counts and n_active
 * from reference runs (golden); brute force on tiny inputs (tests). */
typedef struct or_csr {
  int scale_estimate;      /* log_orig_global_verts_ */
  int64_t num_verts;       /* 2^scale_estimate */
  int64_t num_active;      /* num_global_verts_: vertices with >= 1 non-loop edge */
  int64_t num_tuples;      /* input tuples */
  int64_t num_self_loops;  /* tuples with v0 == v1 */
  int64_t directed_before; /* 2 * non-loop tuples (before dedup) */
  int64_t nnz;             /* directed entries after dedup */
  uint64_t* offsets;       /* [num_active+1] */
  uint32_t* cols;          /* [nnz], new ids */
  uint32_t* new2old;       /* [num_verts] */
  uint32_t* old2new;       /* [num_verts] */
  uint32_t* degree_raw;    /* [num_verts] by new id: non-loop tuple endpoints */
  uint32_t* self_loops;    /* [num_verts] by new id: self-loop tuples */
} or_csr;

or_csr* or_csr_build(const or_packed_edge* edges, int64_t m);
void or_csr_free(or_csr* g);
/* has_edge(v) on an original id: Graph2DCSR::has_edge, graph_constructor.hpp:101-110 */
int or_csr_has_edge(const or_csr* g, int64_t v);
/* find_max_used_vertex, mpi/benchmark_helper.hpp:510-528 */
int64_t or_find_max_used_vertex(const or_csr* g);

/* find_roots, mpi/benchmark_helper.hpp:475-508.  Returns number of roots.
 * Pinned: the 16 reference roots at SCALE 16 (golden). */
int or_find_roots(const or_csr* g, uint64_t userseed1, uint64_t userseed2, int num_roots,
                  int64_t* roots);

/* ---------------- BFS (mpi/bfs.hpp) ---------------- */

#define OR_MAX_LEVELS 256

typedef struct or_bfs_trace {
  int num_levels;                     /* levels executed, including the last empty one */
  int direction[OR_MAX_LEVELS];       /* 1 = top-down, 0 = bottom-up, index = level */
  int64_t nq[OR_MAX_LEVELS];          /* vertices discovered at that level */
  double bytes[OR_MAX_LEVELS];        /* algorithmic bytes of that level */
  int64_t edges_examined[OR_MAX_LEVELS];
  double total_bytes;                 /* B_r incl. 1 B per reached vertex */
  int64_t reached;                    /* vertices with a level (incl. root) */
  int64_t teps_edges;                 /* e_r by per-vertex tuple weights */
} or_bfs_trace;

/* Level-synchronous direction-optimizing BFS with the reference alpha/beta
 * rule: BfsBase::run_bfs (mpi/bfs.hpp:2651-2940, switch :2799-2829),
 * initialize_memory/first_expand (:241-276,:477-518), top-down relax
 * (:1248-1364), bottom-up first-visited-neighbour probe (:1839-1967).
 * level_by_orig: uint8 level per original id (0xFF unreached);
 * pred_by_orig: parent | level << 48, or -1 (the reference's pred encoding,
 * mpi/bfs.hpp:1203,1271,2470).  Either output may be NULL.
 * Algorithmic bytes follow the single-root model of the design notes.
 * Pinned: reference per-root level histograms and direction schedule at
 * SCALE 16 (golden), brute-force BFS on tiny graphs, validator rules. */
int or_bfs(const or_csr* g, int64_t root, uint8_t* level_by_orig, int64_t* pred_by_orig,
           or_bfs_trace* trace);

/* Batched algorithmic bytes for a 64-lane bit-packed pull sweep given the
 * per-lane levels (by NEW id, 0xFF unreached). levels is lanes x num_verts.
 * Returns total bytes; per-level bytes into per_level (may be NULL).  The
 * batch's output is its per-level frontier words (written at every level and
 * counted in the 24 n term); the uint8 level array of a root is unpacked
 * from them on request, outside the batch, so no per-(root, vertex) output
 * byte is added.
 * pull_lanes / push_lanes (indexed by level, may be NULL): with per-root
 * directions only the pulled lanes count as needed, and a vertex whose
 * frontier word holds a pushed lane adds 8 + 12 deg (NULL: every lane pulls). */
double or_msbfs_bytes(const or_csr* g, int lanes, const uint8_t* levels_new, int* num_levels,
                      double* per_level, const uint64_t* pull_lanes, const uint64_t* push_lanes);

/* ---------------- Validation (mpi/validate.hpp) ---------------- */

/* BfsValidation::validate, mpi/validate.hpp:506-802 with helpers :809-995.
 * Returns 1 if valid, 0 otherwise; *edge_visit_count = traversed tuples.
 * msg receives the first error.  Pinned: pf_nedge table, planted corruptions
 * in tests (each rule has a failing case). */
int or_validate(const or_packed_edge* edges, int64_t m, int64_t nglobalverts,
                int64_t nlocalverts, int64_t root, const int64_t* pred,
                int64_t* edge_visit_count, char* msg, int msglen);

/* get_statistics, mpi/logfile.h:19-44: out = {min, q1, median, q3, max, mean, std}. */
void or_statistics(const double* x, int n, double out[7]);

#ifdef __cplusplus
}
#endif

#endif
