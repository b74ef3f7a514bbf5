/*
 * g5.h -- C ABI of the MI355X-native Graph500 BFS hot path.
 *
 * One context (g5_ctx) owns one GPU's copy of an edge list, the CSR built from
 * it and the BFS state.  Every entry point returns 0 on success or a negative
 * G5_ERR_* code; g5_last_error() then gives a message.  No exception crosses
 * this boundary.  Calls on one context must be serialized by the caller (the
 * reference is MPI_THREAD_SINGLE, mpi/utils.hpp:1580); different contexts are
 * independent.  Host pointers passed in are only read/written during the call.
 *
 * Each entry point names the reference interface it replaces (file:line,
 * relative to the suzumura/graph500 tree).
 */
#ifndef G5_H
#define G5_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define G5_OK 0
#define G5_ERR_ARG (-1)      /* invalid argument */
#define G5_ERR_STATE (-2)    /* call out of order (e.g. run_bfs before construct) */
#define G5_ERR_HIP (-3)      /* HIP runtime error */
#define G5_ERR_MEM (-4)      /* device allocation failed */
#define G5_ERR_LIMIT (-5)    /* input beyond a documented limit */

#define G5_MAX_LEVELS 256

typedef struct g5_ctx g5_ctx;

/* 12-byte packed tuple, identical layout to UnweightedPackedEdge
 * (mpi/primitives.hpp:51-64): v0 = v0_low | (high & 0xFFFF) << 32,
 * v1 = v1_low | (high >> 16) << 32. */
typedef struct g5_packed_edge {
  uint32_t v0_low;
  uint32_t v1_low;
  uint32_t high;
} g5_packed_edge;

/* WeightedEdge (mpi/primitives.hpp:87-107): the packed tuple and an int
 * weight, 16 bytes. */
typedef struct g5_weighted_edge {
  uint32_t v0_low;
  uint32_t v1_low;
  uint32_t high;
  int32_t weight;
} g5_weighted_edge;

/* Graph facts after construction (Graph2DCSR members,
 * mpi/graph_constructor.hpp:113-143). */
typedef struct g5_graph_info {
  int32_t scale_estimate;   /* log_orig_global_verts_ */
  int32_t reorder_mode;     /* VERTEX_REORDERING used */
  int64_t num_verts;        /* 2^scale_estimate = pred_size() at 1x1 */
  int64_t num_active;       /* num_global_verts_: vertices with >= 1 edge */
  int64_t num_tuples;       /* input tuples */
  int64_t num_self_loops;   /* tuples with v0 == v1 (dropped) */
  int64_t directed_before;  /* directed entries before duplicate removal */
  int64_t nnz;              /* num_global_edges_: directed entries after removal */
  int64_t max_degree;       /* largest degree counting duplicates */
  int64_t max_used_vertex;  /* find_max_used_vertex, mpi/benchmark_helper.hpp:510-528 */
  double construct_ms;      /* device time of the last g5_construct */
  double generate_ms;       /* device time of the last g5_generate_kronecker */
} g5_graph_info;

/* Per-run trace of the last g5_run_bfs (the reference's VERVOSE per-level
 * report, mpi/bfs.hpp:2836-2879). Index 0 is unused; levels start at 1. */
typedef struct g5_bfs_stats {
  int32_t num_levels;               /* levels executed + 1 (last level finds nothing) */
  int32_t direction[G5_MAX_LEVELS]; /* 1 = top-down, 0 = bottom-up */
  int64_t nq[G5_MAX_LEVELS];        /* vertices discovered at that level */
  double level_ms[G5_MAX_LEVELS];   /* device time of each level's kernels (HIP events) */
  double device_ms;                 /* device time of the whole run (HIP events) */
  int64_t reached;                  /* vertices with a level, root included */
  int32_t bu_kernel[G5_MAX_LEVELS]; /* 1 = the level's vertices were found by the bottom-up kernel
                                       (every bottom-up level, and the top-down levels tail_pull
                                       hands to it; see g5_bfs_tuning) */
} g5_bfs_stats;

/* Trace of the last g5_run_bfs_batch.  Index 0 is unused; levels start at 1. */
typedef struct g5_batch_stats {
  int32_t num_roots;
  int32_t num_levels;               /* levels executed + 1 (last level finds nothing) */
  int32_t direction[G5_MAX_LEVELS]; /* 1 = push (top-down), 0 = vertex-parallel pull (bottom-up),
                                       2 = edge-parallel pull sweep, 3 = some roots push and
                                       the others pull (lane_rule 1) */
  int64_t pairs[G5_MAX_LEVELS];     /* (root, vertex) pairs discovered at that level */
  double level_ms[G5_MAX_LEVELS];   /* device time per level (HIP events) */
  double device_ms;                 /* device time of the whole batch */
  int32_t persistent_grid;          /* workgroups of the persistent launch (0: per-level launches) */
  int32_t persistent_levels;        /* levels run inside the persistent launch */
} g5_batch_stats;

/* Tuning of the batched BFS (set by the autotuner, graph500_amd/autotune.py;
 * defaults: lane_rule 1, alpha 128, pull_grid 8192, sweep_below 0, sweep_grid
 * 4096, persistent 0, persist_levels 16, persist_grid 0, persist_waves 4,
 * pull_filter 0, beta 50, filter_from 0.01, pull_unroll 0, compact_below 0, mixed_push 1). */
typedef struct g5_batch_tuning {
  double alpha;        /* push while the frontier union has fewer than nnz/alpha edges */
  int64_t pull_grid;   /* workgroups of the vertex-parallel pull (256 threads each); pull levels with
                          fewer than 1/8 of the (root, vertex) pairs unreached use at most 2048 */
  double sweep_below;  /* (lane_rule 0) pull levels use the edge-parallel sweep while the (root, vertex)
                          pairs reached so far are below sweep_below x roots x rows
                          (reorder modes 1 and 2; 0 disables the sweep) */
  int64_t sweep_grid;  /* workgroups of the edge-parallel sweep */
  int32_t persistent;      /* 1: all levels in one persistent kernel launch, replayed as a hipGraph
                              (reorder modes 1 and 2); 0: one launch sequence per level */
  int32_t persist_levels;  /* levels the persistent launch covers (frontier arrays kept in HBM);
                              deeper searches continue level by level */
  int64_t persist_grid;    /* workgroups of the persistent launch; 0 = all co-resident ones */
  int32_t persist_waves;   /* register budget of the persistent kernel: 4, 6 or 8 waves per SIMD */
  int32_t pull_filter;     /* 1: the pull reads a neighbour's frontier word only when its bit in a
                              1-bit-per-vertex "frontier word is not zero" map is set; 2: only
                              for neighbours from filter_from x rows on (the lower-degree ones) */
  double beta;             /* pull (never push) once fewer than 1/beta of the (root, vertex) pairs
                              are unreached; 0 disables */
  double filter_from;      /* pull_filter 2 threshold, a fraction of the rows */
  int32_t pull_unroll;     /* probes in flight per lane in the per-level pulls: 0 = 4 per short-row step
                              and 2 per long-row step, 1 = 8/2, 2 = 8/4, 3 = 4/4, 4 = 2/1, 5 = 1/1,
                              6 = 2/2, 7 = 1/2 */
  int32_t lane_rule;       /* 0: one direction per level for all roots (the rule on alpha above);
                              1: a direction per root and level (per-level launches, reorder modes
                              1 and 2): a root pulls when its frontier's edges exceed 1/alpha of the
                              edges of the vertices it has not reached, else pushes; a level with
                              both kinds pushes first, then pulls and merges */
  double compact_below;    /* per-level pulls (persistent = 0) read frontier words from a compact
                              copy holding only the frontier's words when the frontier spans fewer
                              than compact_below x rows vertices; 0 disables */
  int32_t mixed_push;      /* (lane_rule 1) the push of a level where some roots push and the others
                              pull: 0 = over the CSR chunks holding one of the pushing roots'
                              frontier vertices; 1 = over those vertices as a queue, edge-balanced
                              chunks of their rows only */
  int32_t reserved;        /* must be 0 */
} g5_batch_tuning;

/* ---- context ---- */
const char* g5_version(void);
int g5_device_count(int* count);
/* Creates a context on HIP device 'device' with its own stream. */
int g5_create(int device, g5_ctx** out);
/* Frees the context and everything it holds.  Returns G5_ERR_STATE, and
 * frees nothing, while a g5_server created over the context still exists
 * (g5_server_destroy first); a NULL context is a no-op. */
int g5_destroy(g5_ctx* ctx);
/* Message of the last failing call on ctx (empty when none). */
const char* g5_last_error(const g5_ctx* ctx);
/* Waits for all work queued by ctx. */
int g5_synchronize(g5_ctx* ctx);

/* ---- edge list: EdgeListStorage<UnweightedPackedEdge, 8Mi>
 *      (generator/graph_generator.hpp:35-319) held in HBM ---- */
/* Pre-sizes the device edge list (edge_memory_size_, graph_generator.hpp:58). */
int g5_edges_reserve(g5_ctx* ctx, int64_t capacity);
/* Appends host tuples (EdgeListStorage::write, graph_generator.hpp:190-243). */
int g5_edges_append(g5_ctx* ctx, const g5_packed_edge* tuples, int64_t count);
/* Number of tuples held. */
int g5_edges_count(g5_ctx* ctx, int64_t* count);
/* Copies tuples [start, start+count) to the host (EdgeListStorage::read,
 * graph_generator.hpp:120-154). */
int g5_edges_download(g5_ctx* ctx, int64_t start, int64_t count, g5_packed_edge* out);
/* Copies tuples [start, start+count) to the host as WeightedEdge tuples
 * with the weights GraphGeneratorBase::generateWeight gives them
 * (generator/graph_generator.hpp:480-506; edge index = position in the edge
 * list): max_weight <= 0 stores 0xBEEF in every tuple, as the reference's
 * compiled branch does (:499-500); max_weight > 0 draws the per-edge weight
 * of its disabled branch (:483-488): the seed (seed1, seed2) state skipped by
 * (30, 46, edge_index), one draw, modulo max_weight, plus 1.  The weights are
 * made on the GPU. */
int g5_edges_download_weighted(g5_ctx* ctx, int64_t start, int64_t count, int max_weight, uint64_t seed1,
                               uint64_t seed2, g5_weighted_edge* out);
/* Frees the device edge list. */
int g5_edges_release(g5_ctx* ctx);
/* generate_graph_spec2010 / RmatGraphGenerator::generateRange
 * (mpi/benchmark_helper.hpp:237-314, generator/graph_generator.hpp:602-705):
 * replaces the edge list with edgefactor*2^scale Kronecker tuples, bit-exact
 * with the reference splittable-MRG generator.  The reference uses seeds
 * (2,3) and a_num=5700, bc_num=1900 (per 10000). */
int g5_generate_kronecker(g5_ctx* ctx, int scale, int edgefactor, uint64_t seed1,
                          uint64_t seed2, int a_num, int bc_num);
/* generate_graph with an initial-edge type (GraphGeneratorBase, generator/
 * graph_generator.hpp:385-393, 515-524): initial_edges 1 puts 2^scale - 1
 * binary-tree tuples first and the R-MAT tuples after them at their global
 * indices, as generate_graph_spec2012 does (mpi/benchmark_helper.hpp:316-322,
 * a_num 5500, bc_num 100; the reference's weights are not generated);
 * initial_edges 0 is g5_generate_kronecker. */
int g5_generate_graph(g5_ctx* ctx, int scale, int edgefactor, uint64_t seed1, uint64_t seed2,
                      int a_num, int bc_num, int initial_edges);

/* ---- graph: BfsBase::construct -> GraphConstructor2DCSR::construct
 *      (mpi/bfs.hpp:102-113, mpi/graph_constructor.hpp:554-574) ---- */
/* Builds the CSR from the edge list.  reorder_mode is VERTEX_REORDERING
 * (mpi/parameters.h:38-40): 2 = degree order (ties by id), 1 = vertices with
 * edges first, 0 = original order.  release_edges != 0 frees the device edge
 * list afterwards (EdgeListStorage::beginRead(true), graph_generator.hpp:98).
 * G5_ERR_STATE while a g5_server serves the context's graph. */
int g5_construct(g5_ctx* ctx, int reorder_mode, int release_edges);
int g5_graph_info_get(g5_ctx* ctx, g5_graph_info* out);
/* Graph2DCSR::has_edge (mpi/graph_constructor.hpp:101-110): 1, 0 or < 0. */
int g5_has_edge(g5_ctx* ctx, int64_t vertex);
/* Graph2DCSR::pred_size (mpi/graph_constructor.hpp:64). */
int g5_pred_size(g5_ctx* ctx, int64_t* size);
/* find_roots (mpi/benchmark_helper.hpp:475-508): writes up to num_roots roots. */
int g5_find_roots(g5_ctx* ctx, uint64_t seed1, uint64_t seed2, int num_roots, int64_t* roots,
                  int* num_found);
/* Copies the canonical CSR to the host (for parity checks).  Any pointer may
 * be NULL.  offsets: num_rows+1 (num_rows = num_active for modes 1, 2 and
 * num_verts for mode 0), cols: nnz, new2old: num_verts, weight: num_rows. */
int g5_graph_download(g5_ctx* ctx, uint64_t* offsets, uint32_t* cols, uint32_t* new2old,
                      uint32_t* weight);

/* ---- single-root BFS: BfsBase (mpi/bfs.hpp:115-128, run_bfs :2651) ---- */
/* prepare_bfs: allocates BFS state (mpi/bfs.hpp:115-118). */
int g5_prepare_bfs(g5_ctx* ctx);
/* run_bfs(root, pred): direction-optimizing BFS from an original vertex id.
 * Synchronous: returns after the device finished, so host timing brackets it
 * like mpi/main.cc:113-117.  Levels and parents stay in HBM. */
int g5_run_bfs(g5_ctx* ctx, int64_t root);
/* get_pred (mpi/bfs.hpp:122-124): writes pred_size() entries of
 * parent | level << 48, or -1 when unreached (encoding of mpi/bfs.hpp:1203,
 * 1271, 2470). */
int g5_get_pred(g5_ctx* ctx, int64_t* pred);
/* uint8 level per original vertex id, 0xFF when unreached. */
int g5_get_levels(g5_ctx* ctx, uint8_t* levels);
/* Traversed-edge count of the last run: input tuples with a reached endpoint,
 * duplicates and self loops included (mpi/validate.hpp:703-721). */
int g5_teps_edges(g5_ctx* ctx, int64_t* count);
int g5_bfs_stats_get(g5_ctx* ctx, g5_bfs_stats* out);
/* Algorithmic HBM bytes of the last run, per level and in total: top-down
 * level = sum over the frontier of (4 + 8 + 4 deg) + 4 per discovered vertex;
 * bottom-up level = 3 n / 8 + sum over unvisited vertices with edges of
 * (8 + 4 k), k = neighbours probed in ascending order up to the first visited
 * one; plus 1 byte per reached vertex.  No reference counterpart: it is the
 * byte model the roofline figures are computed with (evaluated on the GPU
 * from the final levels, outside any timed region). per_level has
 * G5_MAX_LEVELS entries, index = level. */
int g5_bfs_model_bytes(g5_ctx* ctx, double* per_level, double* total);
/* Single-root tuning (no counterpart in the reference; levels do not depend
 * on it, parents may).  All knobs but small_edges act on bulk top-down levels
 * (more than 2^20 frontier edges); small_edges acts on the small ones.
 *   bin_edges: levels with at least this many frontier edges first group
 *     their (target, source) pairs by target range (0 = never, the default:
 *     twice as slow as the plain level at SCALE 26-28).
 *   td_pieces: while the search grows, rows are cut into this many pieces by
 *     position and taken piece index first (0 = row order, the default).
 *   block_shift: while the search grows, edges are taken by target block of
 *     2^block_shift ids through a block index of the first rows built at the
 *     first run (0 = off, the default; 12..31).
 *   td_head: while the search grows, targets whose degree is at least
 *     td_head / 16 * (nnz / E)^2 (E = the level's frontier edges; a prefix of
 *     the degree-ordered rows) are settled bottom-up first and the expansion
 *     walks only the row suffixes beyond them (default 4; 0 = off; degree
 *     order only).
 *   seg_max: most segments one segment-list level may hold (0 = the count
 *     field's limit, 2^24 - 1, the default; 1..2^24 - 1 lower it).  A level
 *     whose frontier rows (suffix mode) or pieces would exceed it is taken
 *     in plain row order instead.  Only tests lower it, to reach that path
 *     on small graphs.
 *   small_edges: consecutive top-down levels whose frontiers have at most
 *     4096 vertices and at most small_edges edges run in one single-workgroup
 *     launch that applies the direction rule itself, with no host round trip
 *     between them (default 8192; 0 = off; at most 2^20).
 *   tail_pull: a top-down level after the bottom-up phase whose frontier
 *     holds more than 4096 vertices and more than 1/tail_pull as many as
 *     remain unvisited finds its vertices with the bottom-up kernel (every unvisited vertex looks
 *     for a visited neighbour, which can only be on the frontier); the level
 *     keeps its top-down place in the schedule, its vertex set and its
 *     counts (default 4; 0 = off; at most 64). */
typedef struct g5_bfs_tuning {
  int64_t bin_edges;
  int64_t td_pieces;
  int64_t block_shift;
  int64_t td_head;
  int64_t seg_max;
  int64_t small_edges;
  int64_t tail_pull;
} g5_bfs_tuning;
int g5_bfs_set_tuning(g5_ctx* ctx, const g5_bfs_tuning* tuning);
int g5_bfs_get_tuning(g5_ctx* ctx, g5_bfs_tuning* tuning);

/* Validates the last run on the GPU with the rules of BfsValidation::validate
 * (mpi/validate.hpp:506-802), checking every edge of the CSR (all non-loop
 * input tuples, both directions).  errors[6] receives the number of
 * violations of: root is its own parent at level 0, no other self parent,
 * level = parent level + 1, edge endpoints both (un)reached with levels at
 * most one apart, parent is a neighbour, value in range.  *valid is 1 when
 * all are zero. */
int g5_validate(g5_ctx* ctx, int64_t* errors, int* valid);
/* The same checks on a caller's pred array of pred_size() entries
 * (validate_bfs_result, mpi/validate.hpp:1008-1019), with the traversed-edge
 * count of its reached set in *edge_visit_count (may be NULL). */
int g5_validate_pred(g5_ctx* ctx, int64_t root, const int64_t* pred, int64_t* errors, int* valid,
                     int64_t* edge_visit_count);
/* g5_validate_pred without any device allocation: no counterpart in the
 * reference.  It is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): its staging, level, parent and counter buffers live in one
 * 64 MiB array per GPU that the library holds until the process ends, so it
 * serves graphs of up to about 2^22 vertex ids (pred_size() * 8 + 5 bytes
 * per row, plus alignment, within 64 MiB); larger graphs get G5_ERR_LIMIT.
 * Same checks, outputs and errors as g5_validate_pred.  Calls that overlap on
 * one GPU share that array and race: serialize them. */
int g5_validate_pred_noalloc(g5_ctx* ctx, int64_t root, const int64_t* pred, int64_t* errors, int* valid,
                             int64_t* edge_visit_count);
/* end_bfs (mpi/bfs.hpp:126-128): frees BFS state. */
int g5_end_bfs(g5_ctx* ctx);

/* ---- batched BFS: up to 64 roots in one bit-packed sweep ----
 * No counterpart in the reference, which runs its roots one after another
 * (mpi/main.cc:105-150): this is the batch form of BfsBase::run_bfs whose
 * per-root levels equal g5_run_bfs's for the same root. */
int g5_batch_prepare(g5_ctx* ctx);
int g5_batch_set_tuning(g5_ctx* ctx, const g5_batch_tuning* tuning);
int g5_batch_get_tuning(g5_ctx* ctx, g5_batch_tuning* tuning);
/* Runs 1..64 roots (original ids) at once; synchronous. */
int g5_run_bfs_batch(g5_ctx* ctx, const int64_t* roots, int num_roots);
/* uint8 level per original vertex id for root 'lane' of the last batch. */
int g5_batch_get_levels(g5_ctx* ctx, int lane, uint8_t* levels);
/* Traversed-edge count of every root of the last batch (num_roots values). */
int g5_batch_teps_edges(g5_ctx* ctx, int64_t* counts);
int g5_batch_stats_get(g5_ctx* ctx, g5_batch_stats* out);
/* Algorithmic bytes of the last batch under the bit-packed pull model: per
 * level 24 n + sum over vertices lacking some root of (8 + 12 k), k = probes
 * until every lacking root is found.  Under lane_rule 1 only the roots pulled
 * at a level count as lacking, and every vertex whose frontier word holds a
 * pushed root adds 8 + 12 deg.  The batch's output is its per-level frontier
 * words (the write inside 24 n); a root's uint8 levels are unpacked from them
 * by g5_batch_get_levels, outside the batch, so no per-(root, vertex) output
 * byte is counted. */
int g5_batch_model_bytes(g5_ctx* ctx, double* per_level, double* total);
/* Per level of the last batch (G5_MAX_LEVELS entries each, index = level):
 * the roots that pushed and the roots costed as pulling in the byte model
 * (no counterpart in the reference; with lane_rule 0 no root pushes and every
 * root counts as pulling). */
int g5_batch_level_lanes(g5_ctx* ctx, uint64_t* push_lanes, uint64_t* pull_lanes);
int g5_batch_end(g5_ctx* ctx);

/* ---- serving: continuous batching over resident graphs ----
 * (no counterpart in the reference; every answer equals run_bfs's levels
 *  for the same root, mpi/bfs.hpp:2651-2940)
 * A server drives up to 64 graph contexts (each built with reorder mode 1 or
 * 2, all on one device); each graph runs the 64-lane batched BFS one level at
 * a time on its own stream.  At every level boundary finished queries leave
 * their lane and waiting queries take free lanes, while the other lanes go on
 * at whatever level they have reached.  A query holds one answer array
 * (num_rows bytes of levels) from a pool of answer_buffers per graph; a query
 * waits while its graph has no free lane or no free answer array.  With
 * config solo, a query that finds its graph idle and few queries waiting is
 * answered by the single-root kernels instead (see g5_server_config.solo).  The graph
 * contexts must not be used for anything else while a server runs. */
typedef struct g5_server g5_server;

typedef struct g5_server_config {
  int32_t answer_buffers;  /* answer arrays preallocated per graph (HBM admission budget) */
  int32_t pull_filter;     /* as g5_batch_tuning.pull_filter */
  double alpha;            /* lane_rule 1: a search pulls when its frontier has more than nnz/alpha
                              edges, else pushes; lane_rule 3: when its frontier's edges exceed
                              1/alpha of the edges of the vertices it has not reached */
  int32_t min_admit;       /* a graph with searches in flight admits waiting queries into its free
                              lanes at a level boundary while every search in flight still pushes,
                              or once min(min_admit, free answer arrays) of them can start together
                              (searches started together share their pulling levels; 1 admits every
                              query at the first boundary with a free lane) ... */
  int32_t lane_rule;       /* 0: each search switches direction by the reference rule on its own
                              frontier size (mpi/bfs.hpp:2799-2829); 1: by its frontier's edges;
                              2: all searches together, by the frontier union's edges; 3: by its
                              frontier's edges against its unreached edges (as g5_batch_tuning
                              lane_rule 1) */
  double max_wait_s;       /* ... or once its oldest waiting query has waited this long.  An idle
                              graph admits at once. */
  int32_t solo;            /* 1..64: a graph with no search in flight and at most solo waiting
                              queries answers the oldest with the single-root kernels of
                              g5_run_bfs (a host thread of the server drives them, so the other
                              graphs go on), then looks again; a longer queue is batched into
                              lanes.  0: every query takes a lane. */
  int32_t solo_slots;      /* single-root runs of one graph that may go on at once (0 or 1: one;
                              at most 8); each slot beyond the first holds a BFS state of its own
                              (as g5_prepare_bfs allocates) and a stream */
} g5_server_config;

typedef struct g5_query {
  int32_t graph;     /* index into the server's graphs */
  int32_t keep;      /* 1: keep the answer resident until g5_server_release */
  int64_t root;      /* original vertex id */
  double arrival_s;  /* arrival time, seconds after g5_server_run starts (non-decreasing) */
} g5_query;

typedef struct g5_query_result {
  int64_t request;   /* id for g5_server_get_levels / g5_server_release */
  double admit_s;    /* when the query took a lane */
  double done_s;     /* when its answer was complete (latency = done_s - arrival_s) */
  int64_t edges;     /* traversed edges, as the validator counts them */
  int32_t levels;    /* levels its lane took part in */
  int32_t status;    /* 2 = done */
} g5_query_result;

typedef struct g5_server_report {
  double wall_s;            /* from the start of the run to the last answer */
  int64_t completed;
  int64_t edges;            /* sum of traversed edges */
  int64_t levels_launched;  /* engine levels over all graphs */
  int32_t max_active_lanes; /* most lanes active at once in one graph */
  int32_t reserved;
  int64_t admitted_busy;    /* queries admitted into a graph that had searches in flight */
  int64_t admit_events;     /* level boundaries at which some query was admitted */
} g5_server_report;

/* Per-graph engine counters since the server was created. */
typedef struct g5_server_graph_stats {
  int64_t levels;         /* levels launched */
  int64_t push_levels;    /* levels with top-down lanes only */
  int64_t pull_levels;    /* levels with bottom-up lanes only */
  int64_t mixed_levels;   /* levels with both */
  int64_t lane_levels;    /* sum over levels of the active lanes */
  double busy_ms;         /* device time of the levels (HIP events) */
  int64_t completed;      /* queries answered */
  int64_t admitted_busy;  /* queries admitted while other searches of this graph were in flight */
  int64_t admitted_idle;  /* queries admitted into an idle graph */
  int64_t solo;           /* queries answered by the single-root kernels (g5_server_config.solo) */
  int64_t solo_overlap;   /* of those, runs started while another single-root run of this graph
                             was going on (g5_server_config.solo_slots) */
} g5_server_graph_stats;

int g5_server_create(g5_ctx* const* graphs, int num_graphs, const g5_server_config* config, g5_server** out);
int g5_server_graph_stats_get(g5_server* server, int graph, g5_server_graph_stats* out);
void g5_server_destroy(g5_server* server);
const char* g5_server_last_error(const g5_server* server);
/* Open-loop run: query i is submitted at arrival_s; returns when every query
 * has its answer (or fails with G5_ERR_LIMIT after timeout_s). */
int g5_server_run(g5_server* server, const g5_query* queries, int64_t num_queries, double timeout_s,
                  g5_query_result* results, g5_server_report* report);
/* Levels of a kept answer by original id (0xFF unreached), pred_size() entries. */
int g5_server_get_levels(g5_server* server, int64_t request, uint8_t* levels);
int g5_server_release(g5_server* server, int64_t request);

#ifdef __cplusplus
}
#endif

#endif
