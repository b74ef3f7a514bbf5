// g5_bfs.hpp -- C++ drop-in for the reference's BFS class over the g5 C ABI.
//
// g5::BfsOnGPU has the members main.cc and benchmark_helper.hpp use on
// BfsOnCPU / BfsBase (mpi/bfs.hpp:29-130, mpi/bfs_cpu.hpp:13, used by
// mpi/main.cc:58-177):
//   construct(EdgeList*)          mpi/bfs.hpp:102-113
//   prepare_bfs()                 mpi/bfs.hpp:115-118
//   run_bfs(root, pred)           mpi/bfs.hpp:120, :2651
//   get_pred(pred)                mpi/bfs.hpp:122-124
//   end_bfs()                     mpi/bfs.hpp:126-128
//   graph_.has_edge(v), graph_.pred_size(), graph_.log_orig_global_verts_,
//   graph_.num_global_verts_      mpi/graph_constructor.hpp:64-110, :113-143
// Errors are thrown as `const char*`, like throw_exception
// (mpi/utils.hpp:264-276); nothing else throws.
//
// run_bfs returns once the device has finished the traversal, so a host
// clock around it brackets the whole BFS (mpi/main.cc:120-124).  The
// reference fills pred inside run_bfs; here the traversal leaves levels and
// parents in HBM and get_pred(pred) writes the caller's array, which
// main.cc calls before validation and outside the timed region
// (mpi/main.cc:127).  run_bfs also records pred so that get_pred() with no
// argument writes there.
#ifndef G5_BFS_HPP
#define G5_BFS_HPP

#include <stdint.h>

#include <string>

#include "g5.h"

namespace g5 {

class BfsOnGPU;

// The part of Graph2DCSR (mpi/graph_constructor.hpp:28-143) that the driver
// and the root search read.
class GraphView {
 public:
  bool has_edge(int64_t v, bool has_weight = false) const {
    (void)has_weight;
    const int r = g5_has_edge(ctx_, v);
    if (r < 0) throw_last();
    return r == 1;
  }
  int64_t pred_size() const { return int64_t(1) << log_orig_global_verts_; }
  int log_orig_global_verts() const { return log_orig_global_verts_; }

  int log_orig_global_verts_ = 0;  // estimated SCALE
  int64_t num_global_verts_ = 0;   // vertices with at least one edge
  int64_t num_global_edges_ = 0;   // directed CSR entries after duplicate removal
  int64_t max_used_vertex_ = -1;   // find_max_used_vertex (mpi/benchmark_helper.hpp:510-528)

 private:
  friend class BfsOnGPU;
  void throw_last() const {
    static thread_local std::string msg;
    msg = g5_last_error(ctx_);
    throw msg.c_str();
  }
  g5_ctx* ctx_ = nullptr;
};

class BfsOnGPU {
 public:
  // VERTEX_REORDERING (mpi/parameters.h:38-40); the reference build uses 2.
  explicit BfsOnGPU(int device = 0, int reorder_mode = 2) : reorder_mode_(reorder_mode) {
    if (g5_create(device, &ctx_) != G5_OK) throw "g5_create: no such HIP device, or the HIP runtime failed";
    graph_.ctx_ = ctx_;
  }
  virtual ~BfsOnGPU() {
    if (ctx_) g5_destroy(ctx_);
  }
  BfsOnGPU(const BfsOnGPU&) = delete;
  BfsOnGPU& operator=(const BfsOnGPU&) = delete;

  static void printInformation() {}

  // Streams the caller's edge list (EdgeListStorage<UnweightedPackedEdge,..>
  // interface: beginRead / read / endRead, generator/graph_generator.hpp:
  // 98-180) to the device in its own chunks, then builds the CSR there.  The
  // caller keeps its edge list (mpi/main.cc:45-47).
  template <typename EdgeList>
  void construct(EdgeList* edge_list) {
    typedef typename EdgeList::edge_type EdgeType;
    static_assert(sizeof(EdgeType) == sizeof(g5_packed_edge), "edge tuples must be the 12-byte packed layout");
    check(g5_edges_release(ctx_));
    edge_list->beginRead(false);
    EdgeType* buf = nullptr;
    int got;
    while ((got = edge_list->read(&buf)) > 0)
      check(g5_edges_append(ctx_, reinterpret_cast<const g5_packed_edge*>(buf), got));
    edge_list->endRead();
    build(true);
  }

  // Tuples already in host memory (one call, any count).
  void construct(const g5_packed_edge* tuples, int64_t count) {
    check(g5_edges_release(ctx_));
    check(g5_edges_append(ctx_, tuples, count));
    build(true);
  }

  // generate_graph_spec2010 on the device, then construct: the tuples never
  // leave HBM (mpi/benchmark_helper.hpp:237-314).
  void generate_kronecker(int scale, int edgefactor, uint64_t seed1 = 2, uint64_t seed2 = 3) {
    check(g5_generate_kronecker(ctx_, scale, edgefactor, seed1, seed2, 5700, 1900));
  }
  void construct_generated(bool release_edges = true) { build(release_edges); }

  void prepare_bfs() { check(g5_prepare_bfs(ctx_)); }

  void run_bfs(int64_t root, int64_t* pred) {
    check(g5_run_bfs(ctx_, root));
    pred_ = pred;
  }

  void get_pred(int64_t* pred) { check(g5_get_pred(ctx_, pred)); }
  void get_pred() {
    if (!pred_) throw "get_pred: no pred array recorded by run_bfs";
    get_pred(pred_);
  }

  void end_bfs() { check(g5_end_bfs(ctx_)); }

  // Roots as find_roots (mpi/benchmark_helper.hpp:475-508) draws them.
  int find_roots(int64_t* roots, int num_roots, uint64_t seed1 = 2, uint64_t seed2 = 3) {
    int found = 0;
    check(g5_find_roots(ctx_, seed1, seed2, num_roots, roots, &found));
    return found;
  }

  // validate_bfs_result (mpi/validate.hpp:1008-1019) on the GPU, against the
  // resident graph; returns true when pred passes every rule.
  bool validate(int64_t root, const int64_t* pred, int64_t* edge_visit_count) {
    int64_t errors[6];
    int valid = 0;
    check(g5_validate_pred(ctx_, root, pred, errors, &valid, edge_visit_count));
    return valid == 1;
  }

  g5_ctx* context() const { return ctx_; }

  GraphView graph_;

 private:
  void build(bool release_edges) {
    check(g5_construct(ctx_, reorder_mode_, release_edges ? 1 : 0));
    g5_graph_info info;
    check(g5_graph_info_get(ctx_, &info));
    graph_.log_orig_global_verts_ = info.scale_estimate;
    graph_.num_global_verts_ = info.num_active;
    graph_.num_global_edges_ = info.nnz;
    graph_.max_used_vertex_ = info.max_used_vertex;
  }
  void check(int rc) const {
    if (rc != G5_OK) {
      static thread_local std::string msg;
      msg = g5_last_error(ctx_);
      throw msg.c_str();
    }
  }

  g5_ctx* ctx_ = nullptr;
  int reorder_mode_;
  int64_t* pred_ = nullptr;
};

}  // namespace g5

#endif  // G5_BFS_HPP
