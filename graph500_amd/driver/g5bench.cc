// g5bench -- the Graph500 BFS benchmark flow of graph500_bfs
// (mpi/main.cc:34-178) on one GPU through g5::BfsOnGPU.
//
//   g5bench SCALE [edgefactor] [--roots N] [--host-edges] [--reorder M]
//
// Generation (generate_graph_spec2010, mpi/benchmark_helper.hpp:237-314) runs
// on the device; --host-edges instead downloads the tuples and streams them
// back through the EdgeListStorage-style construct(EdgeList*) entry, the path
// a caller holding its own edge list takes.  Each root's run_bfs is timed by
// the host clock (mpi/main.cc:118-124), get_pred fills the caller's pred
// array, and validate_bfs_result's rules (mpi/validate.hpp:1008-1019) run on
// the GPU against the resident graph.  The result block has the keys of
// print_bfs_result (mpi/logfile.h:47-100).
#include <inttypes.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <vector>

#include "g5_bfs.hpp"

namespace {

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Minimum, quartiles, median, maximum, mean and sample standard deviation,
// with the quartile convention of the Graph500 specification
// (mpi/logfile.h:19-45).
struct Stats {
  double min, q1, median, q3, max, mean, std;
};

Stats statistics(const std::vector<double>& x) {
  const size_t n = x.size();
  Stats s{};
  double sum = 0;
  for (double v : x) sum += v;
  s.mean = sum / n;
  double var = 0;
  for (double v : x) var += (v - s.mean) * (v - s.mean);
  s.std = n > 1 ? sqrt(var / (n - 1)) : 0.0;
  std::vector<double> y(x);
  std::sort(y.begin(), y.end());
  auto mid = [&](size_t a, size_t b) { return 0.5 * (y[a] + y[b]); };
  s.min = y.front();
  s.q1 = mid((n - 1) / 4, n / 4);
  s.median = mid((n - 1) / 2, n / 2);
  s.q3 = mid(n - 1 - (n - 1) / 4, n - 1 - n / 4);
  s.max = y.back();
  return s;
}

void print_block(const char* what, const Stats& s, const char* fmt) {
  const char* names[] = {"min", "firstquartile", "median", "thirdquartile", "max", "mean", "stddev"};
  const double vals[] = {s.min, s.q1, s.median, s.q3, s.max, s.mean, s.std};
  for (int i = 0; i < 7; ++i) {
    char key[64];
    snprintf(key, sizeof key, "%s_%s:", names[i], what);
    printf("%-32s", key);
    printf(fmt, vals[i]);
    printf("\n");
  }
}

// A host edge list with the EdgeListStorage read interface
// (generator/graph_generator.hpp:98-180), chunks of 8 Mi tuples.
class HostEdgeList {
 public:
  typedef g5_packed_edge edge_type;
  static const int CHUNK_SIZE = 8 * 1024 * 1024;
  explicit HostEdgeList(std::vector<g5_packed_edge>&& t) : t_(std::move(t)) {}
  int beginRead(bool) {
    pos_ = 0;
    return (int)((t_.size() + CHUNK_SIZE - 1) / CHUNK_SIZE);
  }
  int read(edge_type** p) {
    if (pos_ >= t_.size()) return 0;
    const size_t k = std::min<size_t>(CHUNK_SIZE, t_.size() - pos_);
    *p = t_.data() + pos_;
    pos_ += k;
    return (int)k;
  }
  void endRead() {}

 private:
  std::vector<g5_packed_edge> t_;
  size_t pos_ = 0;
};

int run(int argc, char** argv) {
  if (argc < 2) {
    fprintf(stderr, "usage: %s SCALE [edgefactor] [--roots N] [--host-edges] [--reorder M]\n", argv[0]);
    return 2;
  }
  const int scale = atoi(argv[1]);
  int edgefactor = 16, num_roots = 64, reorder = 2;
  bool host_edges = false;
  for (int i = 2; i < argc; ++i) {
    if (!strcmp(argv[i], "--roots") && i + 1 < argc) num_roots = atoi(argv[++i]);
    else if (!strcmp(argv[i], "--reorder") && i + 1 < argc) reorder = atoi(argv[++i]);
    else if (!strcmp(argv[i], "--host-edges")) host_edges = true;
    else if (argv[i][0] != '-') edgefactor = atoi(argv[i]);
    else {
      fprintf(stderr, "unknown option %s\n", argv[i]);
      return 2;
    }
  }
  if (num_roots < 1 || num_roots > 64) {
    fprintf(stderr, "--roots must be in [1, 64]\n");
    return 2;
  }

  g5::BfsOnGPU bfs(0, reorder);
  printf("Graph generation\n");
  double generation_time = now_s();
  bfs.generate_kronecker(scale, edgefactor);
  g5_synchronize(bfs.context());
  generation_time = now_s() - generation_time;

  printf("Graph construction\n");
  double construction_time;
  if (host_edges) {
    int64_t m = 0;
    g5_edges_count(bfs.context(), &m);
    std::vector<g5_packed_edge> t((size_t)m);
    if (g5_edges_download(bfs.context(), 0, m, t.data()) != G5_OK) throw g5_last_error(bfs.context());
    HostEdgeList list(std::move(t));
    construction_time = now_s();
    bfs.construct(&list);
  } else {
    construction_time = now_s();
    bfs.construct_generated(true);
  }
  g5_synchronize(bfs.context());
  construction_time = now_s() - construction_time;

  std::vector<int64_t> roots(num_roots);
  num_roots = bfs.find_roots(roots.data(), num_roots);
  const int64_t nlocalverts = bfs.graph_.pred_size();
  std::vector<int64_t> pred((size_t)nlocalverts, -1);

  bfs.prepare_bfs();
  std::vector<double> bfs_times(num_roots), validate_times(num_roots), edge_counts(num_roots);
  bool result_ok = true;
  for (int i = 0; i < num_roots; ++i) {
    printf("========== Running BFS %d ==========\n", i);
    double t = now_s();
    bfs.run_bfs(roots[i], pred.data());
    bfs_times[i] = now_s() - t;
    printf("Time for BFS %d is %f\n", i, bfs_times[i]);
    bfs.get_pred(pred.data());
    t = now_s();
    int64_t edge_visit_count = 0;
    result_ok = bfs.validate(roots[i], pred.data(), &edge_visit_count);
    validate_times[i] = now_s() - t;
    edge_counts[i] = (double)edge_visit_count;
    printf("Validate time for BFS %d is %f\n", i, validate_times[i]);
    printf("Number of traversed edges is %" PRId64 "\n", edge_visit_count);
    printf("TEPS for BFS %d is %g\n", i, edge_visit_count / bfs_times[i]);
    if (!result_ok) {
      printf("Validation failed for BFS %d (root %" PRId64 ")\n", i, roots[i]);
      break;
    }
  }
  bfs.end_bfs();

  printf("============= Result ==============\n");
  printf("SCALE:                          %d\n", scale);
  printf("edgefactor:                     %d\n", edgefactor);
  printf("NBFS:                           %d\n", num_roots);
  printf("graph_generation:               %g\n", generation_time);
  printf("num_mpi_processes:              %d\n", 1);
  printf("construction_time:              %g\n", construction_time);
  printf("redistribution_time:            %g\n", 0.0);
  if (!result_ok) {
    printf("No results printed for invalid run.\n");
    return 1;
  }
  print_block("time", statistics(bfs_times), "%.12g");
  print_block("nedge", statistics(edge_counts), "%.11g");
  std::vector<double> spe(num_roots);
  for (int i = 0; i < num_roots; ++i) spe[i] = bfs_times[i] / edge_counts[i];
  const Stats s = statistics(spe);
  printf("%-32s%.12g\n", "min_TEPS:", 1.0 / s.max);
  printf("%-32s%.12g\n", "firstquartile_TEPS:", 1.0 / s.q3);
  printf("%-32s%.12g\n", "median_TEPS:", 1.0 / s.median);
  printf("%-32s%.12g\n", "thirdquartile_TEPS:", 1.0 / s.q1);
  printf("%-32s%.12g\n", "max_TEPS:", 1.0 / s.min);
  printf("%-32s%.12g\n", "harmonic_mean_TEPS:", 1.0 / s.mean);
  // standard error of the harmonic mean (Norris 1940), as the specification asks
  printf("%-32s%.12g\n", "harmonic_stddev_TEPS:", s.std / (s.mean * s.mean * sqrt((double)num_roots - 1)));
  print_block("validate", statistics(validate_times), "%.12g");
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  try {
    return run(argc, argv);
  } catch (const char* msg) {
    fprintf(stderr, "g5bench: %s\n", msg);
    return 1;
  }
}
