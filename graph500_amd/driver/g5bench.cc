// g5bench -- the Graph500 BFS benchmark flow of graph500_bfs
// (mpi/main.cc:34-178) on one GPU through g5::BfsOnGPU.
//
//   g5bench SCALE [edgefactor] [--roots N] [--host-edges] [--reorder M]
//   g5bench --read-log FILE
//
// Generation (generate_graph_spec2010, mpi/benchmark_helper.hpp:237-314) runs
// on the device; --host-edges instead downloads the tuples and streams them
// back through the EdgeListStorage-style construct(EdgeList*) entry, the path
// a caller holding its own edge list takes.  Each root's run_bfs is timed by
// the host clock (mpi/main.cc:118-124), get_pred fills the caller's pred
// array, and validate_bfs_result's rules (mpi/validate.hpp:1008-1019) run on
// the GPU against the resident graph.  The result block has the keys of
// print_bfs_result (mpi/logfile.h:47-100).
//
// Environment, as the reference driver reads it:
//   LOGFILE  per-root results are kept in this file (the reference's binary
//            LogFileFormat, mpi/logfile.h:109-124) after every root; a run
//            whose SCALE and edgefactor match a file's resumes after its
//            last root (read_log_file / update_log_file,
//            mpi/benchmark_helper.hpp:530-604).  --read-log prints such a
//            file as the reference's logreader does (mpi/logreader.cc).
//   TMPFILE  with --host-edges, the edge list is kept in the file
//            TMPFILE-000 instead of host memory, written and read in 8 Mi
//            tuple chunks and removed at the end (EdgeListStorage with a
//            file, generator/graph_generator.hpp:41-83).
#include <inttypes.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <string>
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

// The reference's log record: the same fields in the same order, so files
// are interchangeable with its runnable and logreader.
struct LogTime {
  double bfs_time;
  double validate_time;
  int64_t edge_counts;
};
struct LogFile {
  int32_t scale;
  int32_t edge_factor;
  int32_t mpi_size;
  int32_t num_runs;
  double generation_time;
  double construction_time;
  double redistribution_time;
  LogTime times[64];
};
static_assert(sizeof(LogFile) == 16 + 24 + 64 * 24, "log record layout");

uint32_t bswap32(uint32_t v) { return __builtin_bswap32(v); }
uint64_t bswap64(uint64_t v) { return __builtin_bswap64(v); }
template <class T> void bswap_field(T& x) {
  static_assert(sizeof(T) == 4 || sizeof(T) == 8, "field size");
  if (sizeof(T) == 4) {
    uint32_t u;
    memcpy(&u, &x, 4);
    u = bswap32(u);
    memcpy(&x, &u, 4);
  } else {
    uint64_t u;
    memcpy(&u, &x, 8);
    u = bswap64(u);
    memcpy(&x, &u, 8);
  }
}

// A record written on a machine of the other byte order has an implausible
// run count; it is swapped field by field (mpi/logreader.cc).
void fix_endian(LogFile& log) {
  if (log.num_runs >= 0 && log.num_runs <= 64) return;
  bswap_field(log.scale);
  bswap_field(log.edge_factor);
  bswap_field(log.mpi_size);
  bswap_field(log.num_runs);
  bswap_field(log.generation_time);
  bswap_field(log.construction_time);
  bswap_field(log.redistribution_time);
  for (auto& t : log.times) {
    bswap_field(t.bfs_time);
    bswap_field(t.validate_time);
    bswap_field(t.edge_counts);
  }
}

bool load_log(const char* path, LogFile& log) {
  FILE* fp = fopen(path, "rb");
  if (!fp) return false;
  const bool ok = fread(&log, sizeof log, 1, fp) == 1;
  fclose(fp);
  if (ok) fix_endian(log);
  return ok && log.num_runs >= 0 && log.num_runs <= 64;
}

void save_log(const char* path, const LogFile& log) {
  FILE* fp = fopen(path, "wb");
  if (!fp) {
    printf("Cannot create log file ... skipping\n");
    return;
  }
  fwrite(&log, sizeof log, 1, fp);
  fclose(fp);
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

// The same interface over a file of packed tuples (the reference's TMPFILE
// form): only one chunk is in host memory at a time.
class FileEdgeList {
 public:
  typedef g5_packed_edge edge_type;
  static const int CHUNK_SIZE = 8 * 1024 * 1024;
  explicit FileEdgeList(const std::string& path) : path_(path) {}
  ~FileEdgeList() {
    if (fp_) fclose(fp_);
    remove(path_.c_str());
  }
  // Streams the context's generated tuples into the file chunk by chunk.
  void write_from(g5_ctx* ctx) {
    int64_t m = 0;
    g5_edges_count(ctx, &m);
    fp_ = fopen(path_.c_str(), "w+b");
    if (!fp_) throw "FileEdgeList: cannot create the edge file";
    buf_.resize(CHUNK_SIZE);
    for (int64_t off = 0; off < m; off += CHUNK_SIZE) {
      const int64_t k = std::min<int64_t>(CHUNK_SIZE, m - off);
      if (g5_edges_download(ctx, off, k, buf_.data()) != G5_OK) throw g5_last_error(ctx);
      if (fwrite(buf_.data(), sizeof(edge_type), (size_t)k, fp_) != (size_t)k) throw "FileEdgeList: write failed";
    }
    count_ = m;
  }
  int beginRead(bool) {
    rewind(fp_);
    return (int)((count_ + CHUNK_SIZE - 1) / CHUNK_SIZE);
  }
  int read(edge_type** p) {
    const size_t k = fread(buf_.data(), sizeof(edge_type), CHUNK_SIZE, fp_);
    *p = buf_.data();
    return (int)k;
  }
  void endRead() {}

 private:
  std::string path_;
  FILE* fp_ = nullptr;
  std::vector<edge_type> buf_;
  int64_t count_ = 0;
};


// The statistics part of print_bfs_result (mpi/logfile.h:47-100).
void print_results(const std::vector<double>& bfs_times, const std::vector<double>& validate_times,
                   const std::vector<double>& edge_counts) {
  const size_t n = bfs_times.size();
  print_block("time", statistics(bfs_times), "%.12g");
  print_block("nedge", statistics(edge_counts), "%.11g");
  std::vector<double> spe(n);
  for (size_t i = 0; i < n; ++i) spe[i] = bfs_times[i] / edge_counts[i];
  const Stats s = statistics(spe);
  printf("%-32s%.12g\n", "min_TEPS:", 1.0 / s.max);
  printf("%-32s%.12g\n", "firstquartile_TEPS:", 1.0 / s.q3);
  printf("%-32s%.12g\n", "median_TEPS:", 1.0 / s.median);
  printf("%-32s%.12g\n", "thirdquartile_TEPS:", 1.0 / s.q1);
  printf("%-32s%.12g\n", "max_TEPS:", 1.0 / s.min);
  printf("%-32s%.12g\n", "harmonic_mean_TEPS:", 1.0 / s.mean);
  // standard error of the harmonic mean (Norris 1940), as the specification asks
  printf("%-32s%.12g\n", "harmonic_stddev_TEPS:", n > 1 ? s.std / (s.mean * s.mean * sqrt((double)n - 1)) : 0.0);
  print_block("validate", statistics(validate_times), "%.12g");
}

int read_log_mode(const char* path) {
  LogFile log;
  if (!load_log(path, log)) {
    fprintf(stderr, "File not found.\n");
    return 1;
  }
  std::vector<double> bfs_times(log.num_runs), validate_times(log.num_runs), edge_counts(log.num_runs);
  for (int i = 0; i < log.num_runs; ++i) {
    printf("Time for BFS %d is %f\n", i, log.times[i].bfs_time);
    printf("Validate time for BFS %d is %f\n", i, log.times[i].validate_time);
    printf("TEPS for BFS %d is %g\n", i, log.times[i].edge_counts / log.times[i].bfs_time);
    bfs_times[i] = log.times[i].bfs_time;
    validate_times[i] = log.times[i].validate_time;
    edge_counts[i] = (double)log.times[i].edge_counts;
  }
  printf("=========== %sResult ============\n", log.num_runs == 64 ? "" : "Current ");
  printf("SCALE:                          %d\n", log.scale);
  printf("edgefactor:                     %d\n", log.edge_factor);
  printf("graph_generation:               %.12g\n", log.generation_time);
  printf("num_mpi_processes:              %d\n", log.mpi_size);
  printf("construction_time:              %.12g\n", log.construction_time);
  printf("redistribution_time:            %.12g\n", log.redistribution_time);
  if (log.num_runs > 0) print_results(bfs_times, validate_times, edge_counts);
  return 0;
}

int run(int argc, char** argv) {
  if (argc < 2) {
    fprintf(stderr, "usage: %s SCALE [edgefactor] [--roots N] [--host-edges] [--reorder M]\n"
                    "       %s --read-log FILE\n", argv[0], argv[0]);
    return 2;
  }
  if (!strcmp(argv[1], "--read-log")) {
    if (argc < 3) {
      fprintf(stderr, "--read-log needs a file\n");
      return 2;
    }
    return read_log_mode(argv[2]);
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

  // resume from the log file, as read_log_file (mpi/benchmark_helper.hpp:530-575)
  const char* logname = getenv("LOGFILE");
  LogFile log;
  memset(&log, 0, sizeof log);
  int root_start = 0;
  if (logname && load_log(logname, log)) {
    if (log.scale != scale || log.edge_factor != edgefactor || log.mpi_size != 1) {
      printf("Log file is not match the current run: params:(current),(log): SCALE:%d,%d, edgefactor:%d,%d, "
             "size:%d,%d\n", scale, log.scale, edgefactor, log.edge_factor, 1, log.mpi_size);
      return 1;
    }
    root_start = log.num_runs;
    printf("===== LOG START =====\n");
    printf("graph_generation:               %f s\n", log.generation_time);
    printf("construction_time:              %f s\n", log.construction_time);
    for (int i = 0; i < root_start; ++i) {
      printf("Running BFS %d\n", i);
      printf("Time for BFS %d is %f\n", i, log.times[i].bfs_time);
      printf("Validating BFS %d\n", i);
      printf("Validate time for BFS %d is %f\n", i, log.times[i].validate_time);
      printf("TEPS for BFS %d is %g\n", i, log.times[i].edge_counts / log.times[i].bfs_time);
    }
    printf("=====  LOG END  =====\n");
    if (root_start != 0) printf("Resume from %d th run\n", root_start);
  }

  g5::BfsOnGPU bfs(0, reorder);
  printf("Graph generation\n");
  double generation_time = now_s();
  bfs.generate_kronecker(scale, edgefactor);
  g5_synchronize(bfs.context());
  generation_time = now_s() - generation_time;

  printf("Graph construction\n");
  double construction_time;
  const char* tmpfile = getenv("TMPFILE");
  if (host_edges && tmpfile) {
    FileEdgeList list(std::string(tmpfile) + "-000");
    list.write_from(bfs.context());
    construction_time = now_s();
    bfs.construct(&list);
  } else if (host_edges) {
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

  if (root_start > num_roots) root_start = num_roots;
  if (root_start == 0) {  // init_log (mpi/benchmark_helper.hpp:595-604)
    log.scale = scale;
    log.edge_factor = edgefactor;
    log.mpi_size = 1;
    log.num_runs = 0;
    log.generation_time = generation_time;
    log.construction_time = construction_time;
    log.redistribution_time = 0.0;
  }
  bfs.prepare_bfs();
  std::vector<double> bfs_times(num_roots), validate_times(num_roots), edge_counts(num_roots);
  for (int i = 0; i < root_start; ++i) {
    bfs_times[i] = log.times[i].bfs_time;
    validate_times[i] = log.times[i].validate_time;
    edge_counts[i] = (double)log.times[i].edge_counts;
  }
  bool result_ok = true;
  for (int i = root_start; i < num_roots; ++i) {
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
    if (logname) {  // update_log_file (mpi/benchmark_helper.hpp:577-593)
      const int run_num = log.num_runs++;
      log.times[run_num].bfs_time = bfs_times[i];
      log.times[run_num].validate_time = validate_times[i];
      log.times[run_num].edge_counts = edge_visit_count;
      save_log(logname, log);
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
  print_results(bfs_times, validate_times, edge_counts);
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
