// ref_probe.cc -- dump driver linked against the reference CPU path.
//
// TEST INFRASTRUCTURE ONLY (oracle side).  Built by oracle/ref.mk with the
// reference headers included where they lie (-I<reference>/mpi); nothing of
// the reference is copied here.  It runs the reference's own generator,
// constructor, root search, BFS and validator at one SCALE and writes what the
// golden fixtures need into OUTDIR:
//   tuples.bin      12-byte packed tuples in generation order
//   info.txt        estimated SCALE, n_active, directed entries after dedup
//   roots.txt       one root per line
//   levels_<i>.bin  uint8 BFS level per original vertex id (0xFF unreached)
//   nedge.txt       validator traversed-edge count per root
// Per-level directions come from the reference's verbose output on stderr.
//
// Usage: refprobe SCALE EDGEFACTOR OUTDIR [spec2012]
// spec2012: generate_graph_spec2012 (initiator 5500/100, binary-tree initial
// edges; the unweighted tuple type, so no weights) instead of spec2010.
#include <mpi.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <stdint.h>
#include <inttypes.h>
#include <math.h>
#include <string>

#include "parameters.h"
#include "utils_core.h"
#include "primitives.hpp"
#include "utils.hpp"
#include "../generator/graph_generator.hpp"
#include "graph_constructor.hpp"
#include "validate.hpp"
#include "benchmark_helper.hpp"
#include "bfs.hpp"
#include "bfs_cpu.hpp"

static FILE* open_out(const std::string& dir, const char* name, const char* mode) {
  std::string p = dir + "/" + name;
  FILE* f = fopen(p.c_str(), mode);
  if (!f) {
    fprintf(stderr, "refprobe: cannot open %s\n", p.c_str());
    exit(2);
  }
  return f;
}

int main(int argc, char** argv) {
  if (argc < 4) {
    fprintf(stderr, "usage: %s SCALE EDGEFACTOR OUTDIR [spec2012]\n", argv[0]);
    return 1;
  }
  const int scale = atoi(argv[1]);
  const int ef = atoi(argv[2]);
  const std::string out = argv[3];
  const bool spec2012 = argc > 4 && !strcmp(argv[4], "spec2012");
  setup_globals(argc, argv, scale, ef);

  typedef EdgeListStorage<UnweightedPackedEdge, 8 * 1024 * 1024> EdgeList;
  EdgeList edges((int64_t(1) << scale) * ef / mpi.size_2d, NULL);
  if (spec2012) generate_graph_spec2012(&edges, scale, ef, 255);
  else generate_graph_spec2010(&edges, scale, ef);

  {
    FILE* f = open_out(out, "tuples.bin", "wb");
    int loops = edges.beginRead(false);
    for (int l = 0; l < loops; ++l) {
      UnweightedPackedEdge* buf;
      int n = edges.read(&buf);
      fwrite(buf, sizeof(UnweightedPackedEdge), n, f);
    }
    edges.endRead();
    fclose(f);
  }

  BfsOnCPU* bfs = new BfsOnCPU();
  bfs->construct(&edges);
  redistribute_edge_2d(&edges);

  int64_t roots[PRM::NUM_BFS_ROOTS];
  int nroots = PRM::NUM_BFS_ROOTS;
  find_roots(bfs->graph_, roots, nroots);
  const int64_t max_used = find_max_used_vertex(bfs->graph_);
  const int64_t nlocal = bfs->graph_.pred_size();
  {
    FILE* f = open_out(out, "info.txt", "w");
    fprintf(f, "scale_estimate %d\nnum_active %" PRId64 "\nnnz %" PRId64 "\npred_size %" PRId64
               "\nmax_used_vertex %" PRId64 "\n",
            bfs->graph_.log_orig_global_verts_, bfs->graph_.num_global_verts_,
            bfs->graph_.num_global_edges_, nlocal, max_used);
    fclose(f);
    f = open_out(out, "roots.txt", "w");
    for (int i = 0; i < nroots; ++i) fprintf(f, "%" PRId64 "\n", roots[i]);
    fclose(f);
  }

  int64_t* pred = static_cast<int64_t*>(malloc(nlocal * sizeof(int64_t)));
  uint8_t* lv = static_cast<uint8_t*>(malloc(nlocal));
  bfs->prepare_bfs();
  FILE* fn = open_out(out, "nedge.txt", "w");
  for (int i = 0; i < nroots; ++i) {
    fprintf(stderr, "refprobe: root %d = %" PRId64 "\n", i, roots[i]);
    bfs->run_bfs(roots[i], pred);
    bfs->get_pred(pred);
    for (int64_t v = 0; v < nlocal; ++v) {
      uint16_t d = (uint16_t)((pred[v] >> 48) & 0xFFFF);
      lv[v] = (pred[v] == -1 || d == 0xFFFF) ? 0xFF : (uint8_t)d;
    }
    char name[64];
    snprintf(name, sizeof(name), "levels_%d.bin", i);
    FILE* f = open_out(out, name, "wb");
    fwrite(lv, 1, nlocal, f);
    fclose(f);
    int64_t nedge = 0;
    int ok = validate_bfs_result(&edges, max_used + 1, nlocal, roots[i], pred, &nedge);
    fprintf(fn, "%d %" PRId64 "\n", ok, nedge);
    fflush(fn);
  }
  fclose(fn);
  bfs->end_bfs();
  delete bfs;
  free(pred);
  free(lv);
  cleanup_globals();
  return 0;
}
