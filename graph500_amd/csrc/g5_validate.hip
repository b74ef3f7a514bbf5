// g5_validate.hip -- Graph500 result validation on the GPU.
//
// Restates the rules of BfsValidation::validate (mpi/validate.hpp:506-802,
// helpers :809-995) for the last single-root run, on the device-resident
// levels and parents (new ids) and the CSR:
//   1. the root is its own parent and has level 0               (:530-536)
//   2. no other vertex is its own parent                         (:542-569)
//   3. level(v) == level(parent(v)) + 1 for every reached v      (:920-995)
//   4. every edge joins two reached or two unreached vertices and
//      their levels differ by at most one                        (:598-775)
//   5. every reached non-root vertex has an edge to its parent   (:780-792)
// Rule 4 runs over the CSR, which holds every non-loop input tuple in both
// directions (a self loop always passes it); rule 5 looks the parent up by
// binary search in the sorted row.  Parent values are in range by
// construction for the device result; a caller's pred array is range-checked
// when it is converted (rule :822-838).  The traversed-edge count is the sum
// of the per-vertex tuple weights over reached vertices, halved (:703-721).
#include "g5_ctx.hpp"
#include "g5_frontier.hpp"

#include <mutex>

namespace g5 {

enum { V_ROOT = 0, V_SELF, V_DEPTH, V_EDGE, V_PARENT_EDGE, V_RANGE, V_NUM = 8 };

// pred (parent | level << 48 per original id, -1 unreached; mpi/bfs.hpp:1203,
// 1271, 2470) -> level/parent in new ids.  Entries the checks below cannot see
// are counted here: parent out of range (:822-838), a "no parent" entry with a
// level (:929-936), a reached vertex without edges (it cannot have an edge to
// its parent, :780-792), a level beyond the uint8 range.
__global__ void __launch_bounds__(kBlock) k_val_from_pred(const int64_t* __restrict__ pred, int64_t n,
                                                          const uint32_t* __restrict__ old2new, int64_t rows,
                                                          int64_t root, uint8_t* __restrict__ level,
                                                          uint32_t* __restrict__ parent,
                                                          unsigned long long* __restrict__ err) {
  uint64_t e_range = 0, e_depth = 0, e_pedge = 0, e_root = 0;
  for (int64_t v = (int64_t)blockIdx.x * kBlock + threadIdx.x; v < n; v += (int64_t)gridDim.x * kBlock) {
    const int64_t x = pred[v];
    const uint32_t nv = old2new[v];
    const bool row = (int64_t)nv < rows;
    const int64_t p = (int64_t)((uint64_t)x << 16) >> 16;
    const uint32_t d = (uint32_t)(((uint64_t)x >> 48) & 0xFFFF);
    if (p == -1) {
      e_depth += (d != 0xFFFF);
      if (v == root) ++e_root;
      if (row) level[nv] = 0xFF;
      continue;
    }
    if (p < -1 || p >= n) {
      ++e_range;
      if (row) level[nv] = 0xFF;
      continue;
    }
    if (!row) {
      if (v == root) e_root += (p != root || d != 0);
      else ++e_pedge;
      continue;
    }
    if (d > 254) {
      ++e_depth;
      level[nv] = 0xFF;
      continue;
    }
    level[nv] = (uint8_t)d;
    parent[nv] = old2new[p];
  }
  block_add(e_range, err + V_RANGE);
  block_add(e_depth, err + V_DEPTH);
  block_add(e_pedge, err + V_PARENT_EDGE);
  block_add(e_root, err + V_ROOT);
}

__global__ void __launch_bounds__(kBlock) k_val_vertices(const uint8_t* __restrict__ level,
                                                         const uint32_t* __restrict__ parent,
                                                         const uint64_t* __restrict__ offsets,
                                                         const uint32_t* __restrict__ cols, int64_t rows,
                                                         uint32_t root, unsigned long long* __restrict__ err) {
  uint64_t e_root = 0, e_self = 0, e_depth = 0, e_pedge = 0;
  for (int64_t v = (int64_t)blockIdx.x * kBlock + threadIdx.x; v < rows; v += (int64_t)gridDim.x * kBlock) {
    const uint8_t l = level[v];
    if (l == 0xFF) continue;
    const uint32_t p = parent[v];
    if ((uint32_t)v == root) {
      e_root += (p != root || l != 0);
      continue;
    }
    if (p == (uint32_t)v) {
      ++e_self;
      continue;
    }
    if ((int64_t)p >= rows || level[p] == 0xFF || (int)level[p] + 1 != (int)l) {
      ++e_depth;
      continue;
    }
    // edge (v, p) must exist: binary search in v's sorted row
    uint64_t lo = offsets[v], hi = offsets[v + 1];
    bool found = false;
    while (lo < hi) {
      const uint64_t mid = (lo + hi) >> 1;
      const uint32_t c = cols[mid];
      if (c == p) {
        found = true;
        break;
      }
      if (c < p) lo = mid + 1; else hi = mid;
    }
    e_pedge += !found;
  }
  block_add(e_root, err + V_ROOT);
  block_add(e_self, err + V_SELF);
  block_add(e_depth, err + V_DEPTH);
  block_add(e_pedge, err + V_PARENT_EDGE);
}

// Rule 4 over all directed CSR entries, one wave per row.
__global__ void __launch_bounds__(kBlock) k_val_edges(const uint8_t* __restrict__ level,
                                                      const uint64_t* __restrict__ offsets,
                                                      const uint32_t* __restrict__ cols, int64_t rows,
                                                      unsigned long long* __restrict__ err) {
  const int lane = threadIdx.x & 63;
  const int64_t nwaves = (int64_t)gridDim.x * (kBlock / 64);
  uint64_t bad = 0;
  for (int64_t v = (int64_t)blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6); v < rows; v += nwaves) {
    const int lv = level[v];
    const uint64_t b = offsets[v], e = offsets[v + 1];
    for (uint64_t k = b + lane; k < e; k += 64) {
      const int lw = level[cols[k]];
      const bool in_v = lv != 0xFF, in_w = lw != 0xFF;
      bad += (in_v != in_w) || (in_v && (lv - lw > 1 || lw - lv > 1));
    }
  }
  block_add(bad, err + V_EDGE);
}

// err: V_NUM device counters (cleared here when clear is set).
static void run_checks(g5_ctx* ctx, const uint8_t* level, const uint32_t* parent, int64_t root_orig,
                       bool clear, unsigned long long* err, int64_t* errors_out) {
  Graph& g = ctx->graph;
  hipStream_t s = ctx->stream;
  const int64_t rows = g.n_rows;
  if (clear) G5_HIP(hipMemsetAsync(err, 0, V_NUM * 8, s));
  // a root that is no CSR row (no edges, modes 1 and 2) reaches nothing else
  uint32_t root_new = 0;
  G5_HIP(hipMemcpyAsync(&root_new, g.old2new.p + root_orig, 4, hipMemcpyDeviceToHost, s));
  G5_HIP(hipStreamSynchronize(s));
  if ((int64_t)root_new >= rows) root_new = 0xFFFFFFFFu;
  hipLaunchKernelGGL(k_val_vertices, dim3(grid_of(rows, kBlock, 8192)), dim3(kBlock), 0, s, level, parent,
                     g.offsets.p, g.cols.p, rows, root_new, err);
  G5_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_val_edges, dim3(grid_of(rows, kBlock / 64, 8192)), dim3(kBlock), 0, s, level,
                     g.offsets.p, g.cols.p, rows, err);
  G5_HIP(hipGetLastError());
  uint64_t h[V_NUM];
  G5_HIP(hipMemcpyAsync(h, err, V_NUM * 8, hipMemcpyDeviceToHost, s));
  G5_HIP(hipStreamSynchronize(s));
  for (int i = 0; i < 6; ++i) errors_out[i] = (int64_t)h[i];
}

void bfs_validate(g5_ctx* ctx, int64_t* errors_out) {
  BfsState& b = ctx->bfs;
  G5_CHECK(b.have_result, "g5_validate: no BFS result (call g5_run_bfs first)");
  ctx->acc.alloc(65);
  run_checks(ctx, b.level.p, b.parent.p, b.root_orig, true, (unsigned long long*)ctx->acc.p, errors_out);
}

// The checks of a caller's pred array (host memory, pred_size() entries) on
// device buffers the caller of this function provides: dpred (n words, the
// pred copy), lvl / par (n_rows each) and acc (V_NUM + 1 words: the error
// counters, then the reduction word).  Returns the traversed-edge count of
// the reached set.
static int64_t pred_validate_on(g5_ctx* ctx, int64_t root, const int64_t* host_pred, int64_t* dpred, uint8_t* lvl,
                                uint32_t* par, uint64_t* acc, int64_t* errors_out) {
  Graph& g = ctx->graph;
  hipStream_t s = ctx->stream;
  const int64_t rows = g.n_rows;
  G5_HIP(hipMemcpyAsync(dpred, host_pred, g.n * 8, hipMemcpyHostToDevice, s));
  G5_HIP(hipMemsetAsync(acc, 0, (V_NUM + 1) * 8, s));
  hipLaunchKernelGGL(k_val_from_pred, dim3(grid_of(g.n, kBlock, 8192)), dim3(kBlock), 0, s, dpred, g.n,
                     g.old2new.p, rows, root, lvl, par, (unsigned long long*)acc);
  G5_HIP(hipGetLastError());
  run_checks(ctx, lvl, par, root, false, (unsigned long long*)acc, errors_out);
  const uint64_t sum = device_reduce_sum(ReachedWeight{lvl, g.weight.p}, (size_t)rows, acc + V_NUM, s);
  return (int64_t)(sum / 2);
}

static void pred_validate_args(g5_ctx* ctx, int64_t root, const char* who) {
  Graph& g = ctx->graph;
  G5_CHECK(g.built, std::string(who) + ": no graph (call g5_construct first)");
  G5_CHECK(root >= 0 && root < g.n, std::string(who) + ": root out of range");
}

// Validates a caller's pred array (host memory, pred_size() entries) and
// returns the traversed-edge count of its reached set.
int64_t pred_validate(g5_ctx* ctx, int64_t root, const int64_t* host_pred, int64_t* errors_out) {
  pred_validate_args(ctx, root, "g5_validate_pred");
  Graph& g = ctx->graph;
  DevBuf<int64_t> dpred;
  DevBuf<uint8_t> lvl;
  DevBuf<uint32_t> par;
  dpred.alloc(g.n);
  lvl.alloc(g.n_rows);
  par.alloc(g.n_rows);
  ctx->acc.alloc(V_NUM + 1);
  return pred_validate_on(ctx, root, host_pred, dpred.p, lvl.p, par.p, ctx->acc.p, errors_out);
}

// ---------------------------------------------------------------------------
// The same checks without any allocation (g5_validate_pred_noalloc): every
// device buffer above is carved from `scratch`, one file-scope array per GPU
// that stays resident until the process ends.  Its host-side name is no
// device address: the address each device holds it at comes from
// hipGetSymbolAddress on the first call on that device (with the device
// current) and is kept per device.  Calls that overlap on one GPU share it
// and race; the counters are cleared by every call, and the level and parent
// words are written for every row before they are read.
// ---------------------------------------------------------------------------
constexpr size_t kValScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kValScratchBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static unsigned char* g_scratch_dev[kMaxDevices];  // device address of scratch, per device

static size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

// Bytes of scratch a graph of n ids and `rows` rows needs.
static size_t noalloc_need(int64_t n, int64_t rows) {
  return align256((V_NUM + 1) * 8) + align256((size_t)n * 8) + align256((size_t)rows) + align256((size_t)rows * 4);
}

static unsigned char* scratch_on(int device) {
  G5_CHECK(device >= 0 && device < kMaxDevices, "g5_validate_pred_noalloc: device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[device]) {
    void* p = nullptr;
    G5_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    G5_CHECK(p, "g5_validate_pred_noalloc: no device address for the scratch array");
    g_scratch_dev[device] = (unsigned char*)p;
  }
  return g_scratch_dev[device];
}

int64_t pred_validate_noalloc(g5_ctx* ctx, int64_t root, const int64_t* host_pred, int64_t* errors_out) {
  pred_validate_args(ctx, root, "g5_validate_pred_noalloc");
  Graph& g = ctx->graph;
  const size_t need = noalloc_need(g.n, g.n_rows);
  G5_CHECK_CODE(need <= kValScratchBytes, G5_ERR_LIMIT,
                "g5_validate_pred_noalloc: the graph needs " + std::to_string(need) + " bytes of scratch, more than " +
                    std::to_string(kValScratchBytes) + " (at most about 2^22 vertex ids); use g5_validate_pred");
  unsigned char* base = scratch_on(ctx->device);  // the caller made ctx->device current
  uint64_t* acc = (uint64_t*)base;
  size_t off = align256((V_NUM + 1) * 8);
  int64_t* dpred = (int64_t*)(base + off);
  off += align256((size_t)g.n * 8);
  uint8_t* lvl = base + off;
  off += align256((size_t)g.n_rows);
  uint32_t* par = (uint32_t*)(base + off);
  return pred_validate_on(ctx, root, host_pred, dpred, lvl, par, acc, errors_out);
}

}  // namespace g5
