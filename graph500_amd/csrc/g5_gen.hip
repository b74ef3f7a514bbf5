// g5_gen.hip -- Kronecker (R-MAT) edge generation on gfx950, bit-exact with the
// Graph500 splittable-MRG generator.
//
// Reference behaviour restated (no code shared):
//   MRG recurrence and skip tables  generator/splittable_mrg.c:17-56,115-200
//   seed expansion                  generator/graph_generator.hpp:327-333
//   scramble constants / scramble   generator/graph_generator.hpp:375-384,405-469
//   generateRange/make_one_edge     generator/graph_generator.hpp:612-704
//
// Device design: one lane per tuple.  The lane skips the seed state by
// edge_index * 2^64 (the "middle" exponent word) with one 5x5 transition per
// non-zero byte of edge_index (at most 4 for 2^32 tuples), then draws SCALE
// quadrant choices.  Arithmetic mod 2^31-1 uses 64-bit multiply-accumulate
// (v_mad_u64_u32) and a two-fold Mersenne reduction instead of division.
#include "g5_internal.hpp"

#include <mutex>

namespace g5 {

static constexpr uint32_t kM = 0x7FFFFFFFu;
static constexpr uint32_t kX = 107374182u;
static constexpr uint32_t kY = 104480u;

// ---------------------------------------------------------------------------
// Host: transition table A^(v * 256^j) built from full 5x5 matrix products.
// ---------------------------------------------------------------------------
namespace {

struct Mat5 {
  uint32_t m[5][5];
};

inline uint32_t hmul(uint32_t a, uint32_t b) { return (uint32_t)(((uint64_t)a * b) % kM); }

void mat_mul(const Mat5& a, const Mat5& b, Mat5& r) {
  Mat5 o;
  for (int i = 0; i < 5; ++i)
    for (int j = 0; j < 5; ++j) {
      uint64_t acc = 0;
      for (int k = 0; k < 5; ++k) acc = (acc + (uint64_t)a.m[i][k] * b.m[k][j]) % kM;
      o.m[i][j] = (uint32_t)acc;
    }
  r = o;
}

std::vector<MrgEntry>* g_table = nullptr;
std::once_flag g_table_once;

void build_table() {
  auto* t = new std::vector<MrgEntry>(24 * 256);
  Mat5 p{};  // A^(256^j)
  p.m[0][0] = kX;
  p.m[0][4] = kY;
  for (int i = 1; i < 5; ++i) p.m[i][i - 1] = 1;
  for (int j = 0; j < 24; ++j) {
    if (j > 0)
      for (int s = 0; s < 8; ++s) mat_mul(p, p, p);
    Mat5 cur{};
    for (int i = 0; i < 5; ++i) cur.m[i][i] = 1;
    for (int v = 0; v < 256; ++v) {
      if (v > 0) mat_mul(cur, p, cur);
      // A power is fixed by its bottom row (s t u v w) and its first column.
      MrgEntry e{cur.m[4][0], cur.m[4][1], cur.m[4][2], cur.m[4][3], cur.m[4][4],
                 cur.m[3][0], cur.m[2][0], cur.m[1][0], cur.m[0][0]};
      (*t)[j * 256 + v] = e;
    }
  }
  g_table = t;
}

}  // namespace

const MrgEntry* host_mrg_table() {
  std::call_once(g_table_once, build_table);
  return g_table->data();
}

void host_make_mrg_seed(uint64_t s1, uint64_t s2, MrgState* st) {
  st->z1 = (uint32_t)((s1 & 0x3FFFFFFF) + 1);
  st->z2 = (uint32_t)(((s1 >> 30) & 0x3FFFFFFF) + 1);
  st->z3 = (uint32_t)((s2 & 0x3FFFFFFF) + 1);
  st->z4 = (uint32_t)(((s2 >> 30) & 0x3FFFFFFF) + 1);
  st->z5 = (uint32_t)(((s2 >> 60) << 4) + (s1 >> 60) + 1);
}

// Applies the structured power of A:
//   [d  sy ay by cy]
//   [c  w  sy ay by]
//   [b  v  w  sy ay]
//   [a  u  v  w  sy]
//   [s  t  u  v  w ]
static void host_apply(const MrgEntry& e, MrgState* st) {
  const uint64_t z1 = st->z1, z2 = st->z2, z3 = st->z3, z4 = st->z4, z5 = st->z5;
  const uint64_t sy = hmul(e.s, kY), ay = hmul(e.a, kY), by = hmul(e.b, kY), cy = hmul(e.c, kY);
  st->z1 = (uint32_t)((e.d * z1 % kM + sy * z2 % kM + ay * z3 % kM + by * z4 % kM + cy * z5 % kM) % kM);
  st->z2 = (uint32_t)((e.c * z1 % kM + e.w * z2 % kM + sy * z3 % kM + ay * z4 % kM + by * z5 % kM) % kM);
  st->z3 = (uint32_t)((e.b * z1 % kM + e.v * z2 % kM + e.w * z3 % kM + sy * z4 % kM + ay * z5 % kM) % kM);
  st->z4 = (uint32_t)((e.a * z1 % kM + e.u * z2 % kM + e.v * z3 % kM + e.w * z4 % kM + sy * z5 % kM) % kM);
  st->z5 = (uint32_t)((e.s * z1 % kM + e.t * z2 % kM + e.u * z3 % kM + e.v * z4 % kM + e.w * z5 % kM) % kM);
}

void host_mrg_skip(MrgState* st, uint64_t hi, uint64_t mid, uint64_t lo) {
  const MrgEntry* t = host_mrg_table();
  const uint64_t words[3] = {lo, mid, hi};
  for (int p = 0; p < 3; ++p) {
    uint64_t e = words[p];
    for (int b = 0; e; ++b, e >>= 8) {
      const int v = (int)(e & 0xFF);
      if (v) host_apply(t[(p * 8 + b) * 256 + v], st);
    }
  }
}

uint32_t host_mrg_get_uint(MrgState* st) {
  const uint32_t nz = (uint32_t)(((uint64_t)kX * st->z1 + (uint64_t)kY * st->z5) % kM);
  st->z5 = st->z4;
  st->z4 = st->z3;
  st->z3 = st->z2;
  st->z2 = st->z1;
  st->z1 = nz;
  return nz;
}

// Two draws, scaled by (2^31-1)^-1 and (2^31-1)^-2, first draw first.  This
// translation unit is compiled with -ffp-contract=off so the sum is rounded
// after each product exactly as written.
double host_mrg_get_double(MrgState* st) {
  const uint32_t u1 = host_mrg_get_uint(st);
  const uint32_t u2 = host_mrg_get_uint(st);
  volatile double a = (double)u1 * .000000000465661287524579692;
  volatile double b = (double)u2 * .0000000000000000002168404346990492787;
  return a + b;
}

void host_scramble_values(uint64_t s1, uint64_t s2, uint64_t* v0, uint64_t* v1) {
  MrgState st;
  host_make_mrg_seed(s1, s2, &st);
  host_mrg_skip(&st, 50, 7, 0);
  uint64_t a = host_mrg_get_uint(&st);
  a *= UINT64_C(0xFFFFFFFF);
  a += host_mrg_get_uint(&st);
  uint64_t b = host_mrg_get_uint(&st);
  b *= UINT64_C(0xFFFFFFFF);
  b += host_mrg_get_uint(&st);
  *v0 = a;
  *v1 = b;
}

// ---------------------------------------------------------------------------
// Device
// ---------------------------------------------------------------------------
// Table entry as the device reads it: the 5 distinct column values of each
// row pattern, with the y-multiples precomputed.
struct DevEntry {
  uint32_t d, c, b, a, s, t, u, v, w, sy, ay, by, cy;
  uint32_t pad[3];
};

__device__ __forceinline__ uint32_t mred(uint64_t x) {
  x = (x & kM) + (x >> 31);
  x = (x & kM) + (x >> 31);
  uint32_t r = (uint32_t)x;
  return r >= kM ? r - kM : r;
}

// Five-term dot product mod 2^31-1; each product < 2^62, four of them < 2^64.
__device__ __forceinline__ uint32_t dot5(uint32_t a1, uint32_t b1, uint32_t a2, uint32_t b2,
                                         uint32_t a3, uint32_t b3, uint32_t a4, uint32_t b4,
                                         uint32_t a5, uint32_t b5) {
  uint64_t acc = (uint64_t)a1 * b1 + (uint64_t)a2 * b2 + (uint64_t)a3 * b3 + (uint64_t)a4 * b4;
  acc = (uint64_t)mred(acc) + (uint64_t)a5 * b5;
  return mred(acc);
}

struct DState {
  uint32_t z1, z2, z3, z4, z5;
};

__device__ __forceinline__ void dapply(const DevEntry& e, DState& s) {
  const uint32_t o1 = dot5(e.d, s.z1, e.sy, s.z2, e.ay, s.z3, e.by, s.z4, e.cy, s.z5);
  const uint32_t o2 = dot5(e.c, s.z1, e.w, s.z2, e.sy, s.z3, e.ay, s.z4, e.by, s.z5);
  const uint32_t o3 = dot5(e.b, s.z1, e.v, s.z2, e.w, s.z3, e.sy, s.z4, e.ay, s.z5);
  const uint32_t o4 = dot5(e.a, s.z1, e.u, s.z2, e.v, s.z3, e.w, s.z4, e.sy, s.z5);
  const uint32_t o5 = dot5(e.s, s.z1, e.t, s.z2, e.u, s.z3, e.v, s.z4, e.w, s.z5);
  s.z1 = o1;
  s.z2 = o2;
  s.z3 = o3;
  s.z4 = o4;
  s.z5 = o5;
}

__device__ __forceinline__ uint32_t dstep(DState& s) {
  const uint32_t nz = mred((uint64_t)kX * s.z1 + (uint64_t)kY * s.z5);
  s.z5 = s.z4;
  s.z4 = s.z3;
  s.z3 = s.z2;
  s.z2 = s.z1;
  s.z1 = nz;
  return nz;
}

__device__ __forceinline__ uint64_t dscramble(uint64_t v, int scale, uint64_t val0, uint64_t val1) {
  v += val0 + val1;
  v *= (val0 | UINT64_C(0x4519840211493211));
  v = __builtin_bitreverse64(v) >> (64 - scale);
  v *= (val1 | UINT64_C(0x3050852102C843A5));
  v = __builtin_bitreverse64(v) >> (64 - scale);
  return v;
}

struct GenParams {
  DState seed;
  uint64_t val0, val1;
  int64_t start, count;
  int scale;
  uint32_t a_num, bc_num;
};

// tab: the 8 byte-tables of the middle exponent word (entries [8*256, 16*256)).
__global__ void __launch_bounds__(256) k_kronecker(g5_packed_edge* __restrict__ out,
                                                   const DevEntry* __restrict__ tab,
                                                   GenParams p) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const uint32_t limit = UINT32_C(0xFFFFFFFF) % 10000u;  // 7295: rejection bound
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < p.count; i += stride) {
    const uint64_t idx = (uint64_t)(p.start + i);
    DState st = p.seed;
    uint64_t e = idx;
    for (int b = 0; e; ++b, e >>= 8) {
      const uint32_t v = (uint32_t)(e & 0xFF);
      if (v) dapply(tab[b * 256 + v], st);
    }
    uint64_t bs = 0, bt = 0;
    for (int lvl = p.scale - 1; lvl >= 0; --lvl) {
      uint32_t val = dstep(st);
      while (val < limit) val = dstep(st);
      val %= 10000u;
      int sq;
      if (val < p.bc_num) sq = 1;
      else if (val - p.bc_num < p.bc_num) sq = 2;
      else if (val - 2 * p.bc_num < p.a_num) sq = 0;
      else sq = 3;
      uint32_t so = (uint32_t)sq >> 1, to = (uint32_t)sq & 1;
      if (bs == bt && so > to) {  // clip-and-flip
        so = 0;
        to = 1;
      }
      bs += (uint64_t)so << lvl;
      bt += (uint64_t)to << lvl;
    }
    const uint64_t v0 = dscramble(bs, p.scale, p.val0, p.val1);
    const uint64_t v1 = dscramble(bt, p.scale, p.val0, p.val1);
    g5_packed_edge r;
    r.v0_low = (uint32_t)v0;
    r.v1_low = (uint32_t)v1;
    r.high = (uint32_t)(((v0 >> 32) & 0xFFFF) | ((v1 >> 16) & 0xFFFF0000u));
    out[i] = r;
  }
}

// Binary-tree initial tuples (GraphGeneratorBase::generateInitialEdge,
// generator/graph_generator.hpp:515-524): tuple i joins vertex i+1 to its
// heap parent (i+1)/2, both scrambled.
__global__ void __launch_bounds__(256) k_initial_tree(g5_packed_edge* __restrict__ out, int64_t count, int scale,
                                                      uint64_t val0, uint64_t val1) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) {
    const uint64_t v0 = dscramble((uint64_t)(i + 1), scale, val0, val1);
    const uint64_t v1 = dscramble((uint64_t)((i + 1) / 2), scale, val0, val1);
    g5_packed_edge r;
    r.v0_low = (uint32_t)v0;
    r.v1_low = (uint32_t)v1;
    r.high = (uint32_t)(((v0 >> 32) & 0xFFFF) | ((v1 >> 16) & 0xFFFF0000u));
    out[i] = r;
  }
}

void launch_initial_tree(g5_packed_edge* out, int64_t count, int scale, uint64_t s1, uint64_t s2,
                         hipStream_t stream) {
  if (count <= 0) return;
  uint64_t val0, val1;
  host_scramble_values(s1, s2, &val0, &val1);
  int64_t blocks = (count + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(k_initial_tree, dim3((unsigned)blocks), dim3(256), 0, stream, out, count, scale, val0, val1);
  G5_HIP(hipGetLastError());
}

// Edge weights (GraphGeneratorBase::generateWeight,
// generator/graph_generator.hpp:480-506), written with the tuples as
// WeightedEdge: max_weight <= 0 gives the constant 0xBEEF of the reference's
// compiled branch; otherwise each lane skips 'base' (the seed already skipped
// by (30, 46, 0)) by its edge index in the low exponent word and draws once.
// low_tab: the 8 byte-tables of the low exponent word (entries [0, 8*256)).
__global__ void __launch_bounds__(256) k_weighted(const g5_packed_edge* __restrict__ in, int64_t start_index,
                                                  int64_t count, int max_weight, DState base,
                                                  const DevEntry* __restrict__ low_tab,
                                                  g5_weighted_edge* __restrict__ out) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) {
    const g5_packed_edge t = in[i];
    int32_t wgt = 0xBEEF;
    if (max_weight > 0) {
      DState st = base;
      uint64_t e = (uint64_t)(start_index + i);
      for (int b = 0; e; ++b, e >>= 8) {
        const uint32_t v = (uint32_t)(e & 0xFF);
        if (v) dapply(low_tab[b * 256 + v], st);
      }
      wgt = (int32_t)(dstep(st) % (uint32_t)max_weight) + 1;
    }
    g5_weighted_edge r;
    r.v0_low = t.v0_low;
    r.v1_low = t.v1_low;
    r.high = t.high;
    r.weight = wgt;
    out[i] = r;
  }
}

// Owns the device copies of the middle-word byte tables (tuples) and the
// low-word byte tables (weights).
struct GenTables {
  DevBuf<DevEntry> tab;
  DevBuf<DevEntry> low;
  static void upload(DevBuf<DevEntry>& dst, int first_row) {
    const MrgEntry* t = host_mrg_table();
    std::vector<DevEntry> h(8 * 256);
    for (int j = 0; j < 8; ++j)
      for (int v = 0; v < 256; ++v) {
        const MrgEntry& e = t[(first_row + j) * 256 + v];
        DevEntry d{};
        d.d = e.d; d.c = e.c; d.b = e.b; d.a = e.a;
        d.s = e.s; d.t = e.t; d.u = e.u; d.v = e.v; d.w = e.w;
        d.sy = hmul(e.s, kY); d.ay = hmul(e.a, kY); d.by = hmul(e.b, kY); d.cy = hmul(e.c, kY);
        h[j * 256 + v] = d;
      }
    dst.alloc(h.size());
    G5_HIP(hipMemcpy(dst.p, h.data(), h.size() * sizeof(DevEntry), hipMemcpyHostToDevice));
  }
  void ensure() {
    if (!tab.p) upload(tab, 8);
    if (!low.p) upload(low, 0);
  }
};

static GenTables& tables_for_device(int dev) {
  static std::mutex mu;
  static GenTables* per_dev[64] = {};
  std::lock_guard<std::mutex> lk(mu);
  if (!per_dev[dev]) per_dev[dev] = new GenTables();
  per_dev[dev]->ensure();
  return *per_dev[dev];
}

void launch_kronecker(int device, g5_packed_edge* out, int64_t start, int64_t count, int scale,
                      uint64_t s1, uint64_t s2, int a_num, int bc_num, hipStream_t stream) {
  G5_CHECK(scale >= 1 && scale <= 48, "g5_generate_kronecker: scale must be in [1, 48]");
  G5_CHECK(a_num >= 0 && bc_num >= 0 && a_num + 2 * bc_num <= 10000,
           "g5_generate_kronecker: initiator numerators must satisfy a + 2*bc <= 10000");
  G5_CHECK(start >= 0 && (uint64_t)(start + count) <= (UINT64_C(1) << 62),
           "g5_generate_kronecker: tuple index range too large");
  GenTables& t = tables_for_device(device);
  GenParams p;
  MrgState ms;
  host_make_mrg_seed(s1, s2, &ms);
  p.seed = DState{ms.z1, ms.z2, ms.z3, ms.z4, ms.z5};
  host_scramble_values(s1, s2, &p.val0, &p.val1);
  p.start = start;
  p.count = count;
  p.scale = scale;
  p.a_num = (uint32_t)a_num;
  p.bc_num = (uint32_t)bc_num;
  if (count <= 0) return;
  // Edge indices up to 2^64 use at most the 8 middle-word tables.
  int64_t blocks = (count + 255) / 256;
  if (blocks > 256 * 256) blocks = 256 * 256;
  hipLaunchKernelGGL(k_kronecker, dim3((unsigned)blocks), dim3(256), 0, stream, out, t.tab.p, p);
  G5_HIP(hipGetLastError());
}

void launch_weighted(int device, const g5_packed_edge* in, int64_t start_index, int64_t count, int max_weight,
                     uint64_t s1, uint64_t s2, g5_weighted_edge* out, hipStream_t stream) {
  if (count <= 0) return;
  GenTables& t = tables_for_device(device);
  MrgState ms;
  host_make_mrg_seed(s1, s2, &ms);
  host_mrg_skip(&ms, 30, 46, 0);
  const DState base{ms.z1, ms.z2, ms.z3, ms.z4, ms.z5};
  int64_t blocks = (count + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(k_weighted, dim3((unsigned)blocks), dim3(256), 0, stream, in, start_index, count, max_weight,
                     base, t.low.p, out);
  G5_HIP(hipGetLastError());
}

}  // namespace g5
