// Device side of the block-based triangle counter for gfx950 (CDNA4).
//
// A graph's whole block CSR is resident in HBM. A plan turns a task list into
// a flat table of work items (a task and a chunk of 64 non-empty rows of its
// ab tile); one persistent launch drains that table through an atomic
// head counter, so uneven tasks share the machine instead of being
// serialised launch by launch as in bb_tc.cpp:142-181.
//
// Kernels:
//   count_merge  - (d) wave-cooperative merge: the 64 lanes of a wave flatten
//                  the (u, v) wedges of their 64 rows with a wave prefix sum
//                  and each lane intersects ac(u) with bc(v); counts reduce
//                  across the wave into one 64-bit atomic per item.
//   count_bitmap - (c) LDS bitmap: a workgroup marks ac(u) of its rows in an
//                  LDS bitmap over the k band and probes every bc(v) entry
//                  with one bit test, counting with ballot + popcount.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../host/handles.hpp"

namespace bbtc {

#define HIP_OK(expr)                                                                   \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess)                                                              \
      ::bbtc::fail(BBTC_ERR_HIP, std::string(#expr) + " failed: " + hipGetErrorString(e_));    \
  } while (0)

// ---------------------------------------------------------------------------
// device-side descriptors
// ---------------------------------------------------------------------------
struct TileDesc {
  uint64_t rp;    // offset of the tile's row offsets in rowptr
  uint64_t col;   // offset of the tile's columns in cols
  uint64_t nz;    // offset of the tile's non-empty row list in nzrows
  uint32_t nzr;   // number of non-empty rows
  uint32_t h;     // band height
};
static_assert(sizeof(TileDesc) == 32, "TileDesc layout");

struct TaskDesc {
  uint32_t ab, bc, ac;  // tile ids (i*P+j, j*P+k, i*P+k)
  uint32_t kind;        // 0 merge, 1 bitmap
};

struct CountArgs {
  const uint32_t* __restrict__ rowptr;
  const uint32_t* __restrict__ cols;
  const uint32_t* __restrict__ nzrows;
  const TileDesc* __restrict__ tiles;
  const TaskDesc* __restrict__ tasks;
  const uint64_t* __restrict__ item_prefix;  // ntasks + 1
  uint32_t ntasks;
  uint64_t nitems;
  unsigned long long* per_task;
  unsigned long long* head;  // work counter
};

constexpr int kWave = 64;
constexpr int kRowsPerItem = 64;
constexpr int kBlock = 256;

// ---------------------------------------------------------------------------
// small device helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x, int lane) {
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    const uint32_t y = __shfl_up(x, off, kWave);
    if (lane >= off) x += y;
  }
  return x;
}

__device__ __forceinline__ unsigned long long wave_sum64(unsigned long long x) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) x += __shfl_xor(x, off, kWave);
  return x;
}

// |A ∩ B| for sorted, duplicate-free lists: merge with early exit once one
// cursor passes the other list's last element.
__device__ __forceinline__ uint32_t merge_count(const uint32_t* __restrict__ A, uint32_t na,
                                                const uint32_t* __restrict__ B, uint32_t nb) {
  if (na == 0 || nb == 0) return 0;
  const uint32_t alast = A[na - 1], blast = B[nb - 1];
  uint32_t i = 0, j = 0, c = 0;
  uint32_t x = A[0], y = B[0];
  for (;;) {
    if (x == y) {
      ++c;
      if (++i == na || ++j == nb) break;
      x = A[i];
      y = B[j];
    } else if (x < y) {
      if (y > alast || ++i == na) break;
      x = A[i];
    } else {
      if (x > blast || ++j == nb) break;
      y = B[j];
    }
  }
  return c;
}

// Task whose item range holds `item`: largest t with prefix[t] <= item.
__device__ __forceinline__ uint32_t find_task(const uint64_t* __restrict__ prefix, uint32_t n,
                                              uint64_t item) {
  uint32_t lo = 0, hi = n;  // prefix[lo] <= item < prefix[hi]
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (prefix[mid] <= item) lo = mid;
    else hi = mid;
  }
  return lo;
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
__global__ void zero_counts(unsigned long long* __restrict__ p, uint64_t n,
                            unsigned long long* __restrict__ head) {
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
  if (i == 0) *head = 0;
}

__global__ __launch_bounds__(kBlock) void count_merge(CountArgs a) {
  const int lane = threadIdx.x & (kWave - 1);
  uint32_t t_cur = 0;
  uint64_t t_lo = 1, t_hi = 0;  // cached item range of t_cur (empty at start)
  for (;;) {
    unsigned long long item = 0;
    if (lane == 0) item = atomicAdd(a.head, 1ull);
    item = __shfl(item, 0, kWave);
    if (item >= a.nitems) break;
    if (!(item >= t_lo && item < t_hi)) {
      t_cur = find_task(a.item_prefix, a.ntasks, item);
      t_lo = a.item_prefix[t_cur];
      t_hi = a.item_prefix[t_cur + 1];
    }
    const TaskDesc T = a.tasks[t_cur];
    const TileDesc AB = a.tiles[T.ab];
    const TileDesc BC = a.tiles[T.bc];
    const TileDesc AC = a.tiles[T.ac];
    const uint32_t* __restrict__ ab_rp = a.rowptr + AB.rp;
    const uint32_t* __restrict__ ac_rp = a.rowptr + AC.rp;
    const uint32_t* __restrict__ bc_rp = a.rowptr + BC.rp;
    const uint32_t* __restrict__ ab_c = a.cols + AB.col;
    const uint32_t* __restrict__ ac_c = a.cols + AC.col;
    const uint32_t* __restrict__ bc_c = a.cols + BC.col;

    const uint64_t z = (item - t_lo) * kRowsPerItem + lane;
    const bool valid = z < AB.nzr;
    uint32_t ab0 = 0, ac0 = 0, ac1 = 0, dab = 0;
    if (valid) {
      const uint32_t r = a.nzrows[AB.nz + z];
      ab0 = ab_rp[r];
      const uint32_t ab1 = ab_rp[r + 1];
      ac0 = ac_rp[r];
      ac1 = ac_rp[r + 1];
      dab = (ac1 > ac0) ? ab1 - ab0 : 0u;
    }
    const uint32_t incl = wave_incl_scan(dab, lane);
    const uint32_t excl = incl - dab;
    const uint32_t W = __shfl(incl, kWave - 1, kWave);

    unsigned long long cnt = 0;
    for (uint32_t base = 0; base < W; base += kWave) {
      const uint32_t w = base + lane;
      // owner lane: number of lanes whose inclusive prefix is <= w
      int q = 0;
#pragma unroll
      for (int s = kWave / 2; s >= 1; s >>= 1) {
        const uint32_t pv = __shfl(incl, q + s - 1, kWave);
        if (pv <= w) q += s;
      }
      q = q > kWave - 1 ? kWave - 1 : q;
      const uint32_t q_ab0 = __shfl(ab0, q, kWave);
      const uint32_t q_excl = __shfl(excl, q, kWave);
      const uint32_t q_ac0 = __shfl(ac0, q, kWave);
      const uint32_t q_ac1 = __shfl(ac1, q, kWave);
      if (w < W) {
        const uint32_t v = ab_c[q_ab0 + (w - q_excl)];
        const uint32_t b0 = bc_rp[v], b1 = bc_rp[v + 1];
        cnt += merge_count(ac_c + q_ac0, q_ac1 - q_ac0, bc_c + b0, b1 - b0);
      }
    }
    cnt = wave_sum64(cnt);
    if (lane == 0 && cnt) atomicAdd(a.per_task + t_cur, cnt);
  }
}

}  // namespace bbtc

// ---------------------------------------------------------------------------
// host objects
// ---------------------------------------------------------------------------
struct bbtc_ctx {
  int device = 0;
  int num_cu = 0;
  hipStream_t stream = nullptr;
  uint64_t budget = 0;
  uint64_t in_use = 0;
};

struct bbtc_dgraph {
  bbtc_ctx* ctx = nullptr;
  uint32_t P = 0;
  uint64_t bytes = 0;
  uint32_t* rowptr = nullptr;
  uint32_t* cols = nullptr;
  uint32_t* nzrows = nullptr;
  bbtc::TileDesc* tiles = nullptr;
  std::vector<bbtc::TileDesc> htiles;
  std::vector<float> deg;
};

struct bbtc_plan {
  bbtc_ctx* ctx = nullptr;
  bbtc_dgraph* g = nullptr;
  uint64_t ntasks = 0;
  uint64_t nitems = 0;
  uint64_t bitmap_tasks = 0;
  uint64_t bytes = 0;
  bbtc::TaskDesc* tasks = nullptr;
  uint64_t* item_prefix = nullptr;
  unsigned long long* per_task = nullptr;
  unsigned long long* head = nullptr;
  hipStream_t last_stream = nullptr;
};

namespace bbtc {
void set_last_error(const std::string& s);

namespace {

template <class T>
T* dev_alloc(bbtc_ctx* ctx, uint64_t count, uint64_t* acct) {
  const uint64_t bytes = std::max<uint64_t>(count, 1) * sizeof(T);
  if (ctx->budget && ctx->in_use + bytes > ctx->budget)
    fail(BBTC_ERR_BUDGET, "device budget of " + std::to_string(ctx->budget) + " bytes exceeded");
  void* p = nullptr;
  HIP_OK(hipSetDevice(ctx->device));
  HIP_OK(hipMalloc(&p, bytes));
  ctx->in_use += bytes;
  *acct += bytes;
  return static_cast<T*>(p);
}

void dev_free(bbtc_ctx* ctx, void* p) {
  if (p) (void)hipFree(p);
  (void)ctx;
}

void release_graph(bbtc_dgraph* g) {
  if (!g) return;
  dev_free(g->ctx, g->rowptr);
  dev_free(g->ctx, g->cols);
  dev_free(g->ctx, g->nzrows);
  dev_free(g->ctx, g->tiles);
  if (g->ctx) g->ctx->in_use -= std::min(g->ctx->in_use, g->bytes);
  delete g;
}

void release_plan(bbtc_plan* p) {
  if (!p) return;
  dev_free(p->ctx, p->tasks);
  dev_free(p->ctx, p->item_prefix);
  dev_free(p->ctx, p->per_task);
  dev_free(p->ctx, p->head);
  if (p->ctx) p->ctx->in_use -= std::min(p->ctx->in_use, p->bytes);
  delete p;
}

bool skipped(const bbtc_dgraph* g, uint32_t ab, uint32_t bc, uint32_t ac) {
  return g->deg[bc] == 0 || g->deg[ac] == 0 || g->deg[ab] == 0;
}

void launch(bbtc_plan* p, hipStream_t s) {
  bbtc_dgraph* g = p->g;
  const uint64_t zn = std::max<uint64_t>(p->ntasks, 1);
  hipLaunchKernelGGL(zero_counts, dim3(unsigned((zn + 255) / 256)), dim3(256), 0, s, p->per_task,
                     p->ntasks, p->head);
  HIP_OK(hipGetLastError());
  if (p->nitems == 0) return;
  CountArgs a;
  a.rowptr = g->rowptr;
  a.cols = g->cols;
  a.nzrows = g->nzrows;
  a.tiles = g->tiles;
  a.tasks = p->tasks;
  a.item_prefix = p->item_prefix;
  a.ntasks = uint32_t(p->ntasks);
  a.nitems = p->nitems;
  a.per_task = p->per_task;
  a.head = p->head;
  const int waves_per_block = kBlock / kWave;
  const uint64_t want_blocks = (p->nitems + waves_per_block - 1) / waves_per_block;
  const unsigned blocks = unsigned(std::min<uint64_t>(want_blocks, uint64_t(p->ctx->num_cu) * 8));
  hipLaunchKernelGGL(count_merge, dim3(blocks), dim3(kBlock), 0, s, a);
  HIP_OK(hipGetLastError());
  p->last_stream = s;
}

}  // namespace
}  // namespace bbtc

#define DEV_TRY try {
#define DEV_CATCH                                                                \
  }                                                                              \
  catch (const bbtc::Error& e) {                                                 \
    bbtc::set_last_error(e.what());                                              \
    return e.code;                                                               \
  }                                                                              \
  catch (const std::bad_alloc&) {                                                \
    bbtc::set_last_error("host allocation failed");                              \
    return BBTC_ERR_NOMEM;                                                       \
  }                                                                              \
  catch (const std::exception& e) {                                              \
    bbtc::set_last_error(e.what());                                              \
    return BBTC_ERR_INTERNAL;                                                    \
  }                                                                              \
  return BBTC_OK;

#define DEV_REQUIRE(cond, msg)                                                   \
  do {                                                                           \
    if (!(cond)) {                                                               \
      bbtc::set_last_error(msg);                                                 \
      return BBTC_ERR_INVALID;                                                   \
    }                                                                            \
  } while (0)

extern "C" {

bbtc_status bbtc_ctx_create(int device, const bbtc_opts* opts, bbtc_ctx** out) {
  DEV_REQUIRE(out, "null argument");
  DEV_TRY
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    bbtc::fail(BBTC_ERR_NODEVICE, "no HIP device visible");
  if (device < 0 || device >= ndev) bbtc::fail(BBTC_ERR_INVALID, "device index out of range");
  auto* c = new bbtc_ctx;
  c->device = device;
  c->budget = opts ? opts->hbm_budget_bytes : 0;
  try {
    HIP_OK(hipSetDevice(device));
    HIP_OK(hipDeviceGetAttribute(&c->num_cu, hipDeviceAttributeMultiprocessorCount, device));
    HIP_OK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  } catch (...) {
    delete c;
    throw;
  }
  *out = c;
  DEV_CATCH
}

void bbtc_ctx_destroy(bbtc_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) {
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipStreamDestroy(ctx->stream);
  }
  delete ctx;
}

uint64_t bbtc_ctx_bytes_in_use(const bbtc_ctx* ctx) { return ctx ? ctx->in_use : 0; }

bbtc_status bbtc_graph_upload(bbtc_ctx* ctx, const bbtc_layout* Lh, bbtc_dgraph** out) {
  DEV_REQUIRE(ctx && Lh && out, "null argument");
  DEV_TRY
  const bbtc::Layout& L = Lh->L;
  auto* g = new bbtc_dgraph;
  g->ctx = ctx;
  g->P = L.P;
  try {
    const size_t T = size_t(L.P) * L.P;
    g->htiles.resize(T);
    for (size_t t = 0; t < T; ++t) {
      bbtc::TileDesc d{};
      const uint32_t i = uint32_t(t / L.P), k = uint32_t(t % L.P);
      d.h = L.height(i);
      if (i <= k) {
        d.rp = L.rp_base[t];
        d.col = L.col_base[t];
        d.nz = L.nz_base[t];
        d.nzr = L.nz_rows[t];
      }
      g->htiles[t] = d;
    }
    g->deg = L.deg;
    g->rowptr = bbtc::dev_alloc<uint32_t>(ctx, L.rowptr.size(), &g->bytes);
    g->cols = bbtc::dev_alloc<uint32_t>(ctx, L.cols.size(), &g->bytes);
    g->nzrows = bbtc::dev_alloc<uint32_t>(ctx, L.nzrows.size(), &g->bytes);
    g->tiles = bbtc::dev_alloc<bbtc::TileDesc>(ctx, T, &g->bytes);
    HIP_OK(hipMemcpyAsync(g->rowptr, L.rowptr.data(), L.rowptr.size() * 4, hipMemcpyHostToDevice, ctx->stream));
    HIP_OK(hipMemcpyAsync(g->cols, L.cols.data(), L.cols.size() * 4, hipMemcpyHostToDevice, ctx->stream));
    HIP_OK(hipMemcpyAsync(g->nzrows, L.nzrows.data(), L.nzrows.size() * 4, hipMemcpyHostToDevice, ctx->stream));
    HIP_OK(hipMemcpyAsync(g->tiles, g->htiles.data(), T * sizeof(bbtc::TileDesc), hipMemcpyHostToDevice, ctx->stream));
    HIP_OK(hipStreamSynchronize(ctx->stream));
  } catch (...) {
    bbtc::release_graph(g);
    throw;
  }
  *out = g;
  DEV_CATCH
}

uint64_t bbtc_graph_bytes(const bbtc_dgraph* g) { return g ? g->bytes : 0; }

void bbtc_graph_release(bbtc_dgraph* g) {
  if (g && g->ctx) (void)hipSetDevice(g->ctx->device);
  bbtc::release_graph(g);
}

bbtc_status bbtc_plan_create(bbtc_ctx* ctx, bbtc_dgraph* g, const bbtc_task* tasks, uint64_t n,
                             bbtc_kernel kernel, bbtc_plan** out) {
  DEV_REQUIRE(ctx && g && out && (tasks || n == 0), "null argument");
  DEV_REQUIRE(g->ctx == ctx, "graph belongs to another context");
  DEV_REQUIRE(n < 0xFFFFFFFFull, "too many tasks");
  DEV_REQUIRE(kernel == BBTC_KERNEL_AUTO || kernel == BBTC_KERNEL_MERGE || kernel == BBTC_KERNEL_BITMAP,
              "unknown kernel choice");
  for (uint64_t q = 0; q < n; ++q)
    DEV_REQUIRE(tasks[q].i <= tasks[q].j && tasks[q].j <= tasks[q].k && tasks[q].k < g->P,
                "task out of range or not i <= j <= k");
  DEV_TRY
  auto* p = new bbtc_plan;
  p->ctx = ctx;
  p->g = g;
  p->ntasks = n;
  try {
    const uint32_t P = g->P;
    std::vector<bbtc::TaskDesc> td(std::max<uint64_t>(n, 1));
    std::vector<uint64_t> pref(n + 1, 0);
    for (uint64_t q = 0; q < n; ++q) {
      const bbtc_task& x = tasks[q];
      bbtc::TaskDesc d;
      d.ab = x.i * P + x.j;
      d.bc = x.j * P + x.k;
      d.ac = x.i * P + x.k;
      d.kind = 0;
      td[q] = d;
      uint64_t items = 0;
      if (!bbtc::skipped(g, d.ab, d.bc, d.ac))
        items = (uint64_t(g->htiles[d.ab].nzr) + bbtc::kRowsPerItem - 1) / bbtc::kRowsPerItem;
      pref[q + 1] = pref[q] + items;
    }
    p->nitems = pref[n];
    HIP_OK(hipSetDevice(ctx->device));
    p->tasks = bbtc::dev_alloc<bbtc::TaskDesc>(ctx, td.size(), &p->bytes);
    p->item_prefix = bbtc::dev_alloc<uint64_t>(ctx, pref.size(), &p->bytes);
    p->per_task = bbtc::dev_alloc<unsigned long long>(ctx, std::max<uint64_t>(n, 1), &p->bytes);
    p->head = bbtc::dev_alloc<unsigned long long>(ctx, 1, &p->bytes);
    HIP_OK(hipMemcpyAsync(p->tasks, td.data(), td.size() * sizeof(bbtc::TaskDesc), hipMemcpyHostToDevice, ctx->stream));
    HIP_OK(hipMemcpyAsync(p->item_prefix, pref.data(), pref.size() * 8, hipMemcpyHostToDevice, ctx->stream));
    HIP_OK(hipStreamSynchronize(ctx->stream));
  } catch (...) {
    bbtc::release_plan(p);
    throw;
  }
  *out = p;
  DEV_CATCH
}

bbtc_status bbtc_plan_launch(bbtc_plan* plan, void* stream) {
  DEV_REQUIRE(plan, "null argument");
  DEV_TRY
  HIP_OK(hipSetDevice(plan->ctx->device));
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : plan->ctx->stream;
  bbtc::launch(plan, s);
  DEV_CATCH
}

bbtc_status bbtc_plan_results(bbtc_plan* plan, uint64_t* per_task, uint64_t* total) {
  DEV_REQUIRE(plan, "null argument");
  DEV_TRY
  HIP_OK(hipSetDevice(plan->ctx->device));
  hipStream_t s = plan->last_stream ? plan->last_stream : plan->ctx->stream;
  std::vector<uint64_t> h(std::max<uint64_t>(plan->ntasks, 1), 0);
  if (plan->ntasks)
    HIP_OK(hipMemcpyAsync(h.data(), plan->per_task, plan->ntasks * 8, hipMemcpyDeviceToHost, s));
  HIP_OK(hipStreamSynchronize(s));
  uint64_t t = 0;
  for (uint64_t q = 0; q < plan->ntasks; ++q) t += h[q];
  if (per_task && plan->ntasks) std::copy(h.begin(), h.begin() + plan->ntasks, per_task);
  if (total) *total = t;
  DEV_CATCH
}

bbtc_status bbtc_plan_time(bbtc_plan* plan, int iters, float* ms_each) {
  DEV_REQUIRE(plan && iters >= 0 && (ms_each || iters == 0), "bad argument");
  DEV_TRY
  HIP_OK(hipSetDevice(plan->ctx->device));
  hipStream_t s = plan->ctx->stream;
  std::vector<hipEvent_t> ev(2 * size_t(iters));
  for (auto& e : ev) HIP_OK(hipEventCreate(&e));
  try {
    for (int i = 0; i < iters; ++i) {
      HIP_OK(hipEventRecord(ev[2 * i], s));
      bbtc::launch(plan, s);
      HIP_OK(hipEventRecord(ev[2 * i + 1], s));
    }
    HIP_OK(hipStreamSynchronize(s));
    for (int i = 0; i < iters; ++i) HIP_OK(hipEventElapsedTime(&ms_each[i], ev[2 * i], ev[2 * i + 1]));
  } catch (...) {
    for (auto& e : ev) (void)hipEventDestroy(e);
    throw;
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  DEV_CATCH
}

uint64_t bbtc_plan_bitmap_tasks(const bbtc_plan* plan) { return plan ? plan->bitmap_tasks : 0; }

void bbtc_plan_destroy(bbtc_plan* plan) {
  if (!plan) return;
  (void)hipSetDevice(plan->ctx->device);
  if (plan->last_stream) (void)hipStreamSynchronize(plan->last_stream);
  bbtc::release_plan(plan);
}

bbtc_status bbtc_count(bbtc_ctx* ctx, bbtc_dgraph* g, const bbtc_task* tasks, uint64_t n,
                       bbtc_kernel kernel, uint64_t* per_task, uint64_t* total, void* stream) {
  bbtc_plan* p = nullptr;
  bbtc_status st = bbtc_plan_create(ctx, g, tasks, n, kernel, &p);
  if (st != BBTC_OK) return st;
  st = bbtc_plan_launch(p, stream);
  if (st == BBTC_OK) st = bbtc_plan_results(p, per_task, total);
  bbtc_plan_destroy(p);
  return st;
}

}  // extern "C"
