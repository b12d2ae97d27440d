// Device side of the block-based triangle counter for gfx950 (CDNA4): the
// context, device-resident graphs, plans and their launches. The kernels are
// in kernels.hpp.
//
// A graph's whole block CSR is resident in HBM. A plan turns a task list into
// a flat table of work items and chooses, per task, the LDS form of ac(u)
// (sorted list for binary search, or bitmap); one persistent launch drains
// every item through an atomic head counter, so uneven tasks share the
// machine instead of being serialised launch by launch as in
// bb_tc.cpp:142-181.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../host/device_api.hpp"
#include "kernels.hpp"
#include "objects.hpp"

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
  dev_free(g->ctx, g->dense);
  dev_free(g->ctx, g->tiles);
  if (g->ctx) g->ctx->in_use -= std::min(g->ctx->in_use, g->bytes);
  delete g;
}

void release_plan(bbtc_plan* p) {
  if (!p) return;
  dev_free(p->ctx, p->tasks);
  dev_free(p->ctx, p->item_prefix);
  dev_free(p->ctx, p->per_task);
  dev_free(p->ctx, p->heads);
  dev_free(p->ctx, p->cycles);
  if (p->ctx) p->ctx->in_use -= std::min(p->ctx->in_use, p->bytes);
  delete p;
}

bool skipped(const bbtc_dgraph* g, uint32_t ab, uint32_t bc, uint32_t ac) {
  return g->deg[bc] == 0 || g->deg[ac] == 0 || g->deg[ab] == 0;
}

void zero(bbtc_plan* p, hipStream_t s) {
  const uint64_t zn = std::max<uint64_t>(p->ntasks, 1);
  hipLaunchKernelGGL(zero_counts, dim3(unsigned((zn + 255) / 256)), dim3(256), 0, s, p->per_task,
                     p->ntasks);
  HIP_OK(hipGetLastError());
}

// Counts tasks [t0, t1) of the plan on stream s with work counter `slot`.
// Launches on one slot must be ordered (same stream); distinct slots may run
// concurrently.
void launch_range(bbtc_plan* p, uint64_t t0, uint64_t t1, int slot, hipStream_t s) {
  bbtc_dgraph* g = p->g;
  const uint64_t i0 = p->hprefix[t0], i1 = p->hprefix[t1];
  p->last_stream = s;
  if (i1 <= i0) return;
  unsigned long long* head = p->heads + slot;
  hipLaunchKernelGGL(set_head, dim3(1), dim3(1), 0, s, head, (unsigned long long)i0);
  HIP_OK(hipGetLastError());
  CountArgs a;
  a.rowptr = g->rowptr;
  a.cols = g->cols;
  a.nzrows = g->nzrows;
  a.dense = g->dense;
  a.tiles = g->tiles;
  a.tasks = p->tasks;
  a.item_prefix = p->item_prefix;
  a.ntasks = uint32_t(p->ntasks);
  a.item_end = i1;
  a.per_task = p->per_task;
  a.head = head;
  a.cycles = p->profiling ? p->cycles : nullptr;
  const uint64_t want_blocks = (i1 - i0 + kWavesPerBlock - 1) / kWavesPerBlock;
  const uint64_t resident = uint64_t(p->ctx->num_cu) * uint64_t(p->ctx->blocks_per_cu);
  const unsigned blocks = unsigned(std::max<uint64_t>(1, std::min<uint64_t>(want_blocks, resident)));
  hipLaunchKernelGGL(count_probe, dim3(blocks), dim3(kBlock), 0, s, a);
  HIP_OK(hipGetLastError());
}

void launch(bbtc_plan* p, hipStream_t s) {
  zero(p, s);
  launch_range(p, 0, p->ntasks, 0, s);
}

}  // namespace
}  // namespace bbtc

namespace bbtc {
namespace dev {

namespace {
hipStream_t slot_stream(bbtc_ctx* c, int slot) {
  if (slot <= 0) return c->stream;
  if (slot >= kMaxSlots) fail(BBTC_ERR_INVALID, "stream slot out of range");
  if (c->slots.size() < size_t(slot) + 1) c->slots.resize(size_t(slot) + 1, nullptr);
  if (!c->slots[slot]) HIP_OK(hipStreamCreateWithFlags(&c->slots[slot], hipStreamNonBlocking));
  return c->slots[slot];
}
}  // namespace

int max_slots() { return kMaxSlots; }

void plan_zero_sync(bbtc_plan* p) {
  HIP_OK(hipSetDevice(p->ctx->device));
  zero(p, p->ctx->stream);
  HIP_OK(hipStreamSynchronize(p->ctx->stream));
}

void plan_launch_range(bbtc_plan* p, uint64_t t0, uint64_t t1, int slot) {
  if (t0 > t1 || t1 > p->ntasks) fail(BBTC_ERR_INVALID, "task range out of bounds");
  HIP_OK(hipSetDevice(p->ctx->device));
  launch_range(p, t0, t1, slot, slot_stream(p->ctx, slot));
}

uint64_t plan_items(const bbtc_plan* p, uint64_t t0, uint64_t t1) { return p->hprefix[t1] - p->hprefix[t0]; }

uint64_t plan_batch_items(const bbtc_plan* p) {
  return uint64_t(p->ctx->num_cu) * uint64_t(p->ctx->blocks_per_cu) * kWavesPerBlock * 4;
}

bool plan_slot_idle(bbtc_plan* p, int slot) {
  const hipError_t e = hipStreamQuery(slot_stream(p->ctx, slot));
  if (e == hipSuccess) return true;
  if (e == hipErrorNotReady) return false;
  HIP_OK(e);
  return false;
}

void plan_slot_wait(bbtc_plan* p, int slot) { HIP_OK(hipStreamSynchronize(slot_stream(p->ctx, slot))); }

void plan_copy_counts(bbtc_plan* p, uint64_t* out) {
  HIP_OK(hipSetDevice(p->ctx->device));
  HIP_OK(hipDeviceSynchronize());
  if (p->ntasks) HIP_OK(hipMemcpy(out, p->per_task, p->ntasks * 8, hipMemcpyDeviceToHost));
}

}  // namespace dev
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

bbtc_status bbtc_device_count(int* n) {
  DEV_REQUIRE(n, "null argument");
  *n = 0;
  int c = 0;
  if (hipGetDeviceCount(&c) == hipSuccess) *n = c;
  return BBTC_OK;
}

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
    HIP_OK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&c->blocks_per_cu, bbtc::count_probe, bbtc::kBlock, 0));
    if (c->blocks_per_cu < 1) c->blocks_per_cu = 1;
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
  for (hipStream_t s : ctx->slots) {
    if (!s) continue;
    (void)hipStreamSynchronize(s);
    (void)hipStreamDestroy(s);
  }
  delete ctx;
}

uint64_t bbtc_ctx_bytes_in_use(const bbtc_ctx* ctx) { return ctx ? ctx->in_use : 0; }

bbtc_status bbtc_ctx_synchronize(bbtc_ctx* ctx) {
  DEV_REQUIRE(ctx, "null argument");
  DEV_TRY
  HIP_OK(hipSetDevice(ctx->device));
  HIP_OK(hipDeviceSynchronize());
  DEV_CATCH
}

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
      d.dense = ~0ull;
      d.dwpr = L.dwpr[k];
      if (i <= k) {
        d.rp = L.rp_base[t];
        d.col = L.col_base[t];
        d.nz = L.nz_base[t];
        d.nzr = L.nz_rows[t];
        d.dense = L.dense_base[t];
      }
      g->htiles[t] = d;
    }
    g->deg = L.deg;
    g->dwpr = L.dwpr;
    g->rowptr = bbtc::dev_alloc<uint32_t>(ctx, L.rowptr.size(), &g->bytes);
    g->cols = bbtc::dev_alloc<uint32_t>(ctx, L.cols.size(), &g->bytes);
    g->nzrows = bbtc::dev_alloc<uint32_t>(ctx, L.nzrows.size(), &g->bytes);
    g->dense = bbtc::dev_alloc<uint32_t>(ctx, L.dense.size(), &g->bytes);
    g->tiles = bbtc::dev_alloc<bbtc::TileDesc>(ctx, T, &g->bytes);
    HIP_OK(hipMemcpyAsync(g->rowptr, L.rowptr.data(), L.rowptr.size() * 4, hipMemcpyHostToDevice, ctx->stream));
    HIP_OK(hipMemcpyAsync(g->cols, L.cols.data(), L.cols.size() * 4, hipMemcpyHostToDevice, ctx->stream));
    HIP_OK(hipMemcpyAsync(g->nzrows, L.nzrows.data(), L.nzrows.size() * 4, hipMemcpyHostToDevice, ctx->stream));
    HIP_OK(hipMemcpyAsync(g->dense, L.dense.data(), L.dense.size() * 4, hipMemcpyHostToDevice, ctx->stream));
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
  return bbtc_plan_create_tuned(ctx, g, tasks, n, kernel, nullptr, out);
}

bbtc_status bbtc_plan_create_tuned(bbtc_ctx* ctx, bbtc_dgraph* g, const bbtc_task* tasks, uint64_t n,
                                   bbtc_kernel kernel, const uint8_t* choice, bbtc_plan** out) {
  DEV_REQUIRE(ctx && g && out && (tasks || n == 0), "null argument");
  DEV_REQUIRE(g->ctx == ctx, "graph belongs to another context");
  DEV_REQUIRE(n < 0xFFFFFFFFull, "too many tasks");
  DEV_REQUIRE(kernel == BBTC_KERNEL_AUTO || kernel == BBTC_KERNEL_MERGE || kernel == BBTC_KERNEL_BITMAP ||
                  kernel == BBTC_KERNEL_DENSE || kernel == BBTC_KERNEL_RDENSE,
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
      bbtc::TaskDesc d{};
      d.ab = x.i * P + x.j;
      d.bc = x.j * P + x.k;
      d.ac = x.i * P + x.k;
      // LDS form of ac(u): a bitmap over band k takes dwpr words per row; the
      // wave slice holds kLdsWords words.
      const uint32_t wpr = g->dwpr[x.k];
      const uint32_t fit = wpr ? uint32_t(bbtc::kLdsWords) / wpr : uint32_t(bbtc::kWave);
      const bool has_dense = g->htiles[d.bc].dense != ~0ull;
      bbtc_kernel k = kernel;
      if (choice && choice[q] == 1) k = BBTC_KERNEL_MERGE;
      else if (choice && choice[q] == 2) k = BBTC_KERNEL_BITMAP;
      else if (choice && choice[q] == 3) k = BBTC_KERNEL_DENSE;
      else if (choice && choice[q] == 4) k = BBTC_KERNEL_RDENSE;
      uint32_t kind = bbtc::kKindSearch;
      if (k == BBTC_KERNEL_BITMAP) {
        if (fit >= 1) kind = bbtc::kKindBitmap;
      } else if (k == BBTC_KERNEL_DENSE) {
        if (has_dense && fit >= 1) kind = bbtc::kKindDense;
        else if (fit >= 1) kind = bbtc::kKindBitmap;
      } else if (k == BBTC_KERNEL_RDENSE) {
        if (has_dense) kind = bbtc::kKindRdense;
      } else if (k == BBTC_KERNEL_AUTO) {
        // without measurements: the ac(u) list against the bc(v) bitmap row
        // where the bc tile has one, else the ac(u) bitmap where it fits
        if (has_dense) kind = bbtc::kKindRdense;
        else if (fit >= uint32_t(bbtc::kAutoBitmapRows)) kind = bbtc::kKindBitmap;
      }
      const bool use_bitmap = kind == bbtc::kKindBitmap || kind == bbtc::kKindDense;
      d.kind = kind;
      p->hkind.push_back(kind + 1);
      d.rows = use_bitmap ? std::min<uint32_t>(fit, bbtc::kWave) : uint32_t(bbtc::kWave);
      d.wpr = kind == bbtc::kKindSearch ? 0 : wpr;
      td[q] = d;
      uint64_t items = 0;
      if (!bbtc::skipped(g, d.ab, d.bc, d.ac)) {
        items = (uint64_t(g->htiles[d.ab].nzr) + d.rows - 1) / d.rows;
        if (kind != bbtc::kKindSearch) p->bitmap_tasks++;
      }
      pref[q + 1] = pref[q] + items;
    }
    p->nitems = pref[n];
    p->hprefix = pref;
    HIP_OK(hipSetDevice(ctx->device));
    p->tasks = bbtc::dev_alloc<bbtc::TaskDesc>(ctx, td.size(), &p->bytes);
    p->item_prefix = bbtc::dev_alloc<uint64_t>(ctx, pref.size(), &p->bytes);
    p->per_task = bbtc::dev_alloc<unsigned long long>(ctx, std::max<uint64_t>(n, 1), &p->bytes);
    p->heads = bbtc::dev_alloc<unsigned long long>(ctx, bbtc::kMaxSlots, &p->bytes);
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

bbtc_status bbtc_plan_time_tasks(bbtc_plan* plan, float* ms_per_task) {
  DEV_REQUIRE(plan && (ms_per_task || plan->ntasks == 0), "null argument");
  DEV_TRY
  HIP_OK(hipSetDevice(plan->ctx->device));
  hipStream_t s = plan->ctx->stream;
  const uint64_t n = plan->ntasks;
  std::vector<hipEvent_t> ev(n + 1);
  for (auto& e : ev) HIP_OK(hipEventCreate(&e));
  try {
    bbtc::zero(plan, s);
    HIP_OK(hipEventRecord(ev[0], s));
    for (uint64_t t = 0; t < n; ++t) {
      bbtc::launch_range(plan, t, t + 1, 0, s);
      HIP_OK(hipEventRecord(ev[t + 1], s));
    }
    HIP_OK(hipStreamSynchronize(s));
    for (uint64_t t = 0; t < n; ++t) HIP_OK(hipEventElapsedTime(&ms_per_task[t], ev[t], ev[t + 1]));
  } catch (...) {
    for (auto& e : ev) (void)hipEventDestroy(e);
    throw;
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  DEV_CATCH
}

bbtc_status bbtc_plan_profile(bbtc_plan* plan, uint64_t* cycles_per_task) {
  DEV_REQUIRE(plan && (cycles_per_task || plan->ntasks == 0), "null argument");
  DEV_TRY
  HIP_OK(hipSetDevice(plan->ctx->device));
  hipStream_t s = plan->ctx->stream;
  const uint64_t n = std::max<uint64_t>(plan->ntasks, 1);
  if (!plan->cycles) plan->cycles = bbtc::dev_alloc<unsigned long long>(plan->ctx, n, &plan->bytes);
  hipLaunchKernelGGL(bbtc::zero_counts, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, plan->cycles, n);
  HIP_OK(hipGetLastError());
  plan->profiling = true;
  try {
    bbtc::launch(plan, s);
  } catch (...) {
    plan->profiling = false;
    throw;
  }
  plan->profiling = false;
  if (plan->ntasks)
    HIP_OK(hipMemcpyAsync(cycles_per_task, plan->cycles, plan->ntasks * 8, hipMemcpyDeviceToHost, s));
  HIP_OK(hipStreamSynchronize(s));
  DEV_CATCH
}

bbtc_status bbtc_plan_task_kinds(const bbtc_plan* plan, uint8_t* kinds) {
  DEV_REQUIRE(plan && (kinds || plan->ntasks == 0), "null argument");
  for (uint64_t t = 0; t < plan->ntasks; ++t) kinds[t] = uint8_t(plan->hkind[t]);
  return BBTC_OK;
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
