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
#include <cmath>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../host/device_api.hpp"
#include "internal.hpp"
#include "kernels.hpp"

namespace bbtc {
void set_last_error(const std::string& s);

// ---------------------------------------------------------------------------
// Scratch of bbtc_count_noalloc: a fixed array in each device's global
// memory, so that entry point allocates nothing. Its layout (256-byte aligned
// parts): the slot control block, the segment descriptor, the task
// descriptors, the item prefix, the per-task counters and the spill records
// of split items. Host code never uses the symbol itself (its host-side name
// is a stand-in, not a device address): bbtc_ctx_create resolves the
// device's own address once with hipGetSymbolAddress and the kernels receive
// that address.
// ---------------------------------------------------------------------------
constexpr uint64_t align256(uint64_t x) { return (x + 255) / 256 * 256; }
constexpr uint64_t kScratchCtl = 0;
constexpr uint64_t kScratchSeg = align256(kScratchCtl + sizeof(SlotCtl));
constexpr uint64_t kScratchTasks = align256(kScratchSeg + sizeof(SegDesc));
constexpr uint64_t kScratchPrefix = align256(kScratchTasks + kNoAllocMaxTasks * sizeof(TaskDesc));
constexpr uint64_t kScratchCounts = align256(kScratchPrefix + (kNoAllocMaxTasks + 1) * 8);
constexpr uint64_t kScratchSpill = align256(kScratchCounts + kNoAllocMaxTasks * 8);
constexpr uint64_t kScratchSpillBase = align256(kScratchSpill + uint64_t(kNoAllocSpillCap) * sizeof(SpillRec));
constexpr uint64_t kScratchEnd = align256(kScratchSpillBase + uint64_t(kNoAllocSpillCap) * 8);
static_assert(kScratchEnd <= kScratchBytes, "scratch layout exceeds the array");

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
std::mutex scratch_mu;
std::vector<void*> scratch_addr;  // per device; nullptr until resolved
}  // namespace

// The device address of `scratch` on the current device (resolved once per
// device, with that device current).
void* scratch_of_device(int device) {
  std::lock_guard<std::mutex> lk(scratch_mu);
  if (scratch_addr.size() <= size_t(device)) scratch_addr.resize(size_t(device) + 1, nullptr);
  if (!scratch_addr[device]) {
    void* p = nullptr;
    HIP_OK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) fail(BBTC_ERR_HIP, "hipGetSymbolAddress returned no address for the scratch array");
    scratch_addr[device] = p;
  }
  return scratch_addr[device];
}

// ---------------------------------------------------------------------------
// Pinned staging of host-to-device copies.
// ---------------------------------------------------------------------------
constexpr uint64_t kStageBytes = kStageImageBytes;

struct Staging {
  std::mutex mu;
  void* buf[2] = {nullptr, nullptr};
  hipEvent_t ev[2] = {nullptr, nullptr};
  bool pending[2] = {false, false};  // a copy from the buffer may still be in flight
  int next = 0;                      // the buffer the next chunk fills

  // The next buffer, once its previous chunk has left (st.mu held).
  int take() {
    const int b = next;
    next ^= 1;
    if (pending[b]) {
      HIP_OK(hipEventSynchronize(ev[b]));
      pending[b] = false;
    }
    return b;
  }
  // Copies buf[b][0, len) to dst on s and marks the buffer busy until then.
  void send(int b, void* dst, uint64_t len, hipStream_t s) {
    try {
      HIP_OK(hipMemcpyAsync(dst, buf[b], len, hipMemcpyHostToDevice, s));
      HIP_OK(hipEventRecord(ev[b], s));
    } catch (...) {
      // a copy from either buffer may still be in flight: wait for the stream
      // before the buffers can be refilled
      (void)hipStreamSynchronize(s);
      pending[0] = pending[1] = false;
      throw;
    }
    pending[b] = true;
  }
};

namespace {
Staging& staging_of(bbtc_ctx* ctx) {
  {
    std::lock_guard<std::mutex> lk(scratch_mu);
    if (!ctx->h2d) ctx->h2d = new Staging;
  }
  Staging& st = *ctx->h2d;
  std::lock_guard<std::mutex> lk(st.mu);
  for (int b = 0; b < 2; ++b) {
    if (!st.buf[b]) HIP_OK(hipHostMalloc(&st.buf[b], kStageBytes, hipHostMallocDefault));
    if (!st.ev[b]) HIP_OK(hipEventCreateWithFlags(&st.ev[b], hipEventDisableTiming));
  }
  return st;
}
}  // namespace

void copy_to_device(bbtc_ctx* ctx, hipStream_t s, void* dst, const void* src, uint64_t bytes, bool src_pinned) {
  if (bytes < kStageMinBytes || src_pinned) {
    if (bytes) HIP_OK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
    return;
  }
  Staging& st = staging_of(ctx);
  std::lock_guard<std::mutex> lk(st.mu);
  const uint64_t nchunks = (bytes + kStageBytes - 1) / kStageBytes;
  for (uint64_t c = 0; c < nchunks; ++c) {
    const int b = st.take();
    const uint64_t o = c * kStageBytes, len = std::min(kStageBytes, bytes - o);
    parallel_copy(st.buf[b], static_cast<const char*>(src) + o, len);
    st.send(b, static_cast<char*>(dst) + o, len, s);
  }
}

void copy_image_to_device(bbtc_ctx* ctx, hipStream_t s, void* dst, const CopyPiece* pieces, size_t npieces,
                          uint64_t copy_bytes) {
  uint64_t end = 0;
  for (size_t q = 0; q < npieces; ++q) end = std::max(end, pieces[q].offset + pieces[q].bytes);
  if (end > copy_bytes || copy_bytes > kStageBytes)
    fail(BBTC_ERR_INTERNAL, "copy image of " + std::to_string(copy_bytes) + " bytes does not hold its pieces or "
                            "exceeds a staging buffer");
  if (!copy_bytes) return;
  Staging& st = staging_of(ctx);
  std::lock_guard<std::mutex> lk(st.mu);
  const int b = st.take();
  char* img = static_cast<char*>(st.buf[b]);
  for (size_t q = 0; q < npieces; ++q)
    if (pieces[q].bytes) std::memcpy(img + pieces[q].offset, pieces[q].src, pieces[q].bytes);
  st.send(b, dst, copy_bytes, s);
}

namespace {
void pinned_free(void* p) { (void)hipHostFree(p); }

template <class T>
void pin_buf(Buf<T>& b) {
  if (b.pinned() || b.size() == 0) return;
  const size_t bytes = b.size() * sizeof(T);
  void* p = nullptr;
  HIP_OK(hipHostMalloc(&p, bytes, hipHostMallocDefault));
  parallel_copy(p, b.data(), bytes);
  b.adopt(static_cast<T*>(p), b.size(), &pinned_free);
}
}  // namespace

void pin_layout(Layout& L) {
  pin_buf(L.rowptr);
  pin_buf(L.cols);
  pin_buf(L.nzrows);
}

void staging_release(bbtc_ctx* ctx) {
  if (!ctx->h2d) return;
  Staging* st = ctx->h2d;
  for (int b = 0; b < 2; ++b) {
    if (st->pending[b]) (void)hipEventSynchronize(st->ev[b]);
    if (st->ev[b]) (void)hipEventDestroy(st->ev[b]);
    if (st->buf[b]) (void)hipHostFree(st->buf[b]);
  }
  delete st;
  ctx->h2d = nullptr;
}

bool valid_kernel(bbtc_kernel k) {
  return k == BBTC_KERNEL_AUTO || k == BBTC_KERNEL_MERGE || k == BBTC_KERNEL_BITMAP ||
         k == BBTC_KERNEL_DENSE || k == BBTC_KERNEL_RDENSE || k == BBTC_KERNEL_LANE || k == BBTC_KERNEL_ROWS;
}

void enqueue_rows(bbtc_ctx* ctx, const RowArgs& a, hipStream_t s) {
  const uint64_t items = (uint64_t(a.nrows) + kWave - 1) / kWave;
  hipLaunchKernelGGL(reset_slot, dim3(1), dim3(1), 0, s, a.ctl, 0ull, (unsigned long long)items, 0ull);
  HIP_OK(hipGetLastError());
  const uint64_t want = (items + kWavesPerBlock - 1) / kWavesPerBlock;
  const unsigned blocks = unsigned(std::max<uint64_t>(1, std::min<uint64_t>(want, resident_blocks(ctx, true))));
  hipLaunchKernelGGL(count_rows, dim3(blocks), dim3(kBlock), 0, s, a);
  HIP_OK(hipGetLastError());
}

void enqueue_build_rows(const uint32_t* rowptr, const uint32_t* cols, const TileDesc* tiles, const uint32_t* cuts,
                        uint32_t P, uint32_t nrows, uint32_t S, uint32_t* rows, uint32_t* rowlen, hipStream_t s) {
  if (nrows == 0) return;
  const unsigned blocks = unsigned((uint64_t(nrows) + kBlock - 1) / kBlock);
  hipLaunchKernelGGL(build_rows, dim3(blocks), dim3(kBlock), 0, s, rowptr, cols, tiles, cuts, P,
                     uint32_t(__builtin_ctz(S)), reinterpret_cast<u32x4*>(rows), rowlen);
  HIP_OK(hipGetLastError());
}

bool rows_pay(uint32_t P, const std::vector<TileDesc>& tiles, const std::vector<float>& deg) {
  if (P == 0) return false;
  double n = 0, m = 0;
  for (uint32_t i = 0; i < P; ++i) {
    const double h = tiles[size_t(i) * P + i].h;
    n += h;
    for (uint32_t k = i; k < P; ++k) m += double(deg[size_t(i) * P + k]) * h;  // nnz of tile (i, k)
  }
  return n >= double(1u << 20) || P <= 2 || 2.0 * m <= 24.0 * n;
}

uint32_t auto_kind(const TaskWork& s, uint32_t wpr, uint32_t fit, bool has_dense, double* cost) {
  // the form of least estimated cost among those that apply
  uint32_t kind = kKindSearch;
  *cost = 0;
  if (!(s.W > 0)) return kind;
  const FormCosts c = form_costs(s, wpr);
  double best = c.merge;
  if (fit >= 1 && c.bitmap < best) {
    best = c.bitmap;
    kind = kKindBitmap;
  }
  if (has_dense && fit >= 1 && c.dense < best) {
    best = c.dense;
    kind = kKindDense;
  }
  if (has_dense && c.rdense < best) {
    best = c.rdense;
    kind = kKindRdense;
  }
  *cost = best;
  return kind;
}

std::vector<TaskRef> refs_of(uint32_t P, const bbtc_task* tasks, uint64_t n) {
  std::vector<TaskRef> r(n);
  for (uint64_t q = 0; q < n; ++q) {
    const bbtc_task& x = tasks[q];
    r[q] = TaskRef{x.i * P + x.j, x.j * P + x.k, x.i * P + x.k, 0};
  }
  return r;
}

std::vector<TileDesc> tile_table(const Layout& L) {
  const size_t T = size_t(L.P) * L.P;
  std::vector<TileDesc> out(T);
  for (size_t t = 0; t < T; ++t) {
    TileDesc d{};
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
      d.maxd = L.max_row[t];
    }
    out[t] = d;
  }
  return out;
}

void enqueue_zero(unsigned long long* p, uint64_t n, hipStream_t s) {
  const uint64_t zn = std::max<uint64_t>(n, 1);
  hipLaunchKernelGGL(zero_counts, dim3(unsigned((zn + 255) / 256)), dim3(256), 0, s, p, n);
  HIP_OK(hipGetLastError());
}

std::vector<DenseJob> dense_jobs(const Layout& L, uint64_t rp0, uint64_t col0, uint64_t dense0,
                                 uint64_t row0, uint64_t* rows_out) {
  std::vector<DenseJob> jobs;
  uint64_t rows = row0;
  for (uint32_t i = 0; i < L.P; ++i)
    for (uint32_t k = i; k < L.P; ++k) {
      const size_t t = L.tile(i, k);
      if (L.dense_base[t] == ~0ull) continue;
      const uint32_t h = L.height(i);
      if (L.dense_base[t] + uint64_t(h) * L.dwpr[k] > L.dense_words)
        fail(BBTC_ERR_INTERNAL, "bitmap rows outside the layout's bitmap area");
      jobs.push_back(DenseJob{rp0 + L.rp_base[t], col0 + L.col_base[t], dense0 + L.dense_base[t], rows, h, L.dwpr[k]});
      rows += h;
    }
  if (rows_out) *rows_out = rows - row0;
  return jobs;
}

void enqueue_dense_build(const uint32_t* rowptr, const uint32_t* cols, uint32_t* dense, uint64_t dense_words,
                         const DenseJob* jobs_dev, uint64_t njobs, uint64_t rows, hipStream_t s) {
  const uint64_t zb = std::min<uint64_t>((dense_words + kBlock - 1) / kBlock, 65536);
  hipLaunchKernelGGL(fill_zero_u32, dim3(unsigned(std::max<uint64_t>(zb, 1))), dim3(kBlock), 0, s, dense, dense_words);
  HIP_OK(hipGetLastError());
  if (!njobs || !rows) return;
  if (njobs >= 0xFFFFFFFFull) fail(BBTC_ERR_INVALID, "too many bitmap tiles");
  const uint64_t waves = (rows + kWave - 1) / kWave;
  const uint64_t blocks = std::min<uint64_t>((waves + kWavesPerBlock - 1) / kWavesPerBlock, 16384);
  hipLaunchKernelGGL(build_bitmaps, dim3(unsigned(std::max<uint64_t>(blocks, 1))), dim3(kBlock), 0, s, rowptr, cols,
                     dense, jobs_dev, uint32_t(njobs), rows);
  HIP_OK(hipGetLastError());
}

void enqueue_count(bbtc_ctx* ctx, CountArgs a, unsigned long long i0, uint64_t items, bool lane_only,
                   uint64_t chunk, hipStream_t s, uint32_t defer_steps) {
  hipLaunchKernelGGL(reset_slot, dim3(1), dim3(1), 0, s, a.ctl, i0, i0 + items, (unsigned long long)chunk);
  HIP_OK(hipGetLastError());
  a.phase = 0;
  a.defer_steps = defer_steps ? defer_steps
                  : chunk == 0 && items >= kManyItemsPerWave * resident_blocks(ctx) * kWavesPerBlock
                      ? kDeferStepsMany
                      : kDeferSteps;
  const uint64_t want = (items + kWavesPerBlock - 1) / kWavesPerBlock;
  const unsigned blocks = unsigned(std::max<uint64_t>(1, std::min<uint64_t>(want, resident_blocks(ctx, lane_only))));
  if (lane_only) {
    hipLaunchKernelGGL(count_probe_lanes, dim3(blocks), dim3(kBlock), 0, s, a);
    HIP_OK(hipGetLastError());
    return;
  }
  hipLaunchKernelGGL(count_probe, dim3(blocks), dim3(kBlock), 0, s, a);
  HIP_OK(hipGetLastError());
  // spilled chunks of heavy items, over the whole machine
  a.phase = 1;
  hipLaunchKernelGGL(count_probe, dim3(unsigned(resident_blocks(ctx))), dim3(kBlock), 0, s, a);
  HIP_OK(hipGetLastError());
}

std::vector<TaskWork> task_work(const Layout& L, const bbtc_task* tasks, uint64_t n) {
  std::vector<uint64_t> st(n * kTaskStats);
  task_stats(L, tasks, n, st.data(), kRuleSampleUnits);
  return work_from_stats(st.data(), n);
}

std::vector<TaskWork> work_from_stats(const uint64_t* st, uint64_t n) {
  std::vector<TaskWork> w(n);
  for (uint64_t q = 0; q < n; ++q) {
    const uint64_t* x = st + kTaskStats * q;
    w[q] = TaskWork{float(x[0]), float(x[1]), float(x[2]), float(x[3]), float(x[4]), float(x[5])};
  }
  return w;
}

PlanDesc describe_plan(const std::vector<TileDesc>& tiles, const std::vector<float>& deg,
                       const std::vector<TaskRef>& refs, bbtc_kernel kernel, const uint8_t* choice,
                       bool allow_auto_lanes, const std::vector<TaskWork>* work) {
  const uint64_t n = refs.size();
  // per task, the row form is the lane form (a graph-wide launch takes it)
  if (kernel == BBTC_KERNEL_ROWS) kernel = BBTC_KERNEL_AUTO;
  PlanDesc P;
  P.td.assign(std::max<uint64_t>(n, 1), TaskDesc{});
  P.pref.assign(n + 1, 0);
  P.hkind.reserve(n);
  auto short_rows = [&](const TaskRef& x) {
    return std::max(tiles[x.ab].maxd, std::max(tiles[x.bc].maxd, tiles[x.ac].maxd)) <= kLaneMaxRow;
  };
  auto has_work = [&](const TaskRef& x) {
    return !(deg[x.bc] == 0 || deg[x.ac] == 0 || deg[x.ab] == 0) && tiles[x.ab].nzr > 0;
  };
  // BBTC_KERNEL_AUTO takes the lane form for every task when every task with
  // work admits it (short rows throughout: the lane kernel then runs alone)
  bool auto_lanes = allow_auto_lanes && kernel == BBTC_KERNEL_AUTO && n > 0;
  for (uint64_t q = 0; q < n && auto_lanes; ++q)
    if (has_work(refs[q]) && !short_rows(refs[q])) auto_lanes = false;
  for (uint64_t q = 0; q < n; ++q) {
    const TaskRef& x = refs[q];
    TaskDesc d{};
    d.ab = x.ab;
    d.bc = x.bc;
    d.ac = x.ac;
    d.graph = x.graph;
    // LDS form of ac(u): a bitmap over band k takes dwpr words per row; the
    // wave slice holds kLdsWords words.
    const uint32_t wpr = tiles[x.ac].dwpr;
    const uint32_t fit = wpr ? uint32_t(kLdsWords) / wpr : uint32_t(kWave);
    const bool has_dense = tiles[x.bc].dense != ~0ull;
    bbtc_kernel k = kernel;
    const uint32_t form = choice ? choice[q] & kChoiceFormMask : 0u;
    const uint32_t rows_shift = choice ? (choice[q] >> kChoiceRowsShift) & 3u : 0u;
    if (form == 1) k = BBTC_KERNEL_MERGE;
    else if (form == 2) k = BBTC_KERNEL_BITMAP;
    else if (form == 3) k = BBTC_KERNEL_DENSE;
    else if (form == 4) k = BBTC_KERNEL_RDENSE;
    else if (form == 5) k = BBTC_KERNEL_LANE;
    else if (auto_lanes && !form) k = BBTC_KERNEL_LANE;
    uint32_t kind = kKindSearch;
    if (k == BBTC_KERNEL_BITMAP) {
      if (fit >= 1) kind = kKindBitmap;
    } else if (k == BBTC_KERNEL_DENSE) {
      if (has_dense && fit >= 1) kind = kKindDense;
      else if (fit >= 1) kind = kKindBitmap;
    } else if (k == BBTC_KERNEL_RDENSE) {
      if (has_dense) kind = kKindRdense;
    } else if (k == BBTC_KERNEL_LANE) {
      if (short_rows(x)) kind = kKindLane;
    } else if (k == BBTC_KERNEL_AUTO && work) {
      double cost = 0;
      kind = auto_kind((*work)[q], wpr, fit, has_dense, &cost);
    } else if (k == BBTC_KERNEL_AUTO) {
      // without statistics: the ac(u) list against the bc(v) bitmap row
      // where the bc tile has one, else the ac(u) bitmap where it fits
      if (has_dense) kind = kKindRdense;
      else if (fit >= uint32_t(kAutoBitmapRows)) kind = kKindBitmap;
    }
    const bool use_bitmap = kind == kKindBitmap || kind == kKindDense;
    d.kind = kind;
    P.hkind.push_back(kind + 1);
    d.rows = use_bitmap ? std::min<uint32_t>(fit, kWave) : uint32_t(kWave);
    d.rows = std::max<uint32_t>(1, d.rows >> rows_shift);  // the autotuner's rows per item
    d.wpr = kind == kKindSearch || kind == kKindLane ? 0 : wpr;
    P.td[q] = d;
    uint64_t items = 0;
    if (!(deg[d.bc] == 0 || deg[d.ac] == 0 || deg[d.ab] == 0)) {
      items = (uint64_t(tiles[d.ab].nzr) + d.rows - 1) / d.rows;
      if (kind != kKindSearch && kind != kKindLane) P.bitmap_tasks++;
      if (items && kind != kKindLane) P.lane_only = false;
    }
    P.pref[q + 1] = P.pref[q] + items;
  }
  return P;
}

namespace {

void release_graph(bbtc_dgraph* g) {
  if (!g) return;
  dev_free(g->ctx, g->rowptr);
  dev_free(g->ctx, g->cols);
  dev_free(g->ctx, g->nzrows);
  dev_free(g->ctx, g->dense);
  dev_free(g->ctx, g->tiles);
  dev_free(g->ctx, g->slots);
  dev_free(g->ctx, g->rows);
  dev_free(g->ctx, g->rowlen);
  if (g->ctx) g->ctx->in_use -= std::min(g->ctx->in_use, g->bytes);
  delete g;
}

void release_plan(bbtc_plan* p) {
  if (!p) return;
  dev_free(p->ctx, p->tasks);
  dev_free(p->ctx, p->item_prefix);
  dev_free(p->ctx, p->segs);
  dev_free(p->ctx, p->per_task);
  dev_free(p->ctx, p->ctl);
  for (auto& sp : p->spill) {
    dev_free(p->ctx, sp.rec);
    dev_free(p->ctx, sp.base);
  }
  dev_free(p->ctx, p->cycles);
  dev_free(p->ctx, p->lex2plan);
  dev_free(p->ctx, p->cuts);
  if (p->ctx) p->ctx->in_use -= std::min(p->ctx->in_use, p->bytes);
  delete p;
}

void zero(bbtc_plan* p, hipStream_t s) { enqueue_zero(p->per_task, p->ntasks, s); }

// Counts tasks [t0, t1) of the plan on stream s with work counter `slot`.
// Launches on one slot must be ordered (same stream); distinct slots may run
// concurrently.
void launch_range(bbtc_plan* p, uint64_t t0, uint64_t t1, int slot, hipStream_t s,
                  unsigned long long* queue = nullptr, uint32_t queue_lo = 0, uint32_t queue_ahead = 0) {
  const uint64_t i0 = p->hprefix[t0], i1 = p->hprefix[t1];
  p->last_stream = s;
  if (i1 <= i0) return;
  if (p->spill.size() <= size_t(slot)) p->spill.resize(size_t(slot) + 1);
  bbtc_plan::Spill& sp = p->spill[slot];
  if (!sp.rec) {
    sp.rec = dev_alloc<SpillRec>(p->ctx, p->spill_cap, &p->bytes);
    sp.base = dev_alloc<uint64_t>(p->ctx, p->spill_cap, &p->bytes);
  }
  CountArgs a{};
  a.segs = p->segs + (p->profiling ? 1 : 0);
  a.nseg = 1;
  a.item_end = i1;
  a.ctl = p->ctl + slot;
  a.spill = sp.rec;
  a.spill_base = sp.base;
  a.spill_cap = p->spill_cap;
  a.left = p->track_left;
  a.done = p->track_done;
  a.queue = queue;
  a.queue_lo = queue_lo;
  a.queue_ahead = queue_ahead;
  enqueue_count(p->ctx, a, i0, i1 - i0, p->lane_only, p->chunk, s);
}

// The row form's launch over every task of the plan (slot 0).
void launch_rows(bbtc_plan* p, hipStream_t s) {
  p->last_stream = s;
  const bbtc_dgraph* g = p->g;
  RowArgs a{};
  a.rows = g->rows;
  a.rowlen = g->rowlen;
  a.cuts = p->cuts;
  a.lex2plan = p->lex2plan;
  a.per_task = p->per_task;
  a.ctl = p->ctl;
  a.nrows = g->cuts[g->P];
  a.P = g->P;
  a.s_shift = uint32_t(__builtin_ctz(g->rows_S));
  enqueue_rows(p->ctx, a, s);
}

void launch(bbtc_plan* p, hipStream_t s) {
  zero(p, s);
  if (p->rows_form && !p->profiling) launch_rows(p, s);
  else launch_range(p, 0, p->ntasks, 0, s);
}

}  // namespace
}  // namespace bbtc

namespace bbtc {
namespace {

// Row slots of every non-empty upper tile of a single resident graph (the
// lane form reads a short bc(v) list from them in one 16-byte load), built
// from the resident tiles; none when they would take more than kSlotMaxBytes.
void build_graph_slots(bbtc_ctx* ctx, bbtc_dgraph* g) {
  const uint64_t bytes = g->rowptr_words * 16;
  if (bytes > kSlotMaxBytes || (ctx->budget && ctx->in_use + bytes > ctx->budget)) return;
  std::vector<DenseJob> jobs;
  uint64_t rows = 0;
  for (uint32_t i = 0; i < g->P; ++i)
    for (uint32_t k = i; k < g->P; ++k) {
      const TileDesc& d = g->htiles[size_t(i) * g->P + k];
      if (d.nzr == 0) continue;
      if (d.rp + d.h >= g->rowptr_words) fail(BBTC_ERR_INTERNAL, "tile rows outside the row offsets");
      jobs.push_back(DenseJob{d.rp, d.col, 0, rows, d.h, 0});
      rows += d.h;
    }
  if (jobs.empty()) return;
  if (jobs.size() >= 0xFFFFFFFFull) return;
  uint64_t acct = 0;
  uint32_t* slots = reinterpret_cast<uint32_t*>(dev_alloc<u32x4>(ctx, g->rowptr_words, &acct));
  DenseJob* jd = nullptr;
  try {
    jd = dev_alloc<DenseJob>(ctx, jobs.size(), &acct);
    HIP_OK(hipMemcpyAsync(jd, jobs.data(), jobs.size() * sizeof(DenseJob), hipMemcpyHostToDevice, ctx->stream));
    const uint64_t blocks = std::min<uint64_t>((rows + kBlock - 1) / kBlock, 65536);
    hipLaunchKernelGGL(build_slots, dim3(unsigned(std::max<uint64_t>(blocks, 1))), dim3(kBlock), 0, ctx->stream,
                       g->rowptr, g->cols, reinterpret_cast<u32x4*>(slots), jd, uint32_t(jobs.size()), rows);
    HIP_OK(hipGetLastError());
    HIP_OK(hipStreamSynchronize(ctx->stream));
  } catch (...) {
    (void)hipStreamSynchronize(ctx->stream);
    dev_free(ctx, jd);
    dev_free(ctx, slots);
    ctx->in_use -= std::min(ctx->in_use, acct);
    throw;
  }
  dev_free(ctx, jd);
  const uint64_t job_bytes = std::max<uint64_t>(jobs.size(), 1) * sizeof(DenseJob);
  ctx->in_use -= std::min(ctx->in_use, job_bytes);
  g->slots = slots;
  g->bytes += acct - job_bytes;
}

// The row table of a single resident graph (the row form), built from the
// resident tiles once: the longest row sets the slot size; none when a row
// holds more than kRowSlotMax entries, the table would exceed
// kRowTableMaxBytes or the budget, or the graph has more than kRowMaxCuts
// bands.
void build_graph_rows(bbtc_ctx* ctx, bbtc_dgraph* g) {
  g->rows_tried = true;
  const uint32_t P = g->P;
  if (P == 0 || P > kRowMaxCuts) return;
  g->cuts.assign(size_t(P) + 1, 0);
  for (uint32_t b = 0; b < P; ++b) {
    const uint64_t e = uint64_t(g->cuts[b]) + g->htiles[size_t(b) * P + b].h;
    if (e >= 0xFFFFFFFFull) return;  // ids must stay below kRowPad
    g->cuts[b + 1] = uint32_t(e);
  }
  const uint64_t nrows = g->cuts[P];
  if (nrows == 0) return;
  uint64_t acct = 0;
  uint32_t* cuts = nullptr;
  uint32_t* maxlen = nullptr;
  uint32_t* rows = nullptr;
  uint32_t* rowlen = nullptr;
  try {
    cuts = dev_alloc<uint32_t>(ctx, P + 1, &acct);
    maxlen = dev_alloc<uint32_t>(ctx, 1, &acct);
    HIP_OK(hipMemcpyAsync(cuts, g->cuts.data(), (P + 1) * 4, hipMemcpyHostToDevice, ctx->stream));
    const uint32_t zero = 0;
    HIP_OK(hipMemcpyAsync(maxlen, &zero, 4, hipMemcpyHostToDevice, ctx->stream));
    const unsigned blocks = unsigned((nrows + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(row_lengths, dim3(blocks), dim3(kBlock), 0, ctx->stream, g->rowptr, g->tiles, cuts, P, maxlen);
    HIP_OK(hipGetLastError());
    uint32_t longest = 0;
    HIP_OK(hipMemcpyAsync(&longest, maxlen, 4, hipMemcpyDeviceToHost, ctx->stream));
    HIP_OK(hipStreamSynchronize(ctx->stream));
    uint32_t S = 4;
    while (S < longest && S < kRowSlotMax) S <<= 1;
    const uint64_t bytes = nrows * S * 4;
    if (longest <= kRowSlotMax && bytes <= kRowTableMaxBytes && !(ctx->budget && ctx->in_use + bytes > ctx->budget)) {
      uint64_t racct = 0;
      rows = dev_alloc<uint32_t>(ctx, nrows * S, &racct);
      rowlen = dev_alloc<uint32_t>(ctx, nrows, &racct);
      hipLaunchKernelGGL(build_rows, dim3(blocks), dim3(kBlock), 0, ctx->stream, g->rowptr, g->cols, g->tiles, cuts,
                         P, uint32_t(__builtin_ctz(S)), reinterpret_cast<u32x4*>(rows), rowlen);
      HIP_OK(hipGetLastError());
      HIP_OK(hipStreamSynchronize(ctx->stream));
      g->rows = rows;
      g->rowlen = rowlen;
      g->rows_S = S;
      g->bytes += racct;
      rows = nullptr;
      rowlen = nullptr;
    }
  } catch (...) {
    (void)hipStreamSynchronize(ctx->stream);
    dev_free(ctx, rows);
    dev_free(ctx, rowlen);
    dev_free(ctx, cuts);
    dev_free(ctx, maxlen);
    ctx->in_use -= std::min(ctx->in_use, acct);
    throw;
  }
  dev_free(ctx, cuts);
  dev_free(ctx, maxlen);
  ctx->in_use -= std::min(ctx->in_use, acct);
}

// Whether BBTC_KERNEL_AUTO takes the row form for a graph whose tasks all
// admit the lane form. Measured on G(n, m) graphs of 2^12-2^22 vertices and
// average degree 4-64 with 1-16 cuts: the row form wins on large graphs
// (n >= 2^20: 2-2.7x on G(n=2^22, m=2^26), 2x on G(n=2^20, m=2^24)), with
// one or two cuts, and at average degree up to 24; on smaller graphs of
// higher degree with 4 or more cuts the tiled lane form is faster (its
// tiles stay in the caches and its per-task lists are shorter, so its waves
// diverge less), by up to 7x at degree 63 with 8 cuts.
bool rows_pay(const bbtc_dgraph* g) { return ::bbtc::rows_pay(g->P, g->htiles, g->deg); }

// Sets up the row form for plan p over refs (a lane-only plan of a graph
// with a row table): the map from lexicographic task index to plan task.
// Leaves p unchanged when the plan lists a task twice.
void plan_rows(bbtc_ctx* ctx, bbtc_plan* p, const std::vector<TaskRef>& refs) {
  const bbtc_dgraph* g = p->g;
  const uint32_t P = g->P;
  const uint64_t ntask = lex_task(P, P - 1, P - 1, P - 1) + 1;
  if (ntask >= 0xFFFFFFFFull) return;
  std::vector<uint32_t> lex2plan(ntask, ~0u);
  for (size_t q = 0; q < refs.size(); ++q) {
    const uint32_t i = refs[q].ab / P, j = refs[q].ab % P, k = refs[q].bc % P;
    uint32_t& slot = lex2plan[lex_task(P, i, j, k)];
    if (slot != ~0u) return;  // listed twice
    slot = uint32_t(q);
  }
  p->lex2plan = dev_alloc<uint32_t>(ctx, lex2plan.size(), &p->bytes);
  p->cuts = dev_alloc<uint32_t>(ctx, P + 1, &p->bytes);
  HIP_OK(hipMemcpyAsync(p->lex2plan, lex2plan.data(), lex2plan.size() * 4, hipMemcpyHostToDevice, ctx->stream));
  HIP_OK(hipMemcpyAsync(p->cuts, g->cuts.data(), (P + 1) * 4, hipMemcpyHostToDevice, ctx->stream));
  p->rows_form = true;
  for (uint64_t q = 0; q < p->ntasks; ++q)
    if (p->hprefix[q + 1] > p->hprefix[q]) p->hkind[q] = kKindRows + 1;
}

// Work statistics of the tasks of a single resident graph, from the table
// gathered at upload (empty when it has none).
std::vector<TaskWork> graph_work(const bbtc_dgraph* g, const std::vector<TaskRef>& refs) {
  std::vector<TaskWork> w;
  if (g->P == 0 || g->lex_work.empty()) return w;
  const uint32_t P = g->P;
  w.resize(refs.size());
  for (size_t q = 0; q < refs.size(); ++q) {
    const uint32_t i = refs[q].ab / P, j = refs[q].ab % P, k = refs[q].bc % P;
    w[q] = g->lex_work[g->lex_ij[size_t(i) * P + j] + (k - j)];
  }
  return w;
}

// The first task that a whole-plan launch's last claim part holds entirely
// (its items and those of every later task lie in that part), from the
// plan's item prefix; n when there is nothing to count or the first phase
// deals chunks round robin.
uint64_t contested_from(const std::vector<uint64_t>& pref, uint64_t n, uint64_t chunk) {
  const uint64_t len = pref[n];
  if (len == 0 || chunk != 0) return n;
  // the last part with items starts at item q * C; the task holding that item
  // may also have items in the part before, so the contest starts after it
  const unsigned long long C = part_chunk(len, 0);
  const uint64_t first = (len - 1) / C * C;
  const uint64_t t = uint64_t(std::upper_bound(pref.begin(), pref.begin() + n + 1, first) - pref.begin()) - 1;
  return t + 1;
}

}  // namespace

// The order in which a plan lists tasks [0, k) of D (tasks [k, n) keep their
// places). Within each form, tasks that share a bc tile (the tile whose rows
// are read at random) come together, the tiles of most estimated work
// first, each tile's tasks in the caller's order; the forms are then
// interleaved in proportion to the AUTO rule's estimated clocks, each task
// placed at the middle of its share of its form's total. Forms are bound by
// different parts of the machine (the rdense probes by 128-byte line
// fetches, the merge searches by instruction issue and LDS latency), so a
// claim part that works through a mix keeps both busy, and the tasks of one
// bc tile run close together, while its rows are in the L2 and MALL. RMAT
// s26 with 32 pal cuts: 542.6 ms against 580.4 ms in descending weight order
// (forms interleaved, each form in descending order: 558.7 ms; its bc tiles
// in index order 555.3, or by (k, j) 551.1 ms; grouped by ab or ac tile
// instead: 675-680 ms, from load imbalance).
std::vector<uint64_t> mixed_order(const std::vector<TileDesc>& tiles, const PlanDesc& D,
                                  const std::vector<TaskWork>& work, const std::vector<TaskRef>& refs, uint64_t k,
                                  uint64_t n) {
  std::vector<double> cost(k, 0.0);
  std::map<uint32_t, double> tile_cost;  // bc tile -> estimated clocks of its tasks
  for (uint64_t q = 0; q < k; ++q) {
    if (D.pref[q + 1] > D.pref[q]) {
      const TaskDesc& d = D.td[q];
      const FormCosts c = form_costs(work[q], tiles[d.ac].dwpr);
      cost[q] = d.kind == kKindBitmap ? c.bitmap
                : d.kind == kKindDense ? c.dense
                : d.kind == kKindRdense ? c.rdense
                                        : c.merge;
      if (!(cost[q] > 0) || !std::isfinite(cost[q])) cost[q] = 0;
    }
    tile_cost[refs[q].bc] += cost[q];
  }
  std::vector<uint64_t> by_tile(k);
  for (uint64_t q = 0; q < k; ++q) by_tile[q] = q;
  std::stable_sort(by_tile.begin(), by_tile.end(), [&](uint64_t a, uint64_t b) {
    const uint32_t ta = refs[a].bc, tb = refs[b].bc;
    if (ta == tb) return false;
    const double ca = tile_cost[ta], cb = tile_cost[tb];
    return ca != cb ? ca > cb : ta < tb;
  });
  // per form (in form order), each task's position: the middle of its share
  std::vector<std::pair<double, uint64_t>> pos;
  pos.reserve(k);
  for (uint32_t kind = 0; kind <= kKindRows + 1; ++kind) {
    double total = 0;
    for (uint64_t q = 0; q < k; ++q)
      if (D.hkind[q] == kind) total += cost[q];
    double run = 0;
    for (uint64_t q : by_tile)
      if (D.hkind[q] == kind) {
        pos.emplace_back(total > 0 ? (run + cost[q] / 2) / total : 0.0, q);
        run += cost[q];
      }
  }
  std::stable_sort(pos.begin(), pos.end(),
                   [](const std::pair<double, uint64_t>& a, const std::pair<double, uint64_t>& b) {
                     return a.first < b.first;
                   });
  std::vector<uint64_t> order(n);
  for (uint64_t q = 0; q < k; ++q) order[q] = pos[q].second;
  for (uint64_t q = k; q < n; ++q) order[q] = q;
  return order;
}

namespace {

// Builds a plan over tile-addressed tasks of a resident graph. work: the
// tasks' statistics for the AUTO form rule (NULL: those of a single graph's
// upload table, if any).
bbtc_plan* make_plan(bbtc_ctx* ctx, bbtc_dgraph* g, const std::vector<TaskRef>& refs_in, bbtc_kernel kernel,
                     const uint8_t* choice, const std::vector<TaskWork>* work = nullptr, MixSpec mix = {}) {
  const uint64_t n = refs_in.size();
  auto* p = new bbtc_plan;
  p->ctx = ctx;
  p->g = g;
  p->ntasks = n;
  try {
    std::vector<TaskWork> own;
    if (!work) {
      own = graph_work(g, refs_in);
      if (!own.empty()) work = &own;
    }
    PlanDesc D = describe_plan(g->htiles, g->deg, refs_in, kernel, choice, true, work);
    const std::vector<TaskRef>* refs_p = &refs_in;
    std::vector<TaskRef> mixed_refs;
    std::vector<TaskWork> mixed_work;
    if (mix.on && kernel == BBTC_KERNEL_AUTO && !choice && work && !D.lane_only && n > 1) {
      uint64_t k = std::min<uint64_t>(mix.upto, n);
      if (mix.contested) k = std::max<uint64_t>(mix.upto, contested_from(D.pref, n, 0));
      k = std::min<uint64_t>(k, n);
      if (k > 1) {
        p->order = mixed_order(g->htiles, D, *work, refs_in, k, n);
        p->mixed = k;
        mixed_refs.resize(n);
        mixed_work.resize(n);
        for (uint64_t q = 0; q < n; ++q) {
          mixed_refs[q] = refs_in[p->order[q]];
          mixed_work[q] = (*work)[p->order[q]];
        }
        refs_p = &mixed_refs;
        work = &mixed_work;
        D = describe_plan(g->htiles, g->deg, mixed_refs, kernel, choice, true, work);
      }
    }
    const std::vector<TaskRef>& refs = *refs_p;
    const bool want_rows = (kernel == BBTC_KERNEL_ROWS || (kernel == BBTC_KERNEL_AUTO && rows_pay(g))) && !choice &&
                           D.lane_only && n > 0 && g->P > 0;
    if (want_rows && !g->rows_tried) build_graph_rows(ctx, g);
    if (D.lane_only && g->P > 0 && !g->slots) build_graph_slots(ctx, g);
    std::vector<TaskDesc>& td = D.td;
    std::vector<uint64_t>& pref = D.pref;
    p->hkind = std::move(D.hkind);
    p->bitmap_tasks = D.bitmap_tasks;
    p->lane_only = D.lane_only;
    p->nitems = pref[n];
    p->hprefix = pref;
    HIP_OK(hipSetDevice(ctx->device));
    p->tasks = dev_alloc<TaskDesc>(ctx, td.size(), &p->bytes);
    p->item_prefix = dev_alloc<uint64_t>(ctx, pref.size(), &p->bytes);
    p->per_task = dev_alloc<unsigned long long>(ctx, std::max<uint64_t>(n, 1), &p->bytes);
    p->ctl = dev_alloc<SlotCtl>(ctx, kMaxSlots, &p->bytes);
    p->spill_cap = uint32_t(std::min<uint64_t>(std::max<uint64_t>(p->nitems, 1), kSpillCap));
    p->spill.resize(1);
    p->spill[0].rec = dev_alloc<SpillRec>(ctx, p->spill_cap, &p->bytes);
    p->spill[0].base = dev_alloc<uint64_t>(ctx, p->spill_cap, &p->bytes);
    p->cycles = dev_alloc<unsigned long long>(ctx, std::max<uint64_t>(n, 1), &p->bytes);
    p->segs = dev_alloc<SegDesc>(ctx, 2, &p->bytes);
    SegDesc sd[2] = {};
    sd[0].rowptr = g->rowptr;
    sd[0].cols = g->cols;
    sd[0].nzrows = g->nzrows;
    sd[0].dense = g->dense;
    sd[0].tiles = g->tiles;
    sd[0].tasks = p->tasks;
    sd[0].item_prefix = p->item_prefix;
    sd[0].per_task = p->per_task;
    sd[0].cycles = nullptr;
    sd[0].slots = g->slots;
    sd[0].item0 = 0;
    sd[0].item1 = p->nitems;
    sd[0].launch0 = 0;
    sd[0].ntasks = uint32_t(n);
    sd[1] = sd[0];
    sd[1].cycles = p->cycles;
    HIP_OK(hipMemcpyAsync(p->tasks, td.data(), td.size() * sizeof(TaskDesc), hipMemcpyHostToDevice, ctx->stream));
    HIP_OK(hipMemcpyAsync(p->item_prefix, pref.data(), pref.size() * 8, hipMemcpyHostToDevice, ctx->stream));
    HIP_OK(hipMemcpyAsync(p->segs, sd, sizeof(sd), hipMemcpyHostToDevice, ctx->stream));
    if (want_rows && g->rows) plan_rows(ctx, p, refs);
    HIP_OK(hipStreamSynchronize(ctx->stream));
  } catch (...) {
    release_plan(p);
    throw;
  }
  return p;
}

// Per-task form from measured clocks: the task list runs once per form
// (each on the tasks it applies to) with per-item clock counters while all
// those tasks share the device; each task keeps its cheapest form. This
// replaces the fixed kernel threshold of bb_tc.cpp:98.
void autotune_refs(bbtc_ctx* ctx, bbtc_dgraph* g, const std::vector<TaskRef>& refs, uint8_t* choice_out,
                   uint64_t* cycles_out) {
  constexpr int kForms = 5;
  const bbtc_kernel forced[kForms] = {BBTC_KERNEL_MERGE, BBTC_KERNEL_BITMAP, BBTC_KERNEL_DENSE,
                                      BBTC_KERNEL_RDENSE, BBTC_KERNEL_LANE};
  const uint64_t n = refs.size();
  const uint64_t none = ~0ull;
  std::vector<uint64_t> best(n, none);
  std::vector<uint8_t> choice(n, 1);
  if (cycles_out)
    for (uint64_t x = 0; x < kForms * n; ++x) cycles_out[x] = none;
  bool all_lane = false;  // the lane form applies to every task with work
  for (int f = 0; f < kForms; ++f) {
    std::vector<uint32_t> kinds;
    {
      bbtc_plan* p = make_plan(ctx, g, refs, forced[f], nullptr);
      kinds = p->hkind;
      if (forced[f] == BBTC_KERNEL_LANE) all_lane = p->lane_only && n > 0;
      if (f == 0) {  // warm caches and code once
        try {
          launch(p, ctx->stream);
          HIP_OK(hipStreamSynchronize(ctx->stream));
        } catch (...) {
          release_plan(p);
          throw;
        }
      }
      release_plan(p);
    }
    std::vector<uint64_t> idx;
    std::vector<TaskRef> sub;
    for (uint64_t q = 0; q < n; ++q)
      if (kinds[q] == uint32_t(f + 1)) {  // the form really applies to this task
        idx.push_back(q);
        sub.push_back(refs[q]);
      }
    if (sub.empty()) continue;
    bbtc_plan* p = make_plan(ctx, g, sub, forced[f], nullptr);
    std::vector<uint64_t> cyc(sub.size());
    const bbtc_status st = bbtc_plan_profile(p, cyc.data());
    release_plan(p);
    if (st != BBTC_OK) fail(st, bbtc_last_error());
    for (size_t x = 0; x < idx.size(); ++x) {
      const uint64_t q = idx[x];
      if (cycles_out) cycles_out[kForms * q + f] = cyc[x];
      // The lane form runs in its own kernel only when every task of a
      // launch uses it, so it is weighed as a whole plan below rather than
      // per task.
      if (forced[f] != BBTC_KERNEL_LANE && cyc[x] < best[q]) {
        best[q] = cyc[x];
        choice[q] = uint8_t(f + 1);
      }
    }
  }
  if (all_lane) {
    // the per-task choice against every task in the lane form, by launch time
    auto fastest_ms = [&](bbtc_kernel k, const uint8_t* ch) {
      bbtc_plan* p = make_plan(ctx, g, refs, k, ch);
      float ms[3] = {0.f, 0.f, 0.f};
      const bbtc_status st = bbtc_plan_time(p, 3, ms);
      release_plan(p);
      if (st != BBTC_OK) fail(st, bbtc_last_error());
      return std::min(ms[1], ms[2]);  // the first launch warms caches
    };
    const float per_task = fastest_ms(BBTC_KERNEL_AUTO, choice.data());
    const float lanes = fastest_ms(BBTC_KERNEL_LANE, nullptr);
    if (lanes < per_task) std::fill(choice.begin(), choice.end(), uint8_t(5));
  }
  // Rows per item: the chosen forms counted with 64 (or as many as fit the
  // LDS slice), half and a quarter of that many rows per work item, all tasks
  // sharing the device in each pass; each task keeps its cheapest.
  if (n > 0) {
    std::vector<uint64_t> best_rows(n, none);
    std::vector<uint8_t> shift(n, 0);
    for (uint32_t sh = 0; sh < kRowsShifts; ++sh) {
      std::vector<uint8_t> ch(n);
      for (uint64_t q = 0; q < n; ++q) ch[q] = uint8_t(choice[q] | (sh << kChoiceRowsShift));
      bbtc_plan* p = make_plan(ctx, g, refs, BBTC_KERNEL_AUTO, ch.data());
      std::vector<uint64_t> cyc(n);
      const bbtc_status st = bbtc_plan_profile(p, cyc.data());
      release_plan(p);
      if (st != BBTC_OK) fail(st, bbtc_last_error());
      for (uint64_t q = 0; q < n; ++q)
        if (cyc[q] < best_rows[q]) {
          best_rows[q] = cyc[q];
          shift[q] = uint8_t(sh);
        }
    }
    for (uint64_t q = 0; q < n; ++q) choice[q] = uint8_t(choice[q] | (shift[q] << kChoiceRowsShift));
  }
  std::copy(choice.begin(), choice.end(), choice_out);
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

bool graph_task_work(const bbtc_dgraph* g, const bbtc_task* tasks, uint64_t n, double* out) {
  if (g->P == 0 || g->lex_work.empty()) return false;
  const uint32_t P = g->P;
  for (uint64_t q = 0; q < n; ++q) {
    const bbtc_task& x = tasks[q];
    const TaskWork& s = g->lex_work[g->lex_ij[size_t(x.i) * P + x.j] + (x.k - x.j)];
    out[q] = 16.0 * s.R + 12.0 * s.W + 4.0 * s.A + 4.0 * s.Pbc;
  }
  return true;
}

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

bool host_atomics(bbtc_ctx* ctx) {
  if (ctx->host_atomics < 0) {
    int v = 0;
    HIP_OK(hipSetDevice(ctx->device));
    HIP_OK(hipDeviceGetAttribute(&v, hipDeviceAttributeHostNativeAtomicSupported, ctx->device));
    ctx->host_atomics = v ? 1 : 0;
  }
  return ctx->host_atomics == 1;
}

uint64_t* queue_alloc(bbtc_ctx* ctx) {
  void* p = nullptr;
  HIP_OK(hipSetDevice(ctx->device));
  HIP_OK(hipHostMalloc(&p, 128, hipHostMallocMapped | hipHostMallocCoherent));
  return static_cast<uint64_t*>(p);
}

void queue_free(uint64_t* q) {
  if (q) (void)hipHostFree(q);
}

bbtc_plan* plan_create_split(bbtc_ctx* ctx, bbtc_dgraph* g, const bbtc_task* tasks, uint64_t n, bbtc_kernel kernel,
                             const uint8_t* choice, uint64_t upto, bool contested) {
  if (!ctx || !g || (!tasks && n)) fail(BBTC_ERR_INVALID, "null argument");
  if (g->ctx != ctx) fail(BBTC_ERR_INVALID, "graph belongs to another context");
  if (n >= 0xFFFFFFFFull) fail(BBTC_ERR_INVALID, "too many tasks");
  if (!valid_kernel(kernel)) fail(BBTC_ERR_INVALID, "unknown kernel choice");
  for (uint64_t q = 0; q < n; ++q)
    if (!(tasks[q].i <= tasks[q].j && tasks[q].j <= tasks[q].k && tasks[q].k < g->P))
      fail(BBTC_ERR_INVALID, "task out of range or not i <= j <= k");
  MixSpec mix;
  mix.on = choice == nullptr;
  mix.upto = upto;
  mix.contested = contested;
  return make_plan(ctx, g, refs_of(g->P, tasks, n), kernel, choice, nullptr, mix);
}

uint64_t plan_mixed(const bbtc_plan* p) { return p->mixed; }

uint64_t plan_contested_from(const bbtc_plan* p) {
  // the tasks the plan reordered are never shared with the host
  return std::max<uint64_t>(contested_from(p->hprefix, p->ntasks, p->chunk), p->mixed);
}

void plan_launch_shared(bbtc_plan* p, uint64_t* queue, uint32_t queue_lo) {
  HIP_OK(hipSetDevice(p->ctx->device));
  void* dq = nullptr;
  HIP_OK(hipHostGetDevicePointer(&dq, queue, 0));
  // a claim reaches one item batch per resident wave past the task that needs it
  const uint64_t ahead = resident_blocks(p->ctx, p->lane_only) * kWavesPerBlock * kItemBatch;
  launch_range(p, 0, p->ntasks, 0, p->ctx->stream, static_cast<unsigned long long*>(dq), queue_lo,
               uint32_t(std::min<uint64_t>(ahead, 0xFFFFFFFFull)));
}

void plan_copy_counts(bbtc_plan* p, uint64_t* out) {
  HIP_OK(hipSetDevice(p->ctx->device));
  HIP_OK(hipDeviceSynchronize());
  if (!p->ntasks) return;
  std::vector<uint64_t> h(p->ntasks);
  HIP_OK(hipMemcpy(h.data(), p->per_task, p->ntasks * 8, hipMemcpyDeviceToHost));
  to_caller_order(p, h.data(), out);
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
    HIP_OK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&c->blocks_per_cu_lane, bbtc::count_probe_lanes,
                                                        bbtc::kBlock, 0));
    if (c->blocks_per_cu_lane < 1) c->blocks_per_cu_lane = 1;
    HIP_OK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    c->scratch = static_cast<unsigned char*>(bbtc::scratch_of_device(device));
  } catch (...) {
    if (c->stream) (void)hipStreamDestroy(c->stream);
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
  bbtc::staging_release(ctx);
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

bbtc_status bbtc_ctx_read_bandwidth(bbtc_ctx* ctx, uint64_t bytes, int iters, double* gbps_out) {
  DEV_REQUIRE(ctx && iters > 0 && gbps_out, "bad argument");
  DEV_REQUIRE(bytes >= (1ull << 20), "read at least 1 MiB");
  DEV_TRY
  HIP_OK(hipSetDevice(ctx->device));
  const uint64_t words = bytes / 16;
  const unsigned blocks = unsigned(std::max(1, ctx->num_cu) * 64);
  uint64_t acct = 0;
  auto* buf = bbtc::dev_alloc<bbtc::u32x4>(ctx, words, &acct);
  unsigned* sink = nullptr;
  std::vector<hipEvent_t> ev(2 * size_t(iters) + 2, nullptr);
  auto cleanup = [&] {
    for (auto& e : ev)
      if (e) (void)hipEventDestroy(e);
    if (sink) (void)hipFree(sink);
    bbtc::dev_free(ctx, buf);
    ctx->in_use -= std::min(ctx->in_use, acct);
  };
  try {
    HIP_OK(hipMalloc(&sink, size_t(blocks) * sizeof(unsigned)));
    hipStream_t s = ctx->stream;
    HIP_OK(hipMemsetAsync(buf, 0, words * 16, s));
    for (auto& e : ev) HIP_OK(hipEventCreate(&e));
    // one untimed pass, then the timed ones
    hipLaunchKernelGGL(bbtc::stream_read, dim3(blocks), dim3(bbtc::kBlock), 0, s, buf, words, sink);
    HIP_OK(hipGetLastError());
    for (int i = 0; i < iters; ++i) {
      HIP_OK(hipEventRecord(ev[2 * i], s));
      hipLaunchKernelGGL(bbtc::stream_read, dim3(blocks), dim3(bbtc::kBlock), 0, s, buf, words, sink);
      HIP_OK(hipEventRecord(ev[2 * i + 1], s));
    }
    HIP_OK(hipStreamSynchronize(s));
    for (int i = 0; i < iters; ++i) {
      float ms = 0.f;
      HIP_OK(hipEventElapsedTime(&ms, ev[2 * i], ev[2 * i + 1]));
      gbps_out[i] = double(words) * 16.0 / (double(ms) * 1e-3) / 1e9;
    }
  } catch (...) {
    cleanup();
    throw;
  }
  cleanup();
  DEV_CATCH
}

bbtc_status bbtc_tasks_shard(const bbtc_dgraph* g, const bbtc_task* tasks, uint64_t n, uint32_t shards,
                             uint32_t* shard_out, double* cost_out) {
  DEV_REQUIRE(g && (tasks || !n) && (shard_out || !n), "null argument");
  DEV_REQUIRE(shards >= 1, "shards must be at least 1");
  for (uint64_t q = 0; q < n; ++q)
    DEV_REQUIRE(tasks[q].i <= tasks[q].j && tasks[q].j <= tasks[q].k && tasks[q].k < g->P,
                "task out of range or not i <= j <= k");
  DEV_TRY
  const uint32_t P = g->P;
  std::vector<double> cost(n, 0.0);
  for (uint64_t q = 0; q < n; ++q) {
    const bbtc_task& x = tasks[q];
    const size_t ab = size_t(x.i) * P + x.j, bc = size_t(x.j) * P + x.k, ac = size_t(x.i) * P + x.k;
    if (g->deg[ab] == 0 || g->deg[bc] == 0 || g->deg[ac] == 0) continue;
    if (!g->lex_work.empty()) {
      const uint32_t wpr = g->htiles[ac].dwpr;
      const uint32_t fit = wpr ? uint32_t(bbtc::kLdsWords) / wpr : uint32_t(bbtc::kWave);
      const bool has_dense = g->htiles[bc].dense != ~0ull;
      (void)bbtc::auto_kind(g->lex_work[g->lex_ij[size_t(x.i) * P + x.j] + (x.k - x.j)], wpr, fit, has_dense,
                            &cost[q]);
    } else {
      // the reference's task weight nnz(ab) * max(deg(bc), deg(ac)) (tasks.hpp:21-23)
      cost[q] = double(g->deg[ab]) * g->htiles[ab].h * std::max(g->deg[bc], g->deg[ac]);
    }
  }
  // longest first onto the least loaded share (the lowest-numbered of equal
  // ones). Preferring, among shares within 1% of the least load, the one
  // that held most of a task's tiles cut the tile bytes of a GPU's share at
  // RMAT s26 with 8 shares from 6.3 to 4.0 GB, but its slowest share took
  // 81.9 against 75.5 ms: tasks that share tiles share their estimates'
  // errors. Every share fits one GPU's 288 GB whole, so balance wins.
  std::vector<uint64_t> order(n);
  for (uint64_t q = 0; q < n; ++q) order[q] = q;
  std::stable_sort(order.begin(), order.end(), [&](uint64_t a, uint64_t b) { return cost[a] > cost[b]; });
  std::vector<double> load(shards, 0.0);
  for (uint64_t q : order) {
    const uint32_t s = uint32_t(std::min_element(load.begin(), load.end()) - load.begin());
    shard_out[q] = s;
    load[s] += cost[q];
  }
  if (cost_out) std::copy(cost.begin(), cost.end(), cost_out);
  DEV_CATCH
}

bbtc_status bbtc_layout_pin(bbtc_ctx* ctx, bbtc_layout* Lh) {
  DEV_REQUIRE(ctx && Lh, "null argument");
  DEV_TRY
  bbtc::require_host_arrays(Lh->L, "bbtc_layout_pin");
  HIP_OK(hipSetDevice(ctx->device));
  bbtc::pin_layout(Lh->L);
  DEV_CATCH
}

int bbtc_layout_pinned(const bbtc_layout* Lh) {
  if (!Lh || !Lh->L.host_arrays) return 0;
  const bbtc::Layout& L = Lh->L;
  return (L.rowptr.size() == 0 || L.rowptr.pinned()) && (L.cols.size() == 0 || L.cols.pinned()) &&
                 (L.nzrows.size() == 0 || L.nzrows.pinned())
             ? 1
             : 0;
}

bbtc_status bbtc_graph_upload(bbtc_ctx* ctx, const bbtc_layout* Lh, bbtc_dgraph** out) {
  DEV_REQUIRE(ctx && Lh && out, "null argument");
  DEV_TRY
  const bbtc::Layout& L = Lh->L;
  bbtc::require_host_arrays(L, "bbtc_graph_upload");
  auto* g = new bbtc_dgraph;
  g->ctx = ctx;
  g->P = L.P;
  try {
    const size_t T = size_t(L.P) * L.P;
    g->htiles = bbtc::tile_table(L);
    g->deg = L.deg;
    g->dwpr = L.dwpr;
    if (L.P <= bbtc::kRuleMaxCuts) {
      // every task's work statistics, for the AUTO form rule of later plans
      const uint32_t P = L.P;
      g->lex_ij.assign(size_t(P) * P, 0);
      std::vector<bbtc_task> all;
      for (uint32_t i = 0; i < P; ++i)
        for (uint32_t j = i; j < P; ++j) {
          g->lex_ij[size_t(i) * P + j] = all.size();
          for (uint32_t k = j; k < P; ++k) all.push_back(bbtc_task{i, j, k});
        }
      g->lex_work = bbtc::task_work(L, all.data(), all.size());
    }
    g->rowptr = bbtc::dev_alloc<uint32_t>(ctx, L.rowptr.size(), &g->bytes);
    g->rowptr_words = L.rowptr.size();
    g->cols = bbtc::dev_alloc<uint32_t>(ctx, L.cols.size(), &g->bytes);
    g->nzrows = bbtc::dev_alloc<uint32_t>(ctx, L.nzrows.size(), &g->bytes);
    g->dense = bbtc::dev_alloc<uint32_t>(ctx, L.dense_words, &g->bytes);
    g->tiles = bbtc::dev_alloc<bbtc::TileDesc>(ctx, T, &g->bytes);
    bbtc::copy_to_device(ctx, ctx->stream, g->rowptr, L.rowptr.data(), L.rowptr.size() * 4, L.rowptr.pinned());
    bbtc::copy_to_device(ctx, ctx->stream, g->cols, L.cols.data(), L.cols.size() * 4, L.cols.pinned());
    bbtc::copy_to_device(ctx, ctx->stream, g->nzrows, L.nzrows.data(), L.nzrows.size() * 4, L.nzrows.pinned());
    HIP_OK(hipMemcpyAsync(g->tiles, g->htiles.data(), T * sizeof(bbtc::TileDesc), hipMemcpyHostToDevice, ctx->stream));
    // bitmap rows of the dense tiles, built from the uploaded tiles
    uint64_t rows = 0, acct = 0;
    const std::vector<bbtc::DenseJob> jobs = bbtc::dense_jobs(L, 0, 0, 0, 0, &rows);
    bbtc::DenseJob* jd = jobs.empty() ? nullptr : bbtc::dev_alloc<bbtc::DenseJob>(ctx, jobs.size(), &acct);
    try {
      if (jd)
        HIP_OK(hipMemcpyAsync(jd, jobs.data(), jobs.size() * sizeof(bbtc::DenseJob), hipMemcpyHostToDevice,
                              ctx->stream));
      bbtc::enqueue_dense_build(g->rowptr, g->cols, g->dense, L.dense_words, jd, jobs.size(), rows, ctx->stream);
      HIP_OK(hipStreamSynchronize(ctx->stream));
    } catch (...) {
      (void)hipStreamSynchronize(ctx->stream);
      bbtc::dev_free(ctx, jd);
      ctx->in_use -= std::min(ctx->in_use, acct);
      throw;
    }
    bbtc::dev_free(ctx, jd);
    ctx->in_use -= std::min(ctx->in_use, acct);
  } catch (...) {
    // copies straight from a pinned layout may still be in flight
    (void)hipStreamSynchronize(ctx->stream);
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
  DEV_REQUIRE(bbtc::valid_kernel(kernel), "unknown kernel choice");
  for (uint64_t q = 0; q < n; ++q)
    DEV_REQUIRE(tasks[q].i <= tasks[q].j && tasks[q].j <= tasks[q].k && tasks[q].k < g->P,
                "task out of range or not i <= j <= k");
  DEV_TRY
  // the plan lists AUTO tasks in its own order (mixed_order); results come
  // back in the caller's
  bbtc::MixSpec mix;
  mix.on = choice == nullptr;
  mix.upto = n;
  *out = bbtc::make_plan(ctx, g, bbtc::refs_of(g->P, tasks, n), kernel, choice, nullptr, mix);
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
  if (per_task && plan->ntasks) bbtc::to_caller_order(plan, h.data(), per_task);
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
    std::vector<float> ms(n);
    for (uint64_t t = 0; t < n; ++t) HIP_OK(hipEventElapsedTime(&ms[t], ev[t], ev[t + 1]));
    if (n) bbtc::to_caller_order(plan, ms.data(), ms_per_task);
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
  std::vector<uint64_t> h(n, 0);
  if (plan->ntasks) HIP_OK(hipMemcpyAsync(h.data(), plan->cycles, plan->ntasks * 8, hipMemcpyDeviceToHost, s));
  HIP_OK(hipStreamSynchronize(s));
  if (plan->ntasks) bbtc::to_caller_order(plan, h.data(), cycles_per_task);
  DEV_CATCH
}

bbtc_status bbtc_autotune(bbtc_ctx* ctx, bbtc_dgraph* g, const bbtc_task* tasks, uint64_t n,
                          uint8_t* choice_out, uint64_t* cycles_out) {
  DEV_REQUIRE(ctx && g && (tasks || n == 0) && (choice_out || n == 0), "null argument");
  DEV_REQUIRE(g->ctx == ctx, "graph belongs to another context");
  for (uint64_t q = 0; q < n; ++q)
    DEV_REQUIRE(tasks[q].i <= tasks[q].j && tasks[q].j <= tasks[q].k && tasks[q].k < g->P,
                "task out of range or not i <= j <= k");
  DEV_TRY
  bbtc::autotune_refs(ctx, g, bbtc::refs_of(g->P, tasks, n), choice_out, cycles_out);
  DEV_CATCH
}

bbtc_status bbtc_plan_task_kinds(const bbtc_plan* plan, uint8_t* kinds) {
  DEV_REQUIRE(plan && (kinds || plan->ntasks == 0), "null argument");
  for (uint64_t t = 0; t < plan->ntasks; ++t)
    kinds[plan->order.empty() ? t : plan->order[t]] = uint8_t(plan->hkind[t]);
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

bbtc_status bbtc_count_noalloc(bbtc_ctx* ctx, bbtc_dgraph* g, const bbtc_task* tasks, uint64_t n,
                               bbtc_kernel kernel, uint64_t* per_task, uint64_t* total, void* stream) {
  DEV_REQUIRE(ctx && g && (tasks || n == 0) && (per_task || total), "null argument");
  DEV_REQUIRE(g->ctx == ctx, "graph belongs to another context");
  DEV_REQUIRE(bbtc::valid_kernel(kernel), "unknown kernel choice");
  if (n > bbtc::kNoAllocMaxTasks) {
    bbtc::set_last_error("bbtc_count_noalloc: " + std::to_string(n) + " tasks exceed the scratch limit of " +
                         std::to_string(bbtc::kNoAllocMaxTasks));
    return BBTC_ERR_INVALID;
  }
  for (uint64_t q = 0; q < n; ++q)
    DEV_REQUIRE(tasks[q].i <= tasks[q].j && tasks[q].j <= tasks[q].k && tasks[q].k < g->P,
                "task out of range or not i <= j <= k");
  DEV_TRY
  HIP_OK(hipSetDevice(ctx->device));
  if (!ctx->scratch) bbtc::fail(BBTC_ERR_INTERNAL, "scratch address not resolved");
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
  const std::vector<bbtc::TaskRef> refs = bbtc::refs_of(g->P, tasks, n);
  const std::vector<bbtc::TaskWork> work = bbtc::graph_work(g, refs);
  bbtc::PlanDesc D =
      bbtc::describe_plan(g->htiles, g->deg, refs, kernel, nullptr, true, work.empty() ? nullptr : &work);
  // the tasks in the order bbtc_count's plan runs them; counts go back to
  // the caller's order
  std::vector<uint64_t> order;
  if (kernel == BBTC_KERNEL_AUTO && !work.empty() && !D.lane_only && n > 1) {
    order = bbtc::mixed_order(g->htiles, D, work, refs, n, n);
    std::vector<bbtc::TaskRef> mrefs(n);
    std::vector<bbtc::TaskWork> mwork(n);
    for (uint64_t q = 0; q < n; ++q) {
      mrefs[q] = refs[order[q]];
      mwork[q] = work[order[q]];
    }
    D = bbtc::describe_plan(g->htiles, g->deg, mrefs, kernel, nullptr, true, &mwork);
  }
  const uint64_t nitems = D.pref[n];
  unsigned char* base = ctx->scratch;
  auto* ctl = reinterpret_cast<bbtc::SlotCtl*>(base + bbtc::kScratchCtl);
  auto* seg = reinterpret_cast<bbtc::SegDesc*>(base + bbtc::kScratchSeg);
  auto* td = reinterpret_cast<bbtc::TaskDesc*>(base + bbtc::kScratchTasks);
  auto* pref = reinterpret_cast<uint64_t*>(base + bbtc::kScratchPrefix);
  auto* counts = reinterpret_cast<unsigned long long*>(base + bbtc::kScratchCounts);
  auto* spill = reinterpret_cast<bbtc::SpillRec*>(base + bbtc::kScratchSpill);
  auto* spill_base = reinterpret_cast<uint64_t*>(base + bbtc::kScratchSpillBase);
  bbtc::SegDesc sd{};
  sd.rowptr = g->rowptr;
  sd.cols = g->cols;
  sd.nzrows = g->nzrows;
  sd.dense = g->dense;
  sd.tiles = g->tiles;
  sd.tasks = td;
  sd.item_prefix = pref;
  sd.per_task = counts;
  sd.cycles = nullptr;
  sd.slots = g->slots;  // built by an earlier plan, if any (never here)
  sd.item0 = 0;
  sd.item1 = nitems;
  sd.launch0 = 0;
  sd.ntasks = uint32_t(n);
  if (n) HIP_OK(hipMemcpyAsync(td, D.td.data(), n * sizeof(bbtc::TaskDesc), hipMemcpyHostToDevice, s));
  HIP_OK(hipMemcpyAsync(pref, D.pref.data(), (n + 1) * 8, hipMemcpyHostToDevice, s));
  HIP_OK(hipMemcpyAsync(seg, &sd, sizeof(sd), hipMemcpyHostToDevice, s));
  // the array keeps the previous call's counters: clear them (a fill kernel)
  bbtc::enqueue_zero(counts, n, s);
  if (nitems) {
    bbtc::CountArgs a{};
    a.segs = seg;
    a.nseg = 1;
    a.item_end = nitems;
    a.ctl = ctl;  // reset by the launch sequence
    a.spill = spill;
    a.spill_base = spill_base;
    a.spill_cap = uint32_t(std::min<uint64_t>(nitems, bbtc::kNoAllocSpillCap));
    bbtc::enqueue_count(ctx, a, 0, nitems, D.lane_only, 0, s);
  }
  uint64_t t = 0;
  if (per_task) {
    if (n) HIP_OK(hipMemcpyAsync(per_task, counts, n * 8, hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    for (uint64_t q = 0; q < n; ++q) t += per_task[q];
    if (!order.empty()) {
      const std::vector<uint64_t> mine(per_task, per_task + n);
      for (uint64_t q = 0; q < n; ++q) per_task[order[q]] = mine[q];
    }
  } else {
    // the total without a host array: counters copied back a block at a time
    uint64_t buf[512];
    for (uint64_t q0 = 0; q0 < n; q0 += 512) {
      const uint64_t len = std::min<uint64_t>(512, n - q0);
      HIP_OK(hipMemcpyAsync(buf, counts + q0, len * 8, hipMemcpyDeviceToHost, s));
      HIP_OK(hipStreamSynchronize(s));
      for (uint64_t q = 0; q < len; ++q) t += buf[q];
    }
    HIP_OK(hipStreamSynchronize(s));
  }
  if (total) *total = t;
  DEV_CATCH
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Batches: many graphs, one arena, one persistent launch
// ---------------------------------------------------------------------------
struct bbtc_batch {
  bbtc_ctx* ctx = nullptr;
  bbtc_dgraph* arena = nullptr;  // all graphs' arrays; tile ids are global
  bbtc_plan* plan = nullptr;
  uint64_t ngraphs = 0;
  std::vector<bbtc::TaskRef> refs;     // tasks in launch order
  std::vector<bbtc::TaskWork> work;    // their work statistics (AUTO form rule)
  std::vector<uint32_t> task_ijk;      // i, j, k of each task (3 per task)
  std::vector<unsigned long long> hleft0;
  uint64_t bytes = 0;
  unsigned long long* left0 = nullptr;
  unsigned long long* left = nullptr;
  unsigned long long* done = nullptr;
  unsigned long long* start = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  double wall_khz = 0;
};

namespace bbtc {
namespace {

void batch_drop_capture(bbtc_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
}

void batch_release(bbtc_batch* b) {
  if (!b) return;
  batch_drop_capture(b);
  release_plan(b->plan);
  release_graph(b->arena);
  if (b->ctx) {
    dev_free(b->ctx, b->left0);
    dev_free(b->ctx, b->left);
    dev_free(b->ctx, b->done);
    dev_free(b->ctx, b->start);
    b->ctx->in_use -= std::min(b->ctx->in_use, b->bytes);
  }
  delete b;
}

// New plan for the batch's tasks (choice may be NULL) and per-graph item
// counts for completion tracking.
void batch_replan(bbtc_batch* b, bbtc_kernel kernel, const uint8_t* choice) {
  batch_drop_capture(b);
  if (b->plan) {
    release_plan(b->plan);
    b->plan = nullptr;
  }
  b->plan = make_plan(b->ctx, b->arena, b->refs, kernel, choice, &b->work);
  b->plan->chunk = kListChunkItems;  // graphs early in the list finish early
  b->hleft0.assign(std::max<uint64_t>(b->ngraphs, 1), 0);
  for (uint64_t q = 0; q < b->refs.size(); ++q)
    b->hleft0[b->refs[q].graph] += b->plan->hprefix[q + 1] - b->plan->hprefix[q];
  HIP_OK(hipMemcpyAsync(b->left0, b->hleft0.data(), b->hleft0.size() * 8, hipMemcpyHostToDevice,
                        b->ctx->stream));
  HIP_OK(hipStreamSynchronize(b->ctx->stream));
  b->plan->track_left = b->left;
  b->plan->track_done = b->done;
}

// The launch sequence of one batch count (no host work, no memsets: it can
// be captured in a graph and replayed).
void batch_enqueue(bbtc_batch* b, hipStream_t s) {
  const uint64_t ng = std::max<uint64_t>(b->ngraphs, 1);
  hipLaunchKernelGGL(batch_begin, dim3(unsigned((ng + 255) / 256)), dim3(256), 0, s, b->left0, b->ngraphs,
                     b->left, b->done, b->start);
  HIP_OK(hipGetLastError());
  zero(b->plan, s);
  launch_range(b->plan, 0, b->plan->ntasks, 0, s);
}

}  // namespace
}  // namespace bbtc

extern "C" {

bbtc_status bbtc_batch_create(bbtc_ctx* ctx, const bbtc_layout* const* layouts, uint64_t ngraphs,
                              bbtc_kernel kernel, bbtc_batch_order order_kind, bbtc_batch** out) {
  DEV_REQUIRE(ctx && out && (layouts || ngraphs == 0), "null argument");
  DEV_REQUIRE(order_kind == BBTC_BATCH_HEAVY_FIRST || order_kind == BBTC_BATCH_SMALL_GRAPHS_FIRST,
              "unknown batch order");
  DEV_REQUIRE(bbtc::valid_kernel(kernel), "unknown kernel choice");
  DEV_REQUIRE(ngraphs < 0xFFFFFFFFull, "too many graphs");
  for (uint64_t gi = 0; gi < ngraphs; ++gi) DEV_REQUIRE(layouts[gi], "null layout");
  DEV_TRY
  HIP_OK(hipSetDevice(ctx->device));
  auto* b = new bbtc_batch;
  b->ctx = ctx;
  b->ngraphs = ngraphs;
  try {
    int khz = 0;
    HIP_OK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, ctx->device));
    b->wall_khz = khz > 0 ? double(khz) : 1e5;
    // arena sizes and per-graph bases
    uint64_t n_rp = 0, n_col = 0, n_nz = 0, n_dense = 0, n_tiles = 0;
    std::vector<uint64_t> rp0(ngraphs), col0(ngraphs), nz0(ngraphs), dn0(ngraphs), tile0(ngraphs);
    for (uint64_t gi = 0; gi < ngraphs; ++gi) {
      const bbtc::Layout& L = layouts[gi]->L;
      bbtc::require_host_arrays(L, "bbtc_batch_create");
      rp0[gi] = n_rp;
      col0[gi] = n_col;
      nz0[gi] = n_nz;
      dn0[gi] = n_dense;
      tile0[gi] = n_tiles;
      n_rp += L.rowptr.size();
      n_col += L.cols.size();
      n_nz += L.nzrows.size();
      n_dense += L.dense_words;
      n_tiles += uint64_t(L.P) * L.P;
    }
    if (n_tiles >= 0xFFFFFFFFull) bbtc::fail(BBTC_ERR_INVALID, "too many tiles in one batch");
    auto* A = new bbtc_dgraph;
    A->ctx = ctx;
    A->P = 0;
    b->arena = A;
    A->htiles.resize(std::max<uint64_t>(n_tiles, 1));
    A->deg.assign(std::max<uint64_t>(n_tiles, 1), 0.f);
    A->rowptr = bbtc::dev_alloc<uint32_t>(ctx, n_rp, &A->bytes);
    A->cols = bbtc::dev_alloc<uint32_t>(ctx, n_col, &A->bytes);
    A->nzrows = bbtc::dev_alloc<uint32_t>(ctx, n_nz, &A->bytes);
    A->dense = bbtc::dev_alloc<uint32_t>(ctx, n_dense, &A->bytes);
    A->tiles = bbtc::dev_alloc<bbtc::TileDesc>(ctx, std::max<uint64_t>(n_tiles, 1), &A->bytes);
    struct Order {
      float w;
      uint64_t m;  // edges of the task's graph
      uint64_t q;
    };
    std::vector<Order> order;
    std::vector<bbtc::TaskWork> work;  // per task, in the order of b->refs before sorting
    std::vector<bbtc::DenseJob> jobs;  // bitmap rows of every graph's dense tiles
    uint64_t job_rows = 0;
    for (uint64_t gi = 0; gi < ngraphs; ++gi) {
      const bbtc::Layout& L = layouts[gi]->L;
      const uint32_t P = L.P;
      for (uint32_t i = 0; i < P; ++i)
        for (uint32_t k = 0; k < P; ++k) {
          const size_t t = L.tile(i, k);
          bbtc::TileDesc d{};
          d.h = L.height(i);
          d.dense = ~0ull;
          d.dwpr = L.dwpr[k];
          if (i <= k) {
            d.rp = rp0[gi] + L.rp_base[t];
            d.col = col0[gi] + L.col_base[t];
            d.nz = nz0[gi] + L.nz_base[t];
            d.nzr = L.nz_rows[t];
            d.maxd = L.max_row[t];
            if (L.dense_base[t] != ~0ull) d.dense = dn0[gi] + L.dense_base[t];
          }
          A->htiles[tile0[gi] + t] = d;
          A->deg[tile0[gi] + t] = L.deg[t];
        }
      hipStream_t s = ctx->stream;
      bbtc::copy_to_device(ctx, s, A->rowptr + rp0[gi], L.rowptr.data(), L.rowptr.size() * 4, L.rowptr.pinned());
      bbtc::copy_to_device(ctx, s, A->cols + col0[gi], L.cols.data(), L.cols.size() * 4, L.cols.pinned());
      bbtc::copy_to_device(ctx, s, A->nzrows + nz0[gi], L.nzrows.data(), L.nzrows.size() * 4, L.nzrows.pinned());
      uint64_t rows = 0;
      const std::vector<bbtc::DenseJob> gj = bbtc::dense_jobs(L, rp0[gi], col0[gi], dn0[gi], job_rows, &rows);
      jobs.insert(jobs.end(), gj.begin(), gj.end());
      job_rows += rows;
      // this graph's tasks, weighted as tasks.hpp:21-23 for the cross-graph order
      const size_t first = b->refs.size();
      for (uint32_t i = 0; i < P; ++i)
        for (uint32_t j = i; j < P; ++j)
          for (uint32_t k = j; k < P; ++k) {
            const uint32_t base = uint32_t(tile0[gi]);
            b->refs.push_back(bbtc::TaskRef{base + i * P + j, base + j * P + k, base + i * P + k, uint32_t(gi)});
            b->task_ijk.insert(b->task_ijk.end(), {i, j, k});
            const uint32_t nnz_ab = uint32_t(L.nnz[L.tile(i, j)]);
            const float w = nnz_ab * std::max(L.deg[L.tile(j, k)], L.deg[L.tile(i, k)]);
            order.push_back(Order{w, L.m, b->refs.size() - 1});
          }
      std::vector<bbtc_task> mine;
      for (size_t q = first; q < b->refs.size(); ++q)
        mine.push_back(bbtc_task{b->task_ijk[3 * q], b->task_ijk[3 * q + 1], b->task_ijk[3 * q + 2]});
      const std::vector<bbtc::TaskWork> w = bbtc::task_work(L, mine.data(), mine.size());
      work.insert(work.end(), w.begin(), w.end());
    }
    HIP_OK(hipMemcpyAsync(A->tiles, A->htiles.data(), A->htiles.size() * sizeof(bbtc::TileDesc),
                          hipMemcpyHostToDevice, ctx->stream));
    {
      uint64_t acct = 0;
      bbtc::DenseJob* jd = jobs.empty() ? nullptr : bbtc::dev_alloc<bbtc::DenseJob>(ctx, jobs.size(), &acct);
      try {
        if (jd)
          HIP_OK(hipMemcpyAsync(jd, jobs.data(), jobs.size() * sizeof(bbtc::DenseJob), hipMemcpyHostToDevice,
                                ctx->stream));
        bbtc::enqueue_dense_build(A->rowptr, A->cols, A->dense, std::max<uint64_t>(n_dense, 1), jd, jobs.size(),
                                  job_rows, ctx->stream);
        HIP_OK(hipStreamSynchronize(ctx->stream));
      } catch (...) {
        (void)hipStreamSynchronize(ctx->stream);
        bbtc::dev_free(ctx, jd);
        ctx->in_use -= std::min(ctx->in_use, acct);
        throw;
      }
      bbtc::dev_free(ctx, jd);
      ctx->in_use -= std::min(ctx->in_use, acct);
    }
    HIP_OK(hipStreamSynchronize(ctx->stream));
    // heavy tasks first (NaN weights of empty bands last), over all graphs or
    // within each graph with the smallest graphs first
    std::stable_sort(order.begin(), order.end(), [order_kind](const Order& l, const Order& r) {
      if (order_kind == BBTC_BATCH_SMALL_GRAPHS_FIRST && l.m != r.m) return l.m < r.m;
      const float a = l.w == l.w ? l.w : -1.f, c = r.w == r.w ? r.w : -1.f;
      return a > c;
    });
    std::vector<bbtc::TaskRef> refs(order.size());
    std::vector<uint32_t> ijk(3 * order.size());
    b->work.resize(order.size());
    for (size_t x = 0; x < order.size(); ++x) {
      refs[x] = b->refs[order[x].q];
      b->work[x] = work[order[x].q];
      for (int c = 0; c < 3; ++c) ijk[3 * x + c] = b->task_ijk[3 * order[x].q + c];
    }
    if (order_kind == BBTC_BATCH_SMALL_GRAPHS_FIRST && kernel == BBTC_KERNEL_AUTO && !refs.empty()) {
      // each graph's tasks (a contiguous block) in the order a one-graph
      // plan runs them: each form grouped by bc tile, the forms interleaved
      const bbtc::PlanDesc D = bbtc::describe_plan(A->htiles, A->deg, refs, kernel, nullptr, true, &b->work);
      std::vector<bbtc::TaskRef> r2(refs.size());
      std::vector<bbtc::TaskWork> w2(refs.size());
      std::vector<uint32_t> ijk2(ijk.size());
      size_t lo = 0;
      while (lo < refs.size()) {
        size_t hi = lo + 1;
        while (hi < refs.size() && refs[hi].graph == refs[lo].graph) ++hi;
        const size_t len = hi - lo;
        bbtc::PlanDesc G;
        G.td.assign(D.td.begin() + long(lo), D.td.begin() + long(hi));
        G.hkind.assign(D.hkind.begin() + long(lo), D.hkind.begin() + long(hi));
        G.pref.resize(len + 1);
        for (size_t x = 0; x <= len; ++x) G.pref[x] = D.pref[lo + x] - D.pref[lo];
        const std::vector<bbtc::TaskRef> gr(refs.begin() + long(lo), refs.begin() + long(hi));
        const std::vector<bbtc::TaskWork> gw(b->work.begin() + long(lo), b->work.begin() + long(hi));
        const std::vector<uint64_t> ord = bbtc::mixed_order(A->htiles, G, gw, gr, len, len);
        for (size_t x = 0; x < len; ++x) {
          r2[lo + x] = refs[lo + ord[x]];
          w2[lo + x] = b->work[lo + ord[x]];
          for (int c = 0; c < 3; ++c) ijk2[3 * (lo + x) + c] = ijk[3 * (lo + ord[x]) + c];
        }
        lo = hi;
      }
      refs.swap(r2);
      b->work.swap(w2);
      ijk.swap(ijk2);
    }
    b->refs.swap(refs);
    b->task_ijk.swap(ijk);
    const uint64_t ng = std::max<uint64_t>(ngraphs, 1);
    b->left0 = bbtc::dev_alloc<unsigned long long>(ctx, ng, &b->bytes);
    b->left = bbtc::dev_alloc<unsigned long long>(ctx, ng, &b->bytes);
    b->done = bbtc::dev_alloc<unsigned long long>(ctx, ng, &b->bytes);
    b->start = bbtc::dev_alloc<unsigned long long>(ctx, 1, &b->bytes);
    bbtc::batch_replan(b, kernel, nullptr);
  } catch (...) {
    // copies straight from pinned layouts may still be in flight
    (void)hipStreamSynchronize(ctx->stream);
    bbtc::batch_release(b);
    throw;
  }
  *out = b;
  DEV_CATCH
}

uint64_t bbtc_batch_ntasks(const bbtc_batch* b) { return b ? b->refs.size() : 0; }
uint64_t bbtc_batch_bytes(const bbtc_batch* b) {
  return b ? b->bytes + (b->arena ? b->arena->bytes : 0) + (b->plan ? b->plan->bytes : 0) : 0;
}

bbtc_status bbtc_batch_tasks(const bbtc_batch* b, uint32_t* graph, bbtc_task* tasks) {
  DEV_REQUIRE(b, "null argument");
  for (uint64_t q = 0; q < b->refs.size(); ++q) {
    if (graph) graph[q] = b->refs[q].graph;
    if (tasks) tasks[q] = bbtc_task{b->task_ijk[3 * q], b->task_ijk[3 * q + 1], b->task_ijk[3 * q + 2]};
  }
  return BBTC_OK;
}

bbtc_status bbtc_batch_task_kinds(const bbtc_batch* b, uint8_t* kinds) {
  DEV_REQUIRE(b && (kinds || b->refs.empty()), "null argument");
  for (uint64_t q = 0; q < b->refs.size(); ++q) kinds[q] = uint8_t(b->plan->hkind[q]);
  return BBTC_OK;
}

bbtc_status bbtc_batch_autotune(bbtc_batch* b, uint8_t* choice_out) {
  DEV_REQUIRE(b, "null argument");
  DEV_TRY
  HIP_OK(hipSetDevice(b->ctx->device));
  std::vector<uint8_t> choice(b->refs.size());
  bbtc::autotune_refs(b->ctx, b->arena, b->refs, choice.data(), nullptr);
  bbtc::batch_replan(b, BBTC_KERNEL_AUTO, choice.data());
  if (choice_out) std::copy(choice.begin(), choice.end(), choice_out);
  DEV_CATCH
}

bbtc_status bbtc_batch_set_forms(bbtc_batch* b, bbtc_kernel kernel, const uint8_t* choice) {
  DEV_REQUIRE(b && bbtc::valid_kernel(kernel), "bad argument");
  DEV_TRY
  HIP_OK(hipSetDevice(b->ctx->device));
  bbtc::batch_replan(b, kernel, choice);
  DEV_CATCH
}

bbtc_status bbtc_batch_capture(bbtc_batch* b) {
  DEV_REQUIRE(b, "null argument");
  DEV_TRY
  HIP_OK(hipSetDevice(b->ctx->device));
  bbtc::batch_drop_capture(b);
  hipStream_t s = b->ctx->stream;
  HIP_OK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  try {
    bbtc::batch_enqueue(b, s);
  } catch (...) {
    hipGraph_t g = nullptr;
    (void)hipStreamEndCapture(s, &g);
    if (g) (void)hipGraphDestroy(g);
    throw;
  }
  HIP_OK(hipStreamEndCapture(s, &b->graph));
  HIP_OK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  DEV_CATCH
}

bbtc_status bbtc_batch_launch(bbtc_batch* b) {
  DEV_REQUIRE(b, "null argument");
  DEV_TRY
  HIP_OK(hipSetDevice(b->ctx->device));
  if (b->exec) HIP_OK(hipGraphLaunch(b->exec, b->ctx->stream));
  else bbtc::batch_enqueue(b, b->ctx->stream);
  DEV_CATCH
}

bbtc_status bbtc_batch_results(bbtc_batch* b, uint64_t* per_graph, uint64_t* per_task,
                               double* latency_ms) {
  DEV_REQUIRE(b, "null argument");
  DEV_TRY
  HIP_OK(hipSetDevice(b->ctx->device));
  hipStream_t s = b->ctx->stream;
  const uint64_t n = b->refs.size(), ng = b->ngraphs;
  std::vector<uint64_t> pt(std::max<uint64_t>(n, 1), 0);
  std::vector<unsigned long long> done(std::max<uint64_t>(ng, 1), 0);
  unsigned long long start = 0;
  if (n) HIP_OK(hipMemcpyAsync(pt.data(), b->plan->per_task, n * 8, hipMemcpyDeviceToHost, s));
  if (ng) HIP_OK(hipMemcpyAsync(done.data(), b->done, ng * 8, hipMemcpyDeviceToHost, s));
  HIP_OK(hipMemcpyAsync(&start, b->start, 8, hipMemcpyDeviceToHost, s));
  HIP_OK(hipStreamSynchronize(s));
  if (per_graph) {
    for (uint64_t gi = 0; gi < ng; ++gi) per_graph[gi] = 0;
    for (uint64_t q = 0; q < n; ++q) per_graph[b->refs[q].graph] += pt[q];
  }
  if (per_task) std::copy(pt.begin(), pt.begin() + n, per_task);
  if (latency_ms)
    for (uint64_t gi = 0; gi < ng; ++gi)
      latency_ms[gi] = done[gi] >= start ? double(done[gi] - start) / b->wall_khz : 0.0;
  DEV_CATCH
}

void bbtc_batch_destroy(bbtc_batch* b) {
  if (!b) return;
  (void)hipSetDevice(b->ctx->device);
  (void)hipStreamSynchronize(b->ctx->stream);
  bbtc::batch_release(b);
}

}  // extern "C"
