// Internals shared by the device translation units (device.hip, server.hip):
// device allocation, the host description of a plan, and the counting
// launch sequence. The kernels themselves are compiled in device.hip only.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "form_rule.hpp"
#include "objects.hpp"

namespace bbtc {

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

inline void dev_free(bbtc_ctx* ctx, void* p) {
  if (p) (void)hipFree(p);
  (void)ctx;
}

bool valid_kernel(bbtc_kernel k);

// A task named by its three tile ids in a graph's tile table, and the graph
// it belongs to (0 for a single graph, the graph's index in a batch).
struct TaskRef {
  uint32_t ab, bc, ac, graph;
};

std::vector<TaskRef> refs_of(uint32_t P, const bbtc_task* tasks, uint64_t n);
// The AUTO rule's form for a task of work statistics s (ac(u) bitmaps of wpr
// words per row, fit rows of them in a wave's LDS slice, a bitmap form of the
// bc tile or not) and, in *cost, its estimated shader clocks (0 without
// wedges).
uint32_t auto_kind(const TaskWork& s, uint32_t wpr, uint32_t fit, bool has_dense, double* cost);
// The row form: its launch (the claim counter reset, then count_rows over
// every row), and its table of N+(v) in S-word slots built from a block CSR
// on the device (band starts cuts[0..P]).
void enqueue_rows(bbtc_ctx* ctx, const RowArgs& a, hipStream_t s);
void enqueue_build_rows(const uint32_t* rowptr, const uint32_t* cols, const TileDesc* tiles, const uint32_t* cuts,
                        uint32_t P, uint32_t nrows, uint32_t S, uint32_t* rows, uint32_t* rowlen, hipStream_t s);
// Whether BBTC_KERNEL_AUTO takes the row form for a graph of P bands with
// these tiles and tile degrees whose tasks all admit the lane form.
bool rows_pay(uint32_t P, const std::vector<TileDesc>& tiles, const std::vector<float>& deg);

// Device tile table of a layout (P*P entries, offsets into its arrays).
std::vector<TileDesc> tile_table(const Layout& L);

// Host side of a plan: the form and rows per item of each task (kernel rule,
// or choice[q] = 1 merge, 2 bitmap, 3 dense, 4 rdense, 5 lane) and the item
// prefix.
// Which tasks make_plan may list in an order of its own (see mixed_order):
// none (on false), the first `upto`, or with `contested` the first
// max(upto, first task a whole-plan launch's last claim part holds).
struct MixSpec {
  bool on = false;
  uint64_t upto = 0;
  bool contested = false;
};

// Per-task values of plan p in its own order -> the caller's order.
template <class T>
void to_caller_order(const bbtc_plan* p, const T* in, T* out) {
  if (p->order.empty()) {
    std::copy(in, in + p->ntasks, out);
    return;
  }
  for (uint64_t q = 0; q < p->ntasks; ++q) out[p->order[q]] = in[q];
}

struct PlanDesc {
  std::vector<TaskDesc> td;       // max(n, 1) entries
  std::vector<uint64_t> pref;     // n + 1 entries
  std::vector<uint32_t> hkind;    // form per task, 1-based as in choice
  uint64_t bitmap_tasks = 0;
  bool lane_only = true;          // every task with work items uses the lane form
};

// The order in which a plan runs tasks [0, k) of D (refs, work: the tasks in
// D's order; tasks [k, n) keep their places): each form's tasks grouped by
// bc tile, the tiles of most estimated work first, and the forms
// interleaved by estimated work. Entry q of the result is D's task order[q].
std::vector<uint64_t> mixed_order(const std::vector<TileDesc>& tiles, const PlanDesc& D,
                                  const std::vector<TaskWork>& work, const std::vector<TaskRef>& refs, uint64_t k,
                                  uint64_t n);

// auto_lanes: BBTC_KERNEL_AUTO may take the lane form for every task.
// work (may be NULL): each task's work statistics, aligned with refs; with
// them BBTC_KERNEL_AUTO takes the form of least estimated cost
// (form_rule.hpp), without them the rule that needs no statistics.
PlanDesc describe_plan(const std::vector<TileDesc>& tiles, const std::vector<float>& deg,
                       const std::vector<TaskRef>& refs, bbtc_kernel kernel, const uint8_t* choice,
                       bool auto_lanes = true, const std::vector<TaskWork>* work = nullptr);

// Work statistics (estimated, form_rule.hpp) of tasks of a host layout.
std::vector<TaskWork> task_work(const Layout& L, const bbtc_task* tasks, uint64_t n);
// The same from task_stats output (kTaskStats numbers per task).
std::vector<TaskWork> work_from_stats(const uint64_t* st, uint64_t n);

// Resident blocks of count_probe (all forms, or lane form only) on the
// context's device.
inline uint64_t resident_blocks(const bbtc_ctx* ctx, bool lane_only = false) {
  return std::max<uint64_t>(1, uint64_t(ctx->num_cu) *
                                   uint64_t(lane_only ? ctx->blocks_per_cu_lane : ctx->blocks_per_cu));
}

// Enqueues one counting launch on s: reset of the slot (first item i0), the
// first phase over at most `blocks` blocks, the second over the whole
// machine. lane_only: every item of the launch is of the lane form, which
// runs in its own kernel instantiation (no LDS, more resident waves) and
// never spills, so there is no second phase.
// chunk: items per chunk of the first-phase parts (0: one chunk per part).
// defer_steps: wave steps after which a first-phase unit spills (0: the rule
// of one-graph launches, kDeferStepsMany for large launches with one chunk
// per part, else kDeferSteps).
void enqueue_count(bbtc_ctx* ctx, CountArgs a, unsigned long long i0, uint64_t items, bool lane_only,
                   uint64_t chunk, hipStream_t s, uint32_t defer_steps = 0);

// Host to device copy on stream s. Copies of at least kStageMinBytes go
// through the context's two pinned staging buffers: the host threads fill
// one buffer while the other's chunk is in flight over PCIe (pinned copies
// run at about twice the rate of pageable ones on MI355X hosts), so src may
// be reused on return; smaller copies are plain hipMemcpyAsync calls. Calls
// on one context are serialised on its staging buffers. With src_pinned (src
// is page-locked memory, as after bbtc_layout_pin) the copy is one
// hipMemcpyAsync straight from src, which must then stay valid until the
// stream has passed it.
constexpr uint64_t kStageMinBytes = 8ull << 20;
void copy_to_device(bbtc_ctx* ctx, hipStream_t s, void* dst, const void* src, uint64_t bytes,
                    bool src_pinned = false);
// Moves a layout's block-CSR arrays into page-locked host memory (no-op for
// arrays already there).
void pin_layout(Layout& L);

// Host-to-device copies below this size ran as blit kernels on the compute
// units (a copy trace on MI355X: 4 KiB copies did, 64 KiB and larger ran on
// the SDMA engines), and such a kernel waits for a free CU slot behind a
// persistent count launch that holds them all, stalling the copies queued
// after it on its stream.
constexpr uint64_t kSdmaMinBytes = 64ull << 10;
constexpr uint64_t kStageImageBytes = 64ull << 20;  // one staging buffer
struct CopyPiece {
  uint64_t offset;  // from the image's start
  const void* src;
  uint64_t bytes;
};
// Packs the pieces into one staging buffer at their offsets and copies its
// first copy_bytes (at least every piece's end, at most kStageImageBytes)
// to dst in one hipMemcpyAsync on s: many small copies become one that the
// SDMA engines take when copy_bytes >= kSdmaMinBytes. The bytes between the
// pieces are undefined on the device. The sources may be reused on return.
void copy_image_to_device(bbtc_ctx* ctx, hipStream_t s, void* dst, const CopyPiece* pieces, size_t npieces,
                          uint64_t copy_bytes);
// Frees the context's staging buffers (after its streams are idle).
void staging_release(bbtc_ctx* ctx);

// Enqueues p[0..n) = 0 on s.
void enqueue_zero(unsigned long long* p, uint64_t n, hipStream_t s);

// Bitmap-row jobs of a layout's dense tiles, with its arrays placed at rp0,
// col0 and dense0 of the device arrays; rows_out: the bitmap rows in all.
std::vector<DenseJob> dense_jobs(const Layout& L, uint64_t rp0, uint64_t col0, uint64_t dense0,
                                 uint64_t row0, uint64_t* rows_out);
// Enqueues on s: dense[0..dense_words) = 0, then the bitmap rows of the jobs
// (jobs_dev holds njobs jobs with rows rows in all) from the uploaded tiles.
void enqueue_dense_build(const uint32_t* rowptr, const uint32_t* cols, uint32_t* dense, uint64_t dense_words,
                         const DenseJob* jobs_dev, uint64_t njobs, uint64_t rows, hipStream_t s);

}  // namespace bbtc
