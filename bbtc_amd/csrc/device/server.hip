// Serving: count requests for many graphs on one device with continuous
// batching.
//
// A request carries a host block CSR and its task list. It arrives at a
// given time and is admitted, in arrival order, once its device footprint
// fits the server's heap. The footprint is the block CSR (row offsets,
// columns, non-empty rows, bitmap rows of dense tiles) plus the plan's task
// table and counters. An upload thread copies admitted graphs into the heap
// on a copy stream while the device counts earlier requests.
//
// The scheduler thread keeps up to `max_inflight` launches queued on the
// compute stream. By default a launch's budget is estimated work: each
// request's items carry the reference's task weight (tasks.hpp:21-23) spread
// evenly over the task's items, and the budget follows the measured work
// rate of finished launches so that a launch takes about `launch_us` (a
// launch's fixed costs stay small next to its work, and an arrival waits at
// most about two launches); with `quantum_items` set, the budget is that
// many items. The oldest resident request with work left keeps `oldest_pct`
// percent of the budget, and the rest is filled with slices of the resident
// requests in order of remaining work (shortest first). A launch's items are
// split into one contiguous range per claim part, as for one graph (each XCD
// on its own range; `list_chunks` deals 4,096-item chunks instead). A request that
// arrives while others run joins the next launch and shares the machine
// with them; a large request is slowed by later arrivals by at most about
// 100 / oldest_pct. With `lane_form`, requests whose tasks all take the lane
// form run in launches of their own (the lane kernel), the two kinds taking
// turns while both have work. A request that the plain path would count in
// the row form gets its row table built after its upload and one count_rows
// launch of its own, taking turns with the shared launches. A request
// finishes with the launch that holds its last slice. Its counters are read
// back and its heap block returns to the heap, which lets the next arrival
// in.
//
// This replaces the reference's per-run task loop (bb_tc.cpp:142-181,
// 200-252), which counts one graph per process.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <condition_variable>
#include <deque>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "internal.hpp"

namespace bbtc {
void set_last_error(const std::string& s);

namespace {

using Clock = std::chrono::steady_clock;

constexpr uint64_t kAlign = 256;
constexpr uint64_t kDefaultHeap = 200000000000ull;  // 200 GB
constexpr uint64_t kDefaultQuantum = 1ull << 18;
// adaptive launch budget: limits, and the default target launch time
constexpr uint64_t kMinQuantum = 1ull << 14;
constexpr uint64_t kMaxQuantum = 1ull << 24;
constexpr double kDefaultLaunchMs = 40.0;
constexpr uint32_t kDefaultOldestPct = 90;
constexpr uint32_t kDefaultInflight = 2;
constexpr uint32_t kDefaultSegments = 1024;

inline uint64_t align_up(uint64_t x) { return (x + kAlign - 1) / kAlign * kAlign; }

// First-fit allocator over one device range, with coalescing frees.
class Heap {
 public:
  void init(uint64_t size) {
    free_.clear();
    if (size) free_[0] = size;
    size_ = size;
  }
  // Offset of a block of `bytes`, or ~0 when no free range is large enough.
  uint64_t alloc(uint64_t bytes) {
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < bytes) continue;
      const uint64_t off = it->first, rest = it->second - bytes;
      free_.erase(it);
      if (rest) free_[off + bytes] = rest;
      used_ += bytes;
      return off;
    }
    return ~0ull;
  }
  void release(uint64_t off, uint64_t bytes) {
    used_ -= bytes;
    auto next = free_.lower_bound(off);
    if (next != free_.end() && off + bytes == next->first) {
      bytes += next->second;
      next = free_.erase(next);
    }
    if (next != free_.begin()) {
      auto prev = std::prev(next);
      if (prev->first + prev->second == off) {
        prev->second += bytes;
        return;
      }
    }
    free_[off] = bytes;
  }
  uint64_t used() const { return used_; }
  uint64_t size() const { return size_; }

 private:
  std::map<uint64_t, uint64_t> free_;
  uint64_t size_ = 0, used_ = 0;
};

struct Request {
  uint64_t id = 0;
  const Layout* L = nullptr;
  std::vector<TileDesc> tiles;
  PlanDesc plan;
  uint64_t ntasks = 0, nitems = 0;
  uint64_t* per_task_out = nullptr;
  double arrival = 0;
  // placement of the arrays inside the request's heap block
  uint64_t o_rp = 0, o_col = 0, o_nz = 0, o_dense = 0, o_tiles = 0, o_tasks = 0, o_pref = 0, o_cnt = 0, o_jobs = 0;
  // the block's first image_end bytes (the tables, and the graph arrays when
  // image_graph) are copied in as one packed image (0: piece by piece)
  uint64_t image_end = 0;
  bool image_graph = false;
  // The row form, where the plain path's AUTO takes it: one count_rows
  // launch over the whole request, with the table of N+(v) in rows_S-word
  // slots built on the build stream after the upload; rcuts are the band
  // starts, lex2plan maps a lexicographic task to the request's task.
  bool rows = false;
  uint32_t rows_S = 0, nrows = 0;
  std::vector<uint32_t> rcuts, lex2plan;
  uint64_t o_rows = 0, o_rowlen = 0, o_rcuts = 0, o_lex2plan = 0;
  RowArgs ra{};
  // the plan's task q is the caller's task order[q] (empty: the caller's order)
  std::vector<uint64_t> order;
  std::vector<DenseJob> jobs;  // bitmap rows built on the device after the upload
  uint64_t job_rows = 0;
  uint64_t bytes = 0;
  uint64_t heap_off = ~0ull;
  hipEvent_t copied = nullptr;    // its copies have landed (copy stream)
  hipEvent_t uploaded = nullptr;  // and its bitmap rows are built (build stream)
  SegDesc seg{};
  uint64_t next = 0;  // next item to put in a launch
  // estimated work of items [0, x): wpref[t] before task t (the reference's
  // task weight, tasks.hpp:21-23, spread evenly over the task's items)
  std::vector<double> wpref;
  uint32_t launches = 0;
  double admit_s = 0, ready_s = 0, start_s = -1, done_s = 0;
  bbtc_status status = BBTC_OK;
  uint64_t total = 0;
};

// Estimated work of items [0, x) of r.
double work_at(const Request& r, uint64_t x) {
  if (x >= r.nitems) return r.wpref.back();
  const std::vector<uint64_t>& p = r.plan.pref;
  const size_t t = size_t(std::upper_bound(p.begin(), p.end(), x) - p.begin()) - 1;
  const uint64_t n = p[t + 1] - p[t];
  return r.wpref[t] + (r.wpref[t + 1] - r.wpref[t]) * double(x - p[t]) / double(n);
}

// Items of r from `from` that hold about `work` (at least one item).
uint64_t items_for_work(const Request& r, uint64_t from, double work) {
  const double target = work_at(r, from) + work;
  if (target >= r.wpref.back()) return r.nitems - from;
  const std::vector<uint64_t>& p = r.plan.pref;
  // task t with wpref[t] <= target < wpref[t + 1] (then it has items)
  const size_t t = size_t(std::upper_bound(r.wpref.begin(), r.wpref.end(), target) - r.wpref.begin()) - 1;
  const double per = (r.wpref[t + 1] - r.wpref[t]) / double(p[t + 1] - p[t]);
  uint64_t x = p[t] + uint64_t(std::ceil((target - r.wpref[t]) / per));
  x = std::min(std::max(x, from + 1), r.nitems);
  return x - from;
}

// Waiting requests of equal arrival time compared for the admission order.
constexpr int kAdmitScan = 1024;

struct Launch {
  hipEvent_t start = nullptr, done = nullptr;  // around the launch on the compute stream
  uint64_t items = 0;
  double work = 0;  // estimated work of its items
  int kind = 0;  // 1: lane-form requests only; 2: one request in the row form
  int ring = 0;
  std::vector<std::shared_ptr<Request>> finishing;
};

}  // namespace
}  // namespace bbtc

struct bbtc_server {
  bbtc_ctx* ctx = nullptr;
  uint64_t quantum[2] = {0, 0};  // item budget of the next launch of each kind (lane form: [1])
  bool adaptive = false;         // budgets in estimated work, following the measured rates
  double work_budget[2] = {0, 0};  // adaptive: work per launch of each kind (0: not yet measured)
  double launch_ms = 0;      // target launch time when adaptive
  uint32_t max_inflight = 0, max_segments = 0;
  bool sequential = false;
  bool lane_turn = false;  // kind of the next launch when both kinds have work
  bool rows_turn = false;  // a row-form request's launch goes next (shared launches had theirs)
  bool lane_form = false;  // AUTO requests may take the lane form
  uint32_t oldest_pct = 90;  // share of each launch's budget kept by the oldest request
  bool list_chunks = false;  // launches dealt to the claim parts in 4,096-item chunks
  uint32_t defer_steps = 0;  // split threshold of launches (0: the one-graph rule)

  char* heap_base = nullptr;
  bbtc::Heap heap;
  // copies in (copy), the bitmap rows of uploaded requests (build, so the
  // next request's copies need not wait behind those kernels), counting
  // launches (compute), counters out (readback)
  hipStream_t copy = nullptr, build = nullptr, compute = nullptr, readback = nullptr;
  bbtc::SlotCtl* ctl = nullptr;
  bbtc::SpillRec* spill = nullptr;
  uint64_t* spill_base = nullptr;
  std::vector<bbtc::SegDesc*> seg_dev;   // per ring slot
  std::vector<bbtc::SegDesc*> seg_host;  // pinned, per ring slot
  std::vector<hipEvent_t> event_pool;

  std::mutex mu;
  std::condition_variable cv_admit, cv_sched, cv_done;
  std::deque<std::shared_ptr<bbtc::Request>> pending;  // arrival order
  std::deque<std::shared_ptr<bbtc::Request>> ready;    // uploading or uploaded
  std::deque<bbtc_server_result> results;
  uint64_t next_id = 1;
  uint64_t submitted = 0, finished = 0;
  uint64_t active = 0;  // admitted and not finished
  bool stop = false;
  bbtc_status fatal = BBTC_OK;
  std::string fatal_msg;
  bbtc_server_stats stats{};
  bbtc::Clock::time_point epoch;

  std::thread admit_thread, sched_thread;

  double now() const { return std::chrono::duration<double>(bbtc::Clock::now() - epoch).count(); }
};

namespace bbtc {
namespace {

void set_fatal(bbtc_server* S, bbtc_status st, const std::string& msg) {
  std::lock_guard<std::mutex> lk(S->mu);
  if (S->fatal == BBTC_OK) {
    S->fatal = st;
    S->fatal_msg = msg;
  }
  S->stop = true;
  S->cv_admit.notify_all();
  S->cv_sched.notify_all();
  S->cv_done.notify_all();
}

// Places the request's arrays in one block and returns its size: the small
// tables first, then the graph arrays, the counters and the bitmap rows. The
// tables, with the graph arrays of a graph of at most kStageMinBytes, are
// copied in as one packed image of at least kSdmaMinBytes, and the counters
// are read back kSdmaMinBytes at least, so that neither becomes a blit
// kernel queued behind a count launch; the block is large enough for both.
uint64_t place(Request& r) {
  const Layout& L = *r.L;
  uint64_t o = 0;
  auto put = [&](uint64_t& at, uint64_t bytes) {
    at = o;
    o += align_up(std::max<uint64_t>(bytes, 1));
  };
  put(r.o_tiles, r.tiles.size() * sizeof(TileDesc));
  put(r.o_tasks, r.plan.td.size() * sizeof(TaskDesc));
  put(r.o_pref, r.plan.pref.size() * 8);
  put(r.o_jobs, r.jobs.size() * sizeof(DenseJob));
  put(r.o_rcuts, r.rcuts.size() * 4);
  put(r.o_lex2plan, r.lex2plan.size() * 4);
  const uint64_t tables_end = o;
  put(r.o_rp, L.rowptr.size() * 4);
  put(r.o_col, L.cols.size() * 4);
  put(r.o_nz, L.nzrows.size() * 4);
  const uint64_t graph_bytes = (L.rowptr.size() + L.cols.size() + L.nzrows.size()) * 4;
  r.image_graph = graph_bytes <= kStageMinBytes && o <= kStageImageBytes;
  r.image_end = r.image_graph ? o : tables_end;
  if (r.image_end > kStageImageBytes) r.image_end = 0;  // tables beyond one staging buffer
  put(r.o_cnt, std::max<uint64_t>(r.ntasks, 1) * 8);
  put(r.o_dense, L.dense_words * 4);
  put(r.o_rows, r.rows ? uint64_t(r.nrows) * r.rows_S * 4 : 0);
  put(r.o_rowlen, r.rows ? uint64_t(r.nrows) * 4 : 0);
  return std::max(o, r.o_cnt + kSdmaMinBytes);
}

// Sets up the row form for request r of layout L where the plain path would
// take it (BBTC_KERNEL_ROWS, or AUTO where rows_pay holds): every task
// admits the lane form, no row is longer than kRowSlotMax, no task is listed
// twice, and the table fits kRowTableMaxBytes.
void plan_request_rows(Request& r, const Layout& L, const std::vector<TaskRef>& refs, bbtc_kernel kernel,
                       const std::function<bool()>& lane_only) {
  const uint32_t P = L.P;
  if (refs.empty() || P == 0 || P > kRowMaxCuts || !L.longest_known || L.longest_row > kRowSlotMax) return;
  if (!(kernel == BBTC_KERNEL_ROWS || (kernel == BBTC_KERNEL_AUTO && rows_pay(P, r.tiles, L.deg)))) return;
  if (!lane_only()) return;
  std::vector<uint32_t> cuts(size_t(P) + 1, 0);
  for (uint32_t b = 0; b < P; ++b) {
    const uint64_t e = uint64_t(cuts[b]) + r.tiles[size_t(b) * P + b].h;
    if (e >= 0xFFFFFFFFull) return;  // ids must stay below kRowPad
    cuts[b + 1] = uint32_t(e);
  }
  uint32_t S = 4;
  while (S < L.longest_row && S < kRowSlotMax) S <<= 1;
  if (cuts[P] == 0 || uint64_t(cuts[P]) * S * 4 > kRowTableMaxBytes) return;
  const uint64_t ntask = lex_task(P, P - 1, P - 1, P - 1) + 1;
  if (ntask >= 0xFFFFFFFFull) return;
  std::vector<uint32_t> lex2plan(ntask, ~0u);
  for (size_t q = 0; q < refs.size(); ++q) {
    const uint32_t i = refs[q].ab / P, j = refs[q].ab % P, k = refs[q].bc % P;
    uint32_t& slot = lex2plan[lex_task(P, i, j, k)];
    if (slot != ~0u) return;  // listed twice
    slot = uint32_t(q);
  }
  r.rows = true;
  r.rows_S = S;
  r.nrows = cuts[P];
  r.rcuts = std::move(cuts);
  r.lex2plan = std::move(lex2plan);
}

void push_result(bbtc_server* S, const Request& r) {
  bbtc_server_result x{};
  x.id = r.id;
  x.total = r.total;
  x.status = r.status;
  x.launches = r.launches;
  x.bytes = r.bytes;
  x.arrival_s = r.arrival;
  x.admit_s = r.admit_s;
  x.ready_s = r.ready_s;
  x.start_s = r.start_s < 0 ? r.ready_s : r.start_s;
  x.done_s = r.done_s;
  S->results.push_back(x);
  S->finished++;
}

// Upload thread: admits arrived requests in arrival order while their
// footprint fits the heap, and copies each into its block.
void admit_loop(bbtc_server* S) {
  try {
    HIP_OK(hipSetDevice(S->ctx->device));
    for (;;) {
      std::shared_ptr<Request> r;
      {
        std::unique_lock<std::mutex> lk(S->mu);
        for (;;) {
          if (S->stop && S->pending.empty()) return;
          if (S->stop) {  // shutting down: cancel what was not admitted
            while (!S->pending.empty()) {
              auto q = S->pending.front();
              S->pending.pop_front();
              q->status = BBTC_ERR_INVALID;
              q->done_s = S->now();
              push_result(S, *q);
            }
            S->cv_done.notify_all();
            return;
          }
          if (S->pending.empty()) {
            S->cv_admit.wait(lk);
            continue;
          }
          const double t = S->now();
          if (S->pending.front()->arrival > t) {
            S->cv_admit.wait_for(lk, std::chrono::duration<double>(S->pending.front()->arrival - t));
            continue;
          }
          if (S->sequential && S->active > 0) {  // plain mode: one request at a time
            S->cv_admit.wait(lk);
            continue;
          }
          // Requests keep their arrival order, but of those that arrived at
          // the same time as the first one waiting, the one of most
          // estimated work goes first (for a batch submitted at once:
          // longest first, so the largest counts overlap the uploads of the
          // rest instead of running alone at the end).
          auto pick = S->pending.begin();
          if (!S->sequential) {
            int scanned = 0;
            for (auto it = std::next(pick); it != S->pending.end() && scanned < kAdmitScan; ++it, ++scanned) {
              if ((*it)->arrival != S->pending.front()->arrival) break;
              if ((*it)->wpref.back() > (*pick)->wpref.back()) pick = it;
            }
          }
          const uint64_t off = S->heap.alloc((*pick)->bytes);
          if (off == ~0ull) {  // wait for a finished request to free memory
            S->cv_admit.wait(lk);
            continue;
          }
          r = *pick;
          S->pending.erase(pick);
          r->heap_off = off;
          r->admit_s = t;
          S->active++;
          S->stats.peak_resident_bytes = std::max<uint64_t>(S->stats.peak_resident_bytes, S->heap.used());
          S->stats.peak_active = std::max<uint32_t>(S->stats.peak_active, uint32_t(S->active));
          break;
        }
      }
      // copy the block CSR and the plan into the block
      const auto c0 = Clock::now();
      const Layout& L = *r->L;
      char* b = S->heap_base + r->heap_off;
      hipStream_t s = S->copy;
      const uint64_t tile_b = r->tiles.size() * sizeof(TileDesc), task_b = r->plan.td.size() * sizeof(TaskDesc);
      const uint64_t pref_b = r->plan.pref.size() * 8, job_b = r->jobs.size() * sizeof(DenseJob);
      if (r->image_end) {
        // the image first: its padding may cover the graph arrays' places,
        // which the copies after it then fill
        std::vector<CopyPiece> img = {{r->o_tiles, r->tiles.data(), tile_b},
                                      {r->o_tasks, r->plan.td.data(), task_b},
                                      {r->o_pref, r->plan.pref.data(), pref_b},
                                      {r->o_jobs, r->jobs.data(), job_b},
                                      {r->o_rcuts, r->rcuts.data(), r->rcuts.size() * 4},
                                      {r->o_lex2plan, r->lex2plan.data(), r->lex2plan.size() * 4}};
        if (r->image_graph) {
          img.push_back({r->o_rp, L.rowptr.data(), L.rowptr.size() * 4});
          img.push_back({r->o_col, L.cols.data(), L.cols.size() * 4});
          img.push_back({r->o_nz, L.nzrows.data(), L.nzrows.size() * 4});
        }
        copy_image_to_device(S->ctx, s, b, img.data(), img.size(), std::max(r->image_end, kSdmaMinBytes));
      } else {
        HIP_OK(hipMemcpyAsync(b + r->o_tiles, r->tiles.data(), tile_b, hipMemcpyHostToDevice, s));
        HIP_OK(hipMemcpyAsync(b + r->o_tasks, r->plan.td.data(), task_b, hipMemcpyHostToDevice, s));
        HIP_OK(hipMemcpyAsync(b + r->o_pref, r->plan.pref.data(), pref_b, hipMemcpyHostToDevice, s));
        if (job_b) HIP_OK(hipMemcpyAsync(b + r->o_jobs, r->jobs.data(), job_b, hipMemcpyHostToDevice, s));
        if (r->rows) {
          HIP_OK(hipMemcpyAsync(b + r->o_rcuts, r->rcuts.data(), r->rcuts.size() * 4, hipMemcpyHostToDevice, s));
          HIP_OK(hipMemcpyAsync(b + r->o_lex2plan, r->lex2plan.data(), r->lex2plan.size() * 4,
                                hipMemcpyHostToDevice, s));
        }
      }
      if (!r->image_graph) {
        copy_to_device(S->ctx, s, b + r->o_rp, L.rowptr.data(), L.rowptr.size() * 4, L.rowptr.pinned());
        copy_to_device(S->ctx, s, b + r->o_col, L.cols.data(), L.cols.size() * 4, L.cols.pinned());
        copy_to_device(S->ctx, s, b + r->o_nz, L.nzrows.data(), L.nzrows.size() * 4, L.nzrows.pinned());
      }
      auto* cnt = reinterpret_cast<unsigned long long*>(b + r->o_cnt);
      auto* jd = reinterpret_cast<DenseJob*>(b + r->o_jobs);
      HIP_OK(hipEventCreateWithFlags(&r->copied, hipEventDisableTiming));
      HIP_OK(hipEventRecord(r->copied, s));
      // the counters and the bitmap rows of the dense tiles, built from the
      // uploaded tiles on the build stream
      hipStream_t bs = S->build;
      HIP_OK(hipStreamWaitEvent(bs, r->copied, 0));
      enqueue_zero(cnt, r->ntasks, bs);
      enqueue_dense_build(reinterpret_cast<const uint32_t*>(b + r->o_rp), reinterpret_cast<const uint32_t*>(b + r->o_col),
                          reinterpret_cast<uint32_t*>(b + r->o_dense), L.dense_words, jd, r->jobs.size(), r->job_rows, bs);
      if (r->rows) {
        RowArgs& ra = r->ra;
        ra.rows = reinterpret_cast<const uint32_t*>(b + r->o_rows);
        ra.rowlen = reinterpret_cast<const uint32_t*>(b + r->o_rowlen);
        ra.cuts = reinterpret_cast<const uint32_t*>(b + r->o_rcuts);
        ra.lex2plan = reinterpret_cast<const uint32_t*>(b + r->o_lex2plan);
        ra.per_task = cnt;
        ra.nrows = r->nrows;
        ra.P = L.P;
        ra.s_shift = uint32_t(__builtin_ctz(r->rows_S));
        enqueue_build_rows(reinterpret_cast<const uint32_t*>(b + r->o_rp), reinterpret_cast<const uint32_t*>(b + r->o_col),
                           reinterpret_cast<const TileDesc*>(b + r->o_tiles), ra.cuts, L.P, r->nrows, r->rows_S,
                           reinterpret_cast<uint32_t*>(b + r->o_rows), reinterpret_cast<uint32_t*>(b + r->o_rowlen), bs);
      }
      SegDesc& g = r->seg;
      g.rowptr = reinterpret_cast<const uint32_t*>(b + r->o_rp);
      g.cols = reinterpret_cast<const uint32_t*>(b + r->o_col);
      g.nzrows = reinterpret_cast<const uint32_t*>(b + r->o_nz);
      g.dense = reinterpret_cast<const uint32_t*>(b + r->o_dense);
      g.tiles = reinterpret_cast<const TileDesc*>(b + r->o_tiles);
      g.tasks = reinterpret_cast<const TaskDesc*>(b + r->o_tasks);
      g.item_prefix = reinterpret_cast<const uint64_t*>(b + r->o_pref);
      g.per_task = cnt;
      g.cycles = nullptr;
      g.ntasks = uint32_t(r->ntasks);
      HIP_OK(hipEventCreateWithFlags(&r->uploaded, hipEventDisableTiming));
      HIP_OK(hipEventRecord(r->uploaded, bs));
      // copies straight from a pinned layout return at once: wait for them,
      // so the upload time counts the copies, as the staged path's does
      if (L.cols.pinned()) HIP_OK(hipEventSynchronize(r->copied));
      {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stats.upload_s += std::chrono::duration<double>(Clock::now() - c0).count();
        S->ready.push_back(r);
      }
      S->cv_sched.notify_all();
    }
  } catch (const Error& e) {
    set_fatal(S, e.code, e.what());
  } catch (const std::exception& e) {
    set_fatal(S, BBTC_ERR_INTERNAL, e.what());
  }
}

// Reads back a finished request's counters and frees its block.
void finish(bbtc_server* S, Request& r) {
  // kSdmaMinBytes at least (the block holds them past the counters)
  const uint64_t rd = std::max<uint64_t>(r.ntasks * 8, kSdmaMinBytes);
  std::vector<uint64_t> per(rd / 8, 0);
  if (r.ntasks) {
    HIP_OK(hipMemcpyAsync(per.data(), r.seg.per_task, rd, hipMemcpyDeviceToHost, S->readback));
    HIP_OK(hipStreamSynchronize(S->readback));
  }
  uint64_t tot = 0;
  for (uint64_t t = 0; t < r.ntasks; ++t) tot += per[t];
  r.total = tot;
  if (r.per_task_out && r.ntasks) {
    if (r.order.empty()) std::copy(per.begin(), per.begin() + r.ntasks, r.per_task_out);
    else
      for (uint64_t q = 0; q < r.ntasks; ++q) r.per_task_out[r.order[q]] = per[q];
  }
  if (r.uploaded) (void)hipEventDestroy(r.uploaded);
  r.uploaded = nullptr;
  if (r.copied) (void)hipEventDestroy(r.copied);
  r.copied = nullptr;
  std::lock_guard<std::mutex> lk(S->mu);
  r.done_s = S->now();
  S->heap.release(r.heap_off, r.bytes);
  S->active--;
  push_result(S, r);
  S->cv_admit.notify_all();
  S->cv_done.notify_all();
}

// Scheduler thread: builds and queues launches, retires finished ones.
void sched_loop(bbtc_server* S) {
  try {
    HIP_OK(hipSetDevice(S->ctx->device));
    std::vector<std::shared_ptr<Request>> act;  // resident, in admission order
    std::deque<Launch> inflight;
    std::vector<int> ring_free;
    for (int i = int(S->max_inflight) - 1; i >= 0; --i) ring_free.push_back(i);
    for (;;) {
      bool progress = false;
      // newly uploaded requests
      {
        std::lock_guard<std::mutex> lk(S->mu);
        while (!S->ready.empty()) {
          auto& r = S->ready.front();
          const hipError_t e = hipEventQuery(r->uploaded);
          if (e == hipErrorNotReady) break;
          HIP_OK(e);
          r->ready_s = S->now();
          act.push_back(r);
          S->ready.pop_front();
          progress = true;
        }
      }
      // requests without work finish at once
      for (size_t x = 0; x < act.size();) {
        if (act[x]->nitems == 0) {
          finish(S, *act[x]);
          act.erase(act.begin() + long(x));
          progress = true;
        } else {
          ++x;
        }
      }
      // retire launches in order
      while (!inflight.empty()) {
        const hipError_t e = hipEventQuery(inflight.front().done);
        if (e == hipErrorNotReady) break;
        HIP_OK(e);
        Launch l = std::move(inflight.front());
        inflight.pop_front();
        for (auto& r : l.finishing) {
          finish(S, *r);
          act.erase(std::find(act.begin(), act.end(), r));
        }
        float ms = 0.f;
        HIP_OK(hipEventElapsedTime(&ms, l.start, l.done));
        {
          std::lock_guard<std::mutex> lk(S->mu);
          S->stats.busy_s += ms * 1e-3;
        }
        if (S->adaptive) {
          // next budget: the work this launch would have done in the target
          // time at its measured rate, at most doubling or halving
          if (ms > 0.01f && l.work > 0) {
            const double want = l.work * S->launch_ms / double(ms);
            double& w = S->work_budget[l.kind];
            w = w > 0 ? std::min(2.0 * w, std::max(0.5 * w, want)) : want;
          }
        }
        S->event_pool.push_back(l.start);
        S->event_pool.push_back(l.done);
        ring_free.push_back(l.ring);
        progress = true;
      }
      // A request in the row form is counted by a count_rows launch of its
      // own (the whole request, oldest first), taking turns with the shared
      // launches while both have work.
      Request* rr = nullptr;
      bool shared_work = false;
      for (auto& r : act)
        if (r->next < r->nitems) {
          if (!r->rows) shared_work = true;
          else if (!rr) rr = r.get();
        }
      if (!ring_free.empty() && rr && (!shared_work || S->rows_turn)) {
        S->rows_turn = false;
        Launch l;
        l.ring = ring_free.back();
        ring_free.pop_back();
        l.kind = 2;
        hipStream_t s = S->compute;
        RowArgs a = rr->ra;
        a.ctl = S->ctl;
        while (S->event_pool.size() < 2) {
          hipEvent_t e;
          HIP_OK(hipEventCreate(&e));
          S->event_pool.push_back(e);
        }
        l.start = S->event_pool.back();
        S->event_pool.pop_back();
        l.done = S->event_pool.back();
        S->event_pool.pop_back();
        l.items = rr->nitems - rr->next;
        HIP_OK(hipEventRecord(l.start, s));
        enqueue_rows(S->ctx, a, s);
        HIP_OK(hipEventRecord(l.done, s));
        rr->next = rr->nitems;
        rr->launches++;
        if (rr->start_s < 0) rr->start_s = S->now();
        for (auto& x : act)
          if (x.get() == rr) l.finishing.push_back(x);
        {
          std::lock_guard<std::mutex> lk(S->mu);
          S->stats.launches++;
          S->stats.segments++;
          S->stats.items += l.items;
          S->stats.row_form_requests++;
          S->stats.peak_segments = std::max<uint32_t>(S->stats.peak_segments, 1u);
        }
        inflight.push_back(std::move(l));
        progress = true;
      } else if (rr) {
        S->rows_turn = true;
      }
      // queue a launch over every request with items left
      if (!ring_free.empty() && shared_work) {
        const int ring = ring_free.back();
        SegDesc* sh = S->seg_host[ring];
        uint32_t nseg = 0;
        uint64_t total = 0;
        bool lane_only = true;  // every segment's request uses the lane form only
        Launch l;
        l.ring = ring;
        const double t = S->now();
        // The oldest request with work left keeps oldest_pct percent of the
        // launch's item budget (which bounds how much the arrivals behind it
        // slow a large request); the rest goes shortest remaining work first.
        std::vector<Request*> cand;  // admission order
        for (auto& r : act)
          if (r->next < r->nitems && !r->rows) cand.push_back(r.get());
        // A launch holds requests of one kind: all in the lane form (run by
        // the lane kernel) or none; the two kinds take turns while both have
        // work.
        bool kinds[2] = {false, false};
        for (Request* r : cand) kinds[r->plan.lane_only ? 1 : 0] = true;
        if (kinds[0] && kinds[1]) {
          const bool lane_turn = S->lane_turn;
          S->lane_turn = !S->lane_turn;
          cand.erase(std::remove_if(cand.begin(), cand.end(),
                                    [&](Request* r) { return r->plan.lane_only != lane_turn; }),
                     cand.end());
        }
        std::vector<uint64_t> take(cand.size(), 0);
        l.kind = !cand.empty() && cand[0]->plan.lane_only ? 1 : 0;
        std::vector<size_t> by_rem(cand.size());
        for (size_t x = 0; x < cand.size(); ++x) by_rem[x] = x;
        if (S->adaptive && !S->sequential && !cand.empty()) {
          // budget in estimated work; the first launch of a kind takes the
          // work of kDefaultQuantum items of its oldest request
          Request* o = cand[0];
          double& wb = S->work_budget[l.kind];
          double budget = wb > 0 ? wb
                                 : work_at(*o, std::min(o->nitems, o->next + kDefaultQuantum)) - work_at(*o, o->next);
          std::vector<double> rem(cand.size());
          for (size_t x = 0; x < cand.size(); ++x) rem[x] = cand[x]->wpref.back() - work_at(*cand[x], cand[x]->next);
          const double mine = std::min(rem[0], budget * S->oldest_pct / 100.0);
          take[0] = items_for_work(*o, o->next, mine);
          budget -= mine;
          std::stable_sort(by_rem.begin(), by_rem.end(), [&](size_t x, size_t y) { return rem[x] < rem[y]; });
          for (size_t x : by_rem) {
            if (budget <= 0) break;
            Request* r = cand[x];
            const uint64_t from = r->next + take[x];
            if (from >= r->nitems) continue;
            const double left = r->wpref.back() - work_at(*r, from);
            const double w = std::min(left, budget);
            take[x] += items_for_work(*r, from, w);
            budget -= w;
          }
        } else {
          uint64_t budget = S->sequential ? ~0ull : S->quantum[l.kind];
          if (!S->sequential && !cand.empty()) {
            take[0] = std::min(cand[0]->nitems - cand[0]->next, budget / 100 * S->oldest_pct);
            budget -= take[0];
          }
          std::stable_sort(by_rem.begin(), by_rem.end(), [&](size_t x, size_t y) {
            return cand[x]->nitems - cand[x]->next < cand[y]->nitems - cand[y]->next;
          });
          for (size_t x : by_rem) {
            const uint64_t more = std::min(cand[x]->nitems - cand[x]->next - take[x], budget);
            take[x] += more;
            budget -= more;
          }
        }
        for (size_t x : by_rem) {
          Request* r = cand[x];
          if (nseg == S->max_segments) break;
          if (take[x] == 0) continue;
          SegDesc d = r->seg;
          d.item0 = r->next;
          d.item1 = r->next + take[x];
          d.launch0 = total;
          sh[nseg++] = d;
          l.work += work_at(*r, r->next + take[x]) - work_at(*r, r->next);
          total += take[x];
          lane_only = lane_only && r->plan.lane_only;
          r->next += take[x];
          r->launches++;
          if (r->start_s < 0) r->start_s = t;
          if (r->next == r->nitems)
            for (auto& a : act)
              if (a.get() == r) l.finishing.push_back(a);
        }
        if (nseg) {
          ring_free.pop_back();
          hipStream_t s = S->compute;
          HIP_OK(hipMemcpyAsync(S->seg_dev[ring], sh, nseg * sizeof(SegDesc), hipMemcpyHostToDevice, s));
          CountArgs a{};
          a.segs = S->seg_dev[ring];
          a.nseg = nseg;
          a.item_end = total;
          a.ctl = S->ctl;
          a.spill = S->spill;
          a.spill_base = S->spill_base;
          a.spill_cap = kSpillCap;
          while (S->event_pool.size() < 2) {
            hipEvent_t e;
            HIP_OK(hipEventCreate(&e));
            S->event_pool.push_back(e);
          }
          l.start = S->event_pool.back();
          S->event_pool.pop_back();
          l.done = S->event_pool.back();
          S->event_pool.pop_back();
          l.items = total;
          HIP_OK(hipEventRecord(l.start, s));
          enqueue_count(S->ctx, a, 0, total, lane_only, S->list_chunks ? kListChunkItems : 0, s,
                        S->defer_steps);
          HIP_OK(hipEventRecord(l.done, s));
          inflight.push_back(std::move(l));
          std::lock_guard<std::mutex> lk(S->mu);
          S->stats.launches++;
          S->stats.segments += nseg;
          S->stats.items += total;
          S->stats.peak_segments = std::max<uint32_t>(S->stats.peak_segments, nseg);
          progress = true;
        }
      }
      if (progress) continue;
      {
        std::unique_lock<std::mutex> lk(S->mu);
        if (S->stop && S->pending.empty() && S->ready.empty() && act.empty() && inflight.empty()) return;
        if (inflight.empty() && S->ready.empty()) {
          S->cv_sched.wait_for(lk, std::chrono::milliseconds(2));
          continue;
        }
      }
      // launches or uploads outstanding: poll shortly
      std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
  } catch (const Error& e) {
    set_fatal(S, e.code, e.what());
  } catch (const std::exception& e) {
    set_fatal(S, BBTC_ERR_INTERNAL, e.what());
  }
}

void server_free(bbtc_server* S) {
  if (!S) return;
  (void)hipSetDevice(S->ctx->device);
  for (auto* p : S->seg_dev) (void)hipFree(p);
  for (auto* p : S->seg_host) (void)hipHostFree(p);
  for (auto e : S->event_pool) (void)hipEventDestroy(e);
  if (S->heap_base) {
    const hipError_t e = hipFree(S->heap_base);
    if (e != hipSuccess) std::fprintf(stderr, "bbtc: freeing the server heap failed: %s\n", hipGetErrorString(e));
    (void)hipDeviceSynchronize();
  }
  if (S->ctl) (void)hipFree(S->ctl);
  if (S->spill) (void)hipFree(S->spill);
  if (S->spill_base) (void)hipFree(S->spill_base);
  for (hipStream_t s : {S->copy, S->build, S->compute, S->readback})
    if (s) (void)hipStreamDestroy(s);
  if (S->ctx) S->ctx->in_use -= std::min(S->ctx->in_use, S->heap.size());
  delete S;
}

}  // namespace
}  // namespace bbtc

#define SRV_TRY try {
#define SRV_CATCH                                                                \
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

extern "C" {

bbtc_status bbtc_server_create(bbtc_ctx* ctx, const bbtc_server_opts* opts, bbtc_server** out) {
  if (!ctx || !out) {
    bbtc::set_last_error("null argument");
    return BBTC_ERR_INVALID;
  }
  *out = nullptr;
  SRV_TRY
  HIP_OK(hipSetDevice(ctx->device));
  auto* S = new bbtc_server;
  S->ctx = ctx;
  try {
    size_t free_b = 0, total_b = 0;
    HIP_OK(hipMemGetInfo(&free_b, &total_b));
    uint64_t heap = opts && opts->heap_bytes ? opts->heap_bytes : bbtc::kDefaultHeap;
    heap = heap / bbtc::kAlign * bbtc::kAlign;
    if (!heap) bbtc::fail(BBTC_ERR_INVALID, "heap size is zero");
    if (ctx->budget && ctx->in_use + heap > ctx->budget)
      bbtc::fail(BBTC_ERR_BUDGET, "server heap exceeds the context's device budget");
    if (heap > total_b)
      bbtc::fail(BBTC_ERR_BUDGET, "server heap of " + std::to_string(heap) + " bytes exceeds device memory (" +
                                      std::to_string(total_b) + " bytes)");
    S->quantum[0] = S->quantum[1] = opts && opts->quantum_items ? opts->quantum_items : bbtc::kDefaultQuantum;
    S->adaptive = !(opts && opts->quantum_items);
    S->launch_ms = opts && opts->launch_us ? opts->launch_us * 1e-3 : bbtc::kDefaultLaunchMs;
    S->max_inflight = opts && opts->max_inflight ? opts->max_inflight : bbtc::kDefaultInflight;
    S->max_segments = opts && opts->max_segments ? opts->max_segments : bbtc::kDefaultSegments;
    S->sequential = opts && opts->sequential;
    S->lane_form = opts && opts->lane_form;
    if (opts && opts->oldest_pct > 100) bbtc::fail(BBTC_ERR_INVALID, "oldest_pct above 100");
    S->oldest_pct = opts && opts->oldest_pct ? opts->oldest_pct : bbtc::kDefaultOldestPct;
    S->list_chunks = opts && opts->list_chunks;
    S->defer_steps = opts ? opts->defer_steps : 0;
    if (S->sequential) S->max_inflight = 1;
    hipError_t he = hipMalloc(&S->heap_base, heap);
    if (he != hipSuccess) {
      // memory freed just before (another server's heap) can become
      // available only once the device has gone idle
      (void)hipGetLastError();
      (void)hipDeviceSynchronize();
      he = hipMalloc(&S->heap_base, heap);
    }
    if (he != hipSuccess) {
      (void)hipGetLastError();
      S->heap_base = nullptr;
      bbtc::fail(BBTC_ERR_BUDGET, "server heap of " + std::to_string(heap) + " bytes could not be allocated (" +
                                      std::to_string(free_b) + " of " + std::to_string(total_b) +
                                      " bytes free before the call)");
    }
    S->heap.init(heap);
    ctx->in_use += heap;
    HIP_OK(hipMalloc(&S->ctl, sizeof(bbtc::SlotCtl)));
    HIP_OK(hipMalloc(&S->spill, sizeof(bbtc::SpillRec) * bbtc::kSpillCap));
    HIP_OK(hipMalloc(&S->spill_base, sizeof(uint64_t) * bbtc::kSpillCap));
    for (uint32_t i = 0; i < S->max_inflight; ++i) {
      bbtc::SegDesc* d = nullptr;
      bbtc::SegDesc* h = nullptr;
      HIP_OK(hipMalloc(&d, sizeof(bbtc::SegDesc) * S->max_segments));
      S->seg_dev.push_back(d);
      HIP_OK(hipHostMalloc(reinterpret_cast<void**>(&h), sizeof(bbtc::SegDesc) * S->max_segments, hipHostMallocDefault));
      S->seg_host.push_back(h);
    }
    HIP_OK(hipStreamCreateWithFlags(&S->copy, hipStreamNonBlocking));
    HIP_OK(hipStreamCreateWithFlags(&S->build, hipStreamNonBlocking));
    HIP_OK(hipStreamCreateWithFlags(&S->compute, hipStreamNonBlocking));
    HIP_OK(hipStreamCreateWithFlags(&S->readback, hipStreamNonBlocking));
    S->epoch = bbtc::Clock::now();
    S->admit_thread = std::thread(bbtc::admit_loop, S);
    S->sched_thread = std::thread(bbtc::sched_loop, S);
  } catch (...) {
    bbtc::server_free(S);
    throw;
  }
  *out = S;
  SRV_CATCH
}

bbtc_status bbtc_server_reset_clock(bbtc_server* S) {
  if (!S) {
    bbtc::set_last_error("null argument");
    return BBTC_ERR_INVALID;
  }
  std::lock_guard<std::mutex> lk(S->mu);
  if (S->submitted != S->finished) {
    bbtc::set_last_error("requests are outstanding");
    return BBTC_ERR_INVALID;
  }
  S->epoch = bbtc::Clock::now();
  S->stats = bbtc_server_stats{};
  return BBTC_OK;
}

double bbtc_server_now(const bbtc_server* S) { return S ? S->now() : 0.0; }

bbtc_status bbtc_server_submit(bbtc_server* S, const bbtc_layout* L, const bbtc_task* tasks, uint64_t ntasks,
                               bbtc_kernel kernel, double arrival_s, uint64_t* per_task_out, uint64_t* id_out) {
  if (!S || !L || (!tasks && ntasks) || !id_out) {
    bbtc::set_last_error("null argument");
    return BBTC_ERR_INVALID;
  }
  if (!bbtc::valid_kernel(kernel)) {
    bbtc::set_last_error("unknown kernel choice");
    return BBTC_ERR_INVALID;
  }
  const uint32_t P = L->L.P;
  for (uint64_t q = 0; q < ntasks; ++q)
    if (!(tasks[q].i <= tasks[q].j && tasks[q].j <= tasks[q].k && tasks[q].k < P)) {
      bbtc::set_last_error("task out of range or not i <= j <= k");
      return BBTC_ERR_INVALID;
    }
  SRV_TRY
  bbtc::require_host_arrays(L->L, "bbtc_server_submit");
  auto r = std::make_shared<bbtc::Request>();
  r->L = &L->L;
  r->tiles = bbtc::tile_table(L->L);
  const std::vector<bbtc::TaskWork> work = bbtc::task_work(L->L, tasks, ntasks);
  const std::vector<bbtc::TaskRef> refs = bbtc::refs_of(P, tasks, ntasks);
  r->plan = bbtc::describe_plan(r->tiles, L->L.deg, refs, kernel, nullptr, S->lane_form, &work);
  bbtc::plan_request_rows(*r, L->L, refs, kernel, [&] {
    // every task in the lane form, as a lane-only plan of the plain path
    return S->lane_form ? r->plan.lane_only
                        : bbtc::describe_plan(r->tiles, L->L.deg, refs, kernel, nullptr, true, &work).lane_only;
  });
  if (!r->rows && kernel == BBTC_KERNEL_AUTO && !r->plan.lane_only && ntasks > 1) {
    // the tasks in the order a one-graph plan runs them (each form's tasks
    // grouped by bc tile, the forms interleaved); the counters go back to
    // the caller's order
    r->order = bbtc::mixed_order(r->tiles, r->plan, work, refs, ntasks, ntasks);
    std::vector<bbtc::TaskRef> mrefs(ntasks);
    std::vector<bbtc::TaskWork> mwork(ntasks);
    for (uint64_t q = 0; q < ntasks; ++q) {
      mrefs[q] = refs[r->order[q]];
      mwork[q] = work[r->order[q]];
    }
    r->plan = bbtc::describe_plan(r->tiles, L->L.deg, mrefs, kernel, nullptr, S->lane_form, &mwork);
  }
  r->ntasks = ntasks;
  r->nitems = r->plan.pref[ntasks];
  r->wpref.assign(ntasks + 1, 0.0);
  for (uint64_t q = 0; q < ntasks; ++q) {
    const uint64_t items = r->plan.pref[q + 1] - r->plan.pref[q];
    double w = 0;
    if (items) {  // nnz(ab) * max(deg(bc), deg(ac)), as the reference sorts tasks
      const bbtc::TaskDesc& d = r->plan.td[q];
      w = double(L->L.nnz[d.ab]) * std::max(double(L->L.deg[d.bc]), double(L->L.deg[d.ac]));
      if (!(w > 0) || !std::isfinite(w)) w = double(items);
    }
    r->wpref[q + 1] = r->wpref[q] + w;
  }
  r->per_task_out = per_task_out;
  r->arrival = arrival_s;
  r->jobs = bbtc::dense_jobs(L->L, 0, 0, 0, 0, &r->job_rows);
  r->bytes = bbtc::place(*r);
  {
    std::lock_guard<std::mutex> lk(S->mu);
    if (S->fatal != BBTC_OK) bbtc::fail(S->fatal, "server stopped: " + S->fatal_msg);
    if (S->stop) bbtc::fail(BBTC_ERR_INVALID, "server is shutting down");
    if (r->bytes > S->heap.size())
      bbtc::fail(BBTC_ERR_BUDGET, "request footprint of " + std::to_string(r->bytes) +
                                      " bytes exceeds the server heap of " + std::to_string(S->heap.size()));
    if (!S->pending.empty() && arrival_s < S->pending.back()->arrival)
      bbtc::fail(BBTC_ERR_INVALID, "arrival times must not decrease");
    r->id = S->next_id++;
    S->pending.push_back(r);
    S->submitted++;
  }
  S->cv_admit.notify_all();
  *id_out = r->id;
  SRV_CATCH
}

bbtc_status bbtc_server_poll(bbtc_server* S, bbtc_server_result* out, uint64_t max, int wait_ms, uint64_t* got) {
  if (!S || !got || (!out && max)) {
    bbtc::set_last_error("null argument");
    return BBTC_ERR_INVALID;
  }
  std::unique_lock<std::mutex> lk(S->mu);
  if (wait_ms != 0 && S->results.empty() && S->fatal == BBTC_OK) {
    auto pred = [&] { return !S->results.empty() || S->fatal != BBTC_OK || S->submitted == S->finished; };
    if (wait_ms < 0) S->cv_done.wait(lk, pred);
    else S->cv_done.wait_for(lk, std::chrono::milliseconds(wait_ms), pred);
  }
  uint64_t n = 0;
  while (n < max && !S->results.empty()) {
    out[n++] = S->results.front();
    S->results.pop_front();
  }
  *got = n;
  if (S->fatal != BBTC_OK) {
    bbtc::set_last_error("server stopped: " + S->fatal_msg);
    return S->fatal;
  }
  return BBTC_OK;
}

bbtc_status bbtc_server_wait_idle(bbtc_server* S) {
  if (!S) {
    bbtc::set_last_error("null argument");
    return BBTC_ERR_INVALID;
  }
  std::unique_lock<std::mutex> lk(S->mu);
  S->cv_done.wait(lk, [&] { return S->submitted == S->finished || S->fatal != BBTC_OK; });
  if (S->fatal != BBTC_OK) {
    bbtc::set_last_error("server stopped: " + S->fatal_msg);
    return S->fatal;
  }
  return BBTC_OK;
}

bbtc_status bbtc_server_get_stats(bbtc_server* S, bbtc_server_stats* out) {
  if (!S || !out) {
    bbtc::set_last_error("null argument");
    return BBTC_ERR_INVALID;
  }
  std::lock_guard<std::mutex> lk(S->mu);
  *out = S->stats;
  out->heap_bytes = S->heap.size();
  return BBTC_OK;
}

void bbtc_server_destroy(bbtc_server* S) {
  if (!S) return;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->stop = true;
  }
  S->cv_admit.notify_all();
  S->cv_sched.notify_all();
  if (S->admit_thread.joinable()) S->admit_thread.join();
  if (S->sched_thread.joinable()) S->sched_thread.join();
  (void)hipSetDevice(S->ctx->device);
  // every stream that touched the heap (compute, copy, read-back) must be
  // idle, or the runtime defers releasing the heap past this call
  (void)hipDeviceSynchronize();
  bbtc::server_free(S);
}

}  // extern "C"
