// Device queue + host tail: the --border split of the reference.
//
// The reference (bbtc.cu:54-62; bb_tc.cpp:142-197) gives tasks
// [0, Separator) to GPU worker threads only, and lets GPU workers (from the
// front, one launch per task) and CPU workers (from the back, one task at a
// time) race for [Separator, |T|) with a per-task atomic_flag. Here one host
// thread feeds the device batches of tasks from the front of the list, each
// batch one persistent launch, with up to `nstreams` batches in flight on
// separate streams; the other OpenMP threads claim single tasks from the
// back. Both sides claim from one packed 64-bit word (next front task, one
// past the next back task), so every task is counted exactly once. A batch
// holds 1/(4 * nstreams) of the items left, and at least enough to keep
// every CU busy (guided self-scheduling: large launches while the host is
// far behind, small ones where the two sides meet). When the host may take
// no task (the whole list is device-only), the device counts it in one
// launch.
#include <omp.h>

#include <algorithm>
#include <atomic>
#include <exception>
#include <mutex>
#include <vector>

#include "device_api.hpp"
#include "handles.hpp"

namespace bbtc {
void set_last_error(const std::string& s);

namespace {

inline uint64_t pack(uint64_t front, uint64_t back_end) { return (front << 32) | back_end; }

struct HybridResult {
  uint64_t total = 0;
  uint64_t gpu_tasks = 0;
};

HybridResult run_hybrid(bbtc_ctx* ctx, bbtc_dgraph* g, const Layout& L, const bbtc_task* tasks,
                        uint64_t n, uint64_t sep, int nthreads, int nstreams, bbtc_kernel kernel,
                        const uint8_t* choice, uint64_t* per_task) {
  bbtc_plan* plan = nullptr;
  const bbtc_status st = bbtc_plan_create_tuned(ctx, g, tasks, n, kernel, choice, &plan);
  if (st != BBTC_OK) fail(st, bbtc_last_error());
  struct Guard {
    bbtc_plan* p;
    ~Guard() { bbtc_plan_destroy(p); }
  } guard{plan};
  dev::plan_zero_sync(plan);

  std::vector<uint64_t> cpu_counts(n, 0);
  std::atomic<uint64_t> state(pack(0, n));
  const int S = std::max(1, std::min(nstreams, dev::max_slots()));
  const uint64_t target = std::max<uint64_t>(1, dev::plan_batch_items(plan));
  const int T = std::max(1, nthreads > 0 ? nthreads : omp_get_max_threads());
  std::exception_ptr err;
  std::mutex errm;

#pragma omp parallel num_threads(T)
  {
    const int tid = omp_get_thread_num();
    if (tid == 0) {
      try {
        int slot = 0;
        for (;;) {
          if (!dev::plan_slot_idle(plan, slot)) dev::plan_slot_wait(plan, slot);
          uint64_t cur = state.load();
          uint64_t f = 0, b = 0, t1 = 0;
          bool got = false;
          for (;;) {
            f = cur >> 32;
            b = cur & 0xFFFFFFFFull;
            if (f >= b) break;
            if (sep >= b) {
              t1 = b;  // nothing left the host may take: one launch
            } else {
              const uint64_t want = std::max(target, dev::plan_items(plan, f, b) / (4 * uint64_t(S)));
              // smallest t1 in (f, b] whose items reach the batch size
              uint64_t lo = f + 1, hi = b;
              while (lo < hi) {
                const uint64_t mid = (lo + hi) / 2;
                if (dev::plan_items(plan, f, mid) >= want) hi = mid;
                else lo = mid + 1;
              }
              t1 = lo;
            }
            if (state.compare_exchange_weak(cur, pack(t1, b))) { got = true; break; }
          }
          if (!got) break;
          dev::plan_launch_range(plan, f, t1, slot);
          slot = (slot + 1) % S;
        }
      } catch (...) {
        std::lock_guard<std::mutex> lk(errm);
        if (!err) err = std::current_exception();
        state.store(pack(0, 0));  // stop the host workers too
      }
    } else {
      CpuWorker w;
      for (;;) {
        uint64_t cur = state.load();
        uint64_t q = 0;
        bool got = false;
        for (;;) {
          const uint64_t f = cur >> 32, b = cur & 0xFFFFFFFFull;
          if (b <= f || b <= sep) break;
          q = b - 1;
          if (state.compare_exchange_weak(cur, pack(f, q))) { got = true; break; }
        }
        if (!got) break;
        cpu_counts[q] = count_task_cpu(L, tasks[q], w);
      }
    }
  }
  if (err) std::rethrow_exception(err);

  std::vector<uint64_t> gpu_counts(n, 0);
  dev::plan_copy_counts(plan, gpu_counts.data());
  HybridResult r;
  r.gpu_tasks = state.load() >> 32;
  for (uint64_t q = 0; q < n; ++q) {
    const uint64_t c = q < r.gpu_tasks ? gpu_counts[q] : cpu_counts[q];
    if (per_task) per_task[q] = c;
    r.total += c;
  }
  return r;
}

}  // namespace
}  // namespace bbtc

extern "C" bbtc_status bbtc_count_hybrid(bbtc_ctx* ctx, bbtc_dgraph* g, const bbtc_layout* L,
                                         const bbtc_task* tasks, uint64_t n, uint64_t separator,
                                         int nthreads, int nstreams, bbtc_kernel kernel,
                                         const uint8_t* choice, uint64_t* per_task, uint64_t* total,
                                         uint64_t* gpu_tasks_out) {
  if (!L || (!tasks && n) || (!ctx) != (!g)) {
    bbtc::set_last_error("null argument (ctx and graph must both be set or both be NULL)");
    return BBTC_ERR_INVALID;
  }
  if (n >= 0xFFFFFFFFull) {
    bbtc::set_last_error("too many tasks");
    return BBTC_ERR_INVALID;
  }
  for (uint64_t q = 0; q < n; ++q)
    if (!(tasks[q].i <= tasks[q].j && tasks[q].j <= tasks[q].k && tasks[q].k < L->L.P)) {
      bbtc::set_last_error("task out of range or not i <= j <= k");
      return BBTC_ERR_INVALID;
    }
  try {
    uint64_t tot = 0, gt = 0;
    if (!ctx) {
      tot = bbtc::count_cpu(L->L, tasks, n, nthreads, per_task);
    } else {
      const auto r = bbtc::run_hybrid(ctx, g, L->L, tasks, n, separator, nthreads, nstreams, kernel,
                                      choice, per_task);
      tot = r.total;
      gt = r.gpu_tasks;
    }
    if (total) *total = tot;
    if (gpu_tasks_out) *gpu_tasks_out = gt;
  } catch (const bbtc::Error& e) {
    bbtc::set_last_error(e.what());
    return e.code;
  } catch (const std::exception& e) {
    bbtc::set_last_error(e.what());
    return BBTC_ERR_INTERNAL;
  }
  return BBTC_OK;
}
