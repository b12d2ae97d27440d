// Device queue + host tail: the --border split of the reference.
//
// The reference (bbtc.cu:54-62; bb_tc.cpp:142-197) gives tasks
// [0, Separator) to GPU worker threads only, and lets GPU workers (from the
// front, one launch per task) and CPU workers (from the back, one task at a
// time) race for [Separator, |T|) with a per-task atomic_flag. Here both
// sides claim from one packed 64-bit word (next front task, one past the
// next back task), so every task is counted exactly once, and the other
// OpenMP threads of the caller's team claim single tasks from the back.
// Where the device's atomics reach host memory (run_shared) the word lives
// in host-coherent memory and the device's own waves move the front inside
// one persistent launch; otherwise (run_hybrid) one host thread moves it on
// the device's behalf, feeding launches of tasks from the front, up to
// `nstreams` in flight on separate streams.
//
// Each task's work is estimated from its sampled work statistics (the
// algorithmic bytes they imply). A host worker takes the back task only
// while one host thread would finish it in half the time the device needs
// for its work still ahead of that task, at the rates measured so far (the
// device's from its finished launches, the host's from its finished tasks;
// fixed first guesses before that); otherwise it stops. The first launch
// covers every task the host workers could not reach at those rates, the
// rest goes in guided batches (1/(4 * nstreams) of the items left, at least
// enough to keep every CU busy), and once no host worker is left, or the
// host may take nothing (the whole list is device-only), the rest is one
// launch.
#include <omp.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <exception>
#include <mutex>
#include <vector>

#include "device_api.hpp"
#include "handles.hpp"

namespace bbtc {
void set_last_error(const std::string& s);

namespace {

inline uint64_t pack(uint64_t front, uint64_t back_end) { return (front << 32) | back_end; }

inline void add_to(std::atomic<double>& a, double x) {
  double cur = a.load();
  while (!a.compare_exchange_weak(cur, cur + x)) {
  }
}

struct HybridResult {
  uint64_t total = 0;
  uint64_t gpu_tasks = 0;
};

// First guesses of the rates, in algorithmic bytes per second: the device
// (RMAT s26 counts at about 7.5e12) and one host thread (somewhat below what
// the host count of the same graph reaches per thread); each is replaced by
// its measured rate, and the rates measured last (in this process) are the
// next count's first guesses.
constexpr double kDeviceRate = 4e12;
constexpr double kHostRate = 1e9;
std::atomic<double> last_device_rate(kDeviceRate), last_host_rate(kHostRate);
// A host worker takes a task only if it needs at most this share of the
// device's remaining time for it.
constexpr double kHostShare = 0.5;
constexpr uint32_t kStatUnits = 16;  // units of 64 rows read per task for the estimates

// The device and the host workers claim from one word in host-coherent
// memory, (front << 32) | back: the device's waves move the front with
// system-scope atomics as they reach tasks [lo, n) (the last part of the
// launch), the host workers take single tasks from the back under the same
// rule as above. One launch counts the whole list; a wave that meets a task
// a host worker took leaves the rest of that part to the host.
template <class Work>
HybridResult run_shared(bbtc_ctx* ctx, bbtc_plan* plan, const Layout& L, const bbtc_task* tasks, uint64_t n,
                        uint64_t lo, int T, const Work& work, std::vector<uint64_t>& cpu_counts,
                        uint64_t* per_task) {
  struct Word {
    uint64_t* p;
    ~Word() { dev::queue_free(p); }
  } word{dev::queue_alloc(ctx)};
  uint64_t* W = word.p;
  __atomic_store_n(W, pack(lo, n), __ATOMIC_SEQ_CST);
  using Clock = std::chrono::steady_clock;
  const double dev_guess = last_device_rate.load(), host_guess = last_host_rate.load();
  std::atomic<double> host_work(0.0), host_secs(0.0), dev_secs(0.0);
  std::atomic<bool> dev_finished(false);
  std::exception_ptr err;
  std::mutex errm;
  Clock::time_point t0{};
  auto host_rate = [&]() {
    const double s2 = host_secs.load();
    return s2 > 0.01 ? host_work.load() / s2 : host_guess;
  };
#pragma omp parallel num_threads(T)
  {
    if (omp_get_thread_num() == 0) {
      try {
        t0 = Clock::now();
        dev::plan_launch_shared(plan, W, uint32_t(lo));
      } catch (...) {
        std::lock_guard<std::mutex> lk(errm);
        if (!err) err = std::current_exception();
        __atomic_store_n(W, pack(0, 0), __ATOMIC_SEQ_CST);
      }
    }
#pragma omp barrier
    if (omp_get_thread_num() == 0) {
      try {
        if (!err) dev::plan_slot_wait(plan, 0);
      } catch (...) {
        std::lock_guard<std::mutex> lk(errm);
        if (!err) err = std::current_exception();
      }
      dev_secs.store(std::chrono::duration<double>(Clock::now() - t0).count());
      dev_finished.store(true);
    } else {
      CpuWorker w;
      for (;;) {
        uint64_t cur = __atomic_load_n(W, __ATOMIC_SEQ_CST);
        uint64_t q = 0;
        bool got = false;
        for (;;) {
          const uint64_t f = cur >> 32, b = cur & 0xFFFFFFFFull;
          if (b <= f || b <= lo || dev_finished.load()) break;
          q = b - 1;
          // the device's work still ahead of q: from where its rate puts it
          // now (at most its front) to q
          const double el = std::chrono::duration<double>(Clock::now() - t0).count();
          uint64_t from = f;
          {
            const double done = el * dev_guess;
            uint64_t a = 0, z = f;  // largest position with work(0, pos) <= done, capped at f
            while (a < z) {
              const uint64_t mid = (a + z + 1) / 2;
              if (work(0, mid) <= done) a = mid;
              else z = mid - 1;
            }
            from = std::min(a, q);
          }
          if (work(q, q + 1) / host_rate() > kHostShare * work(from, q) / dev_guess) break;
          if (__atomic_compare_exchange_n(W, &cur, pack(f, q), false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {
            got = true;
            break;
          }
        }
        if (!got) break;
        const Clock::time_point ts = Clock::now();
        cpu_counts[q] = count_task_cpu(L, tasks[q], w);
        add_to(host_secs, std::chrono::duration<double>(Clock::now() - ts).count());
        add_to(host_work, work(q, q + 1));
      }
    }
  }
  if (err) std::rethrow_exception(err);
  const uint64_t fin = __atomic_load_n(W, __ATOMIC_SEQ_CST);
  const uint64_t front = fin >> 32, back = fin & 0xFFFFFFFFull;
  if (dev_secs.load() > 1e-4) last_device_rate.store(work(0, front) / dev_secs.load());
  if (host_secs.load() > 0.01) last_host_rate.store(host_work.load() / host_secs.load());
  std::vector<uint64_t> gpu_counts(n, 0);
  dev::plan_copy_counts(plan, gpu_counts.data());
  HybridResult r;
  // tasks [0, back) are the device's (those in [front, back) have no items),
  // [back, n) the host workers'
  r.gpu_tasks = back;
  for (uint64_t q = 0; q < n; ++q) {
    const uint64_t c = q < back ? gpu_counts[q] : cpu_counts[q];
    if (per_task) per_task[q] = c;
    r.total += c;
  }
  return r;
}

HybridResult run_hybrid(bbtc_ctx* ctx, bbtc_dgraph* g, const Layout& L, const bbtc_task* tasks,
                        uint64_t n, uint64_t sep_in, int nthreads, int nstreams, bbtc_kernel kernel,
                        const uint8_t* choice, uint64_t* per_task) {
  if (sep_in < n) require_host_arrays(L, "the host tail of bbtc_count_hybrid");
  const int T = std::max(1, nthreads > 0 ? nthreads : omp_get_max_threads());
  // The plan may list in its own order only tasks the host workers never
  // take: all of them without host workers, else those before the border
  // (and, with the shared queue, before the launch's last claim part).
  const bool hosts = T > 1 && sep_in < n;
  bbtc_plan* plan =
      dev::plan_create_split(ctx, g, tasks, n, kernel, choice, hosts ? sep_in : n, hosts && dev::host_atomics(ctx));
  struct Guard {
    bbtc_plan* p;
    ~Guard() { bbtc_plan_destroy(p); }
  } guard{plan};
  dev::plan_zero_sync(plan);
  const uint64_t sep = std::max(sep_in, dev::plan_mixed(plan));
  // estimated work of each task (algorithmic bytes with 4-byte offsets) and
  // its prefix over the list
  std::vector<double> pre(n + 1, 0.0);
  if (T > 1 && sep < n) {
    // from the statistics the upload gathered, else sampled here
    std::vector<double> est(n);
    if (!dev::graph_task_work(g, tasks, n, est.data())) {
      std::vector<uint64_t> w(n * kTaskStats);
      task_stats(L, tasks, n, w.data(), kStatUnits);
      for (uint64_t q = 0; q < n; ++q) {
        const uint64_t* x = w.data() + kTaskStats * q;
        est[q] = 16.0 * x[0] + 12.0 * x[1] + 4.0 * x[2] + 4.0 * x[3];
      }
    }
    for (uint64_t q = 0; q < n; ++q) pre[q + 1] = pre[q] + 1.0 + est[q];
  }
  auto work = [&](uint64_t a, uint64_t b) { return pre[b] - pre[a]; };

  std::vector<uint64_t> cpu_counts(n, 0);
  if (T > 1 && sep < n && dev::host_atomics(ctx)) {
    const uint64_t lo = std::max<uint64_t>(sep, dev::plan_contested_from(plan));
    if (lo < n) return run_shared(ctx, plan, L, tasks, n, lo, T, work, cpu_counts, per_task);
  }
  std::atomic<uint64_t> state(pack(0, n));
  const int S = std::max(1, std::min(nstreams, dev::max_slots()));
  const uint64_t target = std::max<uint64_t>(1, dev::plan_batch_items(plan));
  std::exception_ptr err;
  std::mutex errm;
  using Clock = std::chrono::steady_clock;
  const Clock::time_point t_start = Clock::now();
  auto seconds = [&](Clock::time_point t) { return std::chrono::duration<double>(t - t_start).count(); };
  // measured rates (0: not yet), the end of the device's finished launches,
  // and the host workers still claiming
  std::atomic<double> dev_rate(0.0), host_work(0.0), host_secs(0.0);
  std::atomic<uint64_t> dev_done(0);
  std::atomic<int> host_active(0);
  const double dev_guess = last_device_rate.load(), host_guess = last_host_rate.load();
  auto device_rate = [&]() { const double r = dev_rate.load(); return r > 0 ? r : dev_guess; };
  auto host_rate = [&]() {
    const double s2 = host_secs.load();
    return s2 > 0.01 ? host_work.load() / s2 : host_guess;
  };

#pragma omp parallel num_threads(T)
  {
    const int tid = omp_get_thread_num();
    // host workers of the team actually granted (it may be smaller than T)
    if (tid == 0) host_active.store(omp_get_num_threads() - 1);
#pragma omp barrier
    if (tid == 0) {
      try {
        struct InFlight {
          uint64_t t1 = 0;
          bool busy = false;
        };
        std::vector<InFlight> fl(static_cast<size_t>(S));
        Clock::time_point first_launch{};
        bool any = false;
        int slot = 0;
        auto retire = [&](int sl) {
          if (!fl[sl].busy) return;
          fl[sl].busy = false;
          // launches on different slots may finish out of order; the done
          // mark only moves forward
          uint64_t d = dev_done.load();
          while (fl[sl].t1 > d && !dev_done.compare_exchange_weak(d, fl[sl].t1)) {
          }
          const double el = seconds(Clock::now()) - seconds(first_launch);
          if (el > 1e-4) dev_rate.store(work(0, dev_done.load()) / el);
        };
        for (;;) {
          if (!dev::plan_slot_idle(plan, slot)) dev::plan_slot_wait(plan, slot);
          retire(slot);
          for (int x = 0; x < S; ++x)
            if (fl[x].busy && dev::plan_slot_idle(plan, x)) retire(x);
          uint64_t cur = state.load();
          uint64_t f = 0, b = 0, t1 = 0;
          bool got = false;
          for (;;) {
            f = cur >> 32;
            b = cur & 0xFFFFFFFFull;
            if (f >= b) break;
            if (sep >= b || host_active.load() == 0) {
              t1 = b;  // nothing left the host may take: one launch
            } else if (!any) {
              // first launch: every task past the border the host workers
              // could not reach at the first-guess rates
              const double hosts = double(T - 1) * host_rate();
              uint64_t lo = std::max<uint64_t>(f + 1, std::min(sep, b)), hi = b;
              while (lo < hi) {  // smallest t1 whose tail the host finishes in time
                const uint64_t mid = (lo + hi) / 2;
                if (work(mid, b) / hosts <= kHostShare * work(f, mid) / device_rate()) hi = mid;
                else lo = mid + 1;
              }
              t1 = lo;
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
          if (!any) {
            first_launch = Clock::now();
            any = true;
          }
          dev::plan_launch_range(plan, f, t1, slot);
          fl[slot] = InFlight{t1, true};
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
          // the device's work still ahead of q: from the end of its finished
          // launches (or the front) to q
          const uint64_t from = std::min(q, dev_done.load());
          if (work(q, q + 1) / host_rate() > kHostShare * work(from, q) / device_rate()) break;
          if (state.compare_exchange_weak(cur, pack(f, q))) { got = true; break; }
        }
        if (!got) break;
        const Clock::time_point t0 = Clock::now();
        cpu_counts[q] = count_task_cpu(L, tasks[q], w);
        const double el = std::chrono::duration<double>(Clock::now() - t0).count();
        add_to(host_secs, el);
        add_to(host_work, work(q, q + 1));
      }
      host_active.fetch_sub(1);
    }
  }
  if (err) std::rethrow_exception(err);
  if (dev_rate.load() > 0) last_device_rate.store(dev_rate.load());
  if (host_secs.load() > 0.01) last_host_rate.store(host_work.load() / host_secs.load());

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
