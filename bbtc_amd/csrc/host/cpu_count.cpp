// CPU triangle count over the block layout: the --ngpu 0 path and the host
// tail of --border.
//
// Same task semantics as the reference CPU workers (bb_tc.cpp:118-139,
// 184-197): workers claim tasks from the back of the list with an atomic
// counter, here in units of 1,024 non-empty rows so that several workers
// share a heavy task; a task whose tiles include one with average degree
// > 10000 uses a per-worker stamp table (cpu_dense_kernel.cpp:3-51, merge for
// rows with d_ab < 4). Every other task marks each ac row met by two or more
// ab entries once in a per-worker table and probes its bc rows against it;
// the rest take the reference's sorted merge with early exit
// (cpu_sparse_kernel.cpp:3-59). The per-task counts are the same. Rows come
// from the compact non-empty list of the ab tile instead of a scan over the
// whole band.
#include <omp.h>

#include <algorithm>
#include <atomic>
#include <vector>

#include "graph.hpp"

namespace bbtc {

namespace {

constexpr float kDenseDegree = 10000.f;  // block_csr.h:35

inline uint64_t merge_count(const uint32_t* a, uint32_t na, const uint32_t* b, uint32_t nb) {
  uint32_t i = 0, j = 0;
  uint64_t c = 0;
  if (!na || !nb) return 0;
  const uint32_t alast = a[na - 1], blast = b[nb - 1];
  while (i < na && j < nb) {
    const uint32_t x = a[i], y = b[j];
    if (x == y) { ++i; ++j; ++c; }
    else if (x < y) { if (y > alast) break; ++i; }
    else { if (x > blast) break; ++j; }
  }
  return c;
}

struct TaskView {
  const uint32_t *ab_rp, *ac_rp, *bc_rp, *ab_c, *ac_c, *bc_c, *nz;
  uint32_t nnz_rows;
};

TaskView view(const Layout& L, const bbtc_task& x) {
  const size_t ab = L.tile(x.i, x.j), bc = L.tile(x.j, x.k), ac = L.tile(x.i, x.k);
  const uint32_t* rp = L.rowptr.data();
  const uint32_t* cl = L.cols.data();
  return TaskView{rp + L.rp_base[ab], rp + L.rp_base[ac], rp + L.rp_base[bc],
                  cl + L.col_base[ab], cl + L.col_base[ac], cl + L.col_base[bc],
                  L.nzrows.data() + L.nz_base[ab], L.nz_rows[ab]};
}

// Rows [z0, z1) of the ab tile's non-empty row list. An ac row met by two or
// more ab entries is marked once in the worker's epoch table, and each bc row
// is then probed against the marks (one look-up per bc entry instead of a
// merge through the ac row for every wedge); a bc row more than kMergeRatio
// times longer than the ac row, and the one wedge of a row with a single ab
// entry, is merged with early exit as the reference does. Both give the
// number of common entries.
constexpr uint32_t kMergeRatio = 128;

uint64_t sparse_rows(const TaskView& t, uint32_t z0, uint32_t z1, CpuWorker& w) {
  uint64_t c = 0;
  uint32_t* mark = w.mark.data();
  for (uint32_t z = z0; z < z1; ++z) {
    const uint32_t r = t.nz[z];
    const uint32_t a0 = t.ac_rp[r], a1 = t.ac_rp[r + 1];
    if (a0 == a1) continue;
    const uint32_t na = a1 - a0;
    const uint32_t b0 = t.ab_rp[r], b1 = t.ab_rp[r + 1];
    if (b1 - b0 < 2) {
      for (uint32_t y = b0; y < b1; ++y) {
        const uint32_t v = t.ab_c[y];
        c += merge_count(t.ac_c + a0, na, t.bc_c + t.bc_rp[v], t.bc_rp[v + 1] - t.bc_rp[v]);
      }
      continue;
    }
    if (w.epoch == 0xFFFFFFFFu) {  // every epoch used: clear and start again
      std::fill(w.mark.begin(), w.mark.end(), 0u);
      w.epoch = 0;
    }
    const uint32_t e = ++w.epoch;
    for (uint32_t x = a0; x < a1; ++x) mark[t.ac_c[x]] = e;
    for (uint32_t y = b0; y < b1; ++y) {
      const uint32_t v = t.ab_c[y];
      const uint32_t* bc = t.bc_c + t.bc_rp[v];
      const uint32_t nb = t.bc_rp[v + 1] - t.bc_rp[v];
      if (uint64_t(nb) > uint64_t(kMergeRatio) * na) {
        c += merge_count(t.ac_c + a0, na, bc, nb);
        continue;
      }
      uint32_t k = 0;
      for (uint32_t x = 0; x < nb; ++x) k += mark[bc[x]] == e;
      c += k;
    }
  }
  return c;
}

// Rows [z0, z1) with the worker's stamp table, which holds no row of
// [z0, z1) (it is cleared for every task a worker starts): stamp[c] == r
// marks column c of ac row r.
uint64_t dense_rows(const TaskView& t, uint32_t z0, uint32_t z1, uint32_t* stamp) {
  uint64_t c = 0;
  for (uint32_t z = z0; z < z1; ++z) {
    const uint32_t r = t.nz[z];
    const uint32_t a0 = t.ac_rp[r], a1 = t.ac_rp[r + 1];
    if (a0 == a1) continue;
    const uint32_t b0 = t.ab_rp[r], b1 = t.ab_rp[r + 1];
    if (b1 - b0 < 4) {
      for (uint32_t y = b0; y < b1; ++y) {
        const uint32_t v = t.ab_c[y];
        c += merge_count(t.ac_c + a0, a1 - a0, t.bc_c + t.bc_rp[v], t.bc_rp[v + 1] - t.bc_rp[v]);
      }
      continue;
    }
    for (uint32_t x = a0; x < a1; ++x) stamp[t.ac_c[x]] = r;
    for (uint32_t y = b0; y < b1; ++y) {
      const uint32_t v = t.ab_c[y];
      for (uint32_t x = t.bc_rp[v]; x < t.bc_rp[v + 1]; ++x) c += stamp[t.bc_c[x]] == r;
    }
  }
  return c;
}

bool dense_task(const Layout& L, const bbtc_task& x) {
  return L.deg[L.tile(x.i, x.k)] > kDenseDegree || L.deg[L.tile(x.j, x.k)] > kDenseDegree ||
         L.deg[L.tile(x.i, x.j)] > kDenseDegree;
}

void clear_stamp(std::vector<uint32_t>& stamp, uint32_t width) {
  if (stamp.size() < width) stamp.resize(width);
  std::fill(stamp.begin(), stamp.begin() + width, 0xFFFFFFFFu);
}

// The worker's tables ready for task x (columns of band k index them).
void prepare(const Layout& L, const bbtc_task& x, bool dense, CpuWorker& w) {
  const uint32_t width = L.height(x.k);
  if (dense) {
    clear_stamp(w.stamp, width);
  } else if (w.mark.size() < width) {
    w.mark.assign(width, 0u);
    w.epoch = 0;
  }
}

// Non-empty ab rows per claim unit of the CPU count.
constexpr uint32_t kRowsPerUnit = 1024;

}  // namespace

uint64_t count_task_cpu(const Layout& L, const bbtc_task& x, CpuWorker& w) {
  if (task_skipped(L, x)) return 0;
  const TaskView v = view(L, x);
  const bool dense = dense_task(L, x);
  prepare(L, x, dense, w);
  return dense ? dense_rows(v, 0, v.nnz_rows, w.stamp.data()) : sparse_rows(v, 0, v.nnz_rows, w);
}

// The workers claim units from one counter: the tasks from the back of the
// list (as the reference's workers claim tasks, bb_tc.cpp:184-197), each task
// cut into units of kRowsPerUnit non-empty ab rows, so every worker can share
// a heavy task instead of one worker holding it to the end. A worker adds its
// sum of a task's units to the task's count when it moves on; the sum of
// integers is the same whichever worker counts which rows.
uint64_t count_cpu(const Layout& L, const bbtc_task* tasks, uint64_t n, int nthreads,
                   uint64_t* per_task) {
  require_host_arrays(L, "the CPU count");
  // first[s]: the first unit of the s-th task from the back
  std::vector<uint64_t> first(n + 1, 0);
  for (uint64_t s = 0; s < n; ++s) {
    const bbtc_task& x = tasks[n - 1 - s];
    const uint64_t rows = task_skipped(L, x) ? 0 : L.nz_rows[L.tile(x.i, x.j)];
    first[s + 1] = first[s] + std::max<uint64_t>(1, (rows + kRowsPerUnit - 1) / kRowsPerUnit);
  }
  const uint64_t units = first[n];
  if (per_task) std::fill(per_task, per_task + n, 0);
  std::atomic<uint64_t> next(0);
  std::atomic<uint64_t> total(0);
  const int T = nthreads > 0 ? nthreads : omp_get_max_threads();
#pragma omp parallel num_threads(T)
  {
    CpuWorker w;
    uint64_t mine = 0, acc = 0;
    uint64_t s = 0, cur = n;  // cur: the task whose rows acc holds (n: none)
    TaskView v{};
    bool dense = false, skip = false;
    auto flush = [&]() {
      if (cur < n) {
        if (per_task) __atomic_fetch_add(&per_task[n - 1 - cur], acc, __ATOMIC_RELAXED);
        mine += acc;
      }
      acc = 0;
    };
    for (uint64_t u = next.fetch_add(1); u < units; u = next.fetch_add(1)) {
      while (first[s + 1] <= u) ++s;  // a worker's units only increase
      const bbtc_task& x = tasks[n - 1 - s];
      if (s != cur) {
        flush();
        cur = s;
        skip = task_skipped(L, x);
        if (!skip) {
          v = view(L, x);
          dense = dense_task(L, x);
          prepare(L, x, dense, w);
        }
      }
      if (skip) continue;
      const uint64_t z0 = (u - first[s]) * kRowsPerUnit;
      const uint32_t z1 = uint32_t(std::min<uint64_t>(z0 + kRowsPerUnit, v.nnz_rows));
      acc += dense ? dense_rows(v, uint32_t(z0), z1, w.stamp.data()) : sparse_rows(v, uint32_t(z0), z1, w);
    }
    flush();
    total.fetch_add(mine);
  }
  return total.load();
}

}  // namespace bbtc
