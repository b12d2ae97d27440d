// CPU triangle count over the block layout: the --ngpu 0 path and the host
// tail of --border.
//
// Same task semantics as the reference CPU workers (bb_tc.cpp:118-139,
// 184-197): workers claim tasks from the back of the list with an atomic
// counter; a task whose tiles include one with average degree > 10000 uses a
// per-worker stamp table (cpu_dense_kernel.cpp:3-51, merge for rows with
// d_ab < 4), every other task a sorted merge with early exit
// (cpu_sparse_kernel.cpp:3-59). Rows come from the compact non-empty list of
// the ab tile instead of a scan over the whole band.
#include <omp.h>

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

uint64_t sparse_task(const TaskView& t) {
  uint64_t c = 0;
  for (uint32_t z = 0; z < t.nnz_rows; ++z) {
    const uint32_t r = t.nz[z];
    const uint32_t a0 = t.ac_rp[r], a1 = t.ac_rp[r + 1];
    if (a0 == a1) continue;
    for (uint32_t y = t.ab_rp[r]; y < t.ab_rp[r + 1]; ++y) {
      const uint32_t v = t.ab_c[y];
      c += merge_count(t.ac_c + a0, a1 - a0, t.bc_c + t.bc_rp[v], t.bc_rp[v + 1] - t.bc_rp[v]);
    }
  }
  return c;
}

uint64_t dense_task(const TaskView& t, std::vector<uint32_t>& stamp, uint32_t width) {
  if (stamp.size() < width) stamp.resize(width);
  std::fill(stamp.begin(), stamp.begin() + width, 0xFFFFFFFFu);
  uint64_t c = 0;
  for (uint32_t z = 0; z < t.nnz_rows; ++z) {
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

}  // namespace

uint64_t count_task_cpu(const Layout& L, const bbtc_task& x, std::vector<uint32_t>& stamp) {
  if (task_skipped(L, x)) return 0;
  const TaskView v = view(L, x);
  const bool dense = L.deg[L.tile(x.i, x.k)] > kDenseDegree ||
                     L.deg[L.tile(x.j, x.k)] > kDenseDegree ||
                     L.deg[L.tile(x.i, x.j)] > kDenseDegree;
  return dense ? dense_task(v, stamp, L.height(x.k)) : sparse_task(v);
}

uint64_t count_cpu(const Layout& L, const bbtc_task* tasks, uint64_t n, int nthreads,
                   uint64_t* per_task) {
  std::atomic<int64_t> next(int64_t(n) - 1);
  std::atomic<uint64_t> total(0);
  const int T = nthreads > 0 ? nthreads : omp_get_max_threads();
#pragma omp parallel num_threads(T)
  {
    std::vector<uint32_t> stamp;
    uint64_t mine = 0;
    for (int64_t q = next.fetch_sub(1); q >= 0; q = next.fetch_sub(1)) {
      const uint64_t c = count_task_cpu(L, tasks[q], stamp);
      if (per_task) per_task[q] = c;
      mine += c;
    }
    total.fetch_add(mine);
  }
  return total.load();
}

}  // namespace bbtc
