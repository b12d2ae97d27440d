// Host side of batched counting: block layouts for many graphs at once, and
// the per-graph choice of the number of cuts from measured device work.
#include <omp.h>

#include <chrono>
#include <cmath>
#include <limits>
#include <vector>

#include "handles.hpp"

namespace bbtc {
void set_last_error(const std::string& s);

namespace {
// size class of a graph for the cut rule
inline uint32_t bucket_of(uint64_t m) {
  uint32_t b = 0;
  while ((2ull << b) <= m + 1 && b < 63) ++b;
  return b;
}
}  // namespace
}  // namespace bbtc

extern "C" {

bbtc_status bbtc_layouts_build(const bbtc_hgraph* const* graphs, uint64_t n, const uint32_t* ncut,
                               bbtc_part kind, bbtc_layout** out) {
  if ((!graphs || !ncut || !out) && n) {
    bbtc::set_last_error("null argument");
    return BBTC_ERR_INVALID;
  }
  for (uint64_t x = 0; x < n; ++x) {
    out[x] = nullptr;
    if (!graphs[x] || ncut[x] < 1) {
      bbtc::set_last_error("null graph or ncut < 1");
      return BBTC_ERR_INVALID;
    }
  }
  // one graph per thread; the per-graph work runs single-threaded inside
  bbtc_status err = BBTC_OK;
  std::string msg;
  const int levels = omp_get_max_active_levels();
  omp_set_max_active_levels(1);
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t x = 0; x < int64_t(n); ++x) {
    try {
      auto cuts = bbtc::partition(graphs[x]->g, kind, ncut[x]);
      out[x] = new bbtc_layout{bbtc::build_layout(graphs[x]->g, cuts)};
    } catch (const bbtc::Error& e) {
#pragma omp critical
      {
        err = e.code;
        msg = e.what();
      }
    } catch (const std::exception& e) {
#pragma omp critical
      {
        err = BBTC_ERR_INTERNAL;
        msg = e.what();
      }
    }
  }
  omp_set_max_active_levels(levels);
  if (err != BBTC_OK) {
    for (uint64_t x = 0; x < n; ++x) {
      delete out[x];
      out[x] = nullptr;
    }
    bbtc::set_last_error(msg);
  }
  return err;
}

bbtc_status bbtc_tune_ncut(bbtc_ctx* ctx, const bbtc_hgraph* const* sample, uint64_t n,
                           const uint32_t* cands, uint32_t ncand, bbtc_part kind, int tune_forms,
                           uint32_t* rule, uint64_t* work_out) {
  if (!ctx || (!sample && n) || !cands || !ncand || !rule) {
    bbtc::set_last_error("null argument");
    return BBTC_ERR_INVALID;
  }
  const uint64_t none = std::numeric_limits<uint64_t>::max();
  // work[x][c]: device clocks of sample x counted with cands[c] cuts
  std::vector<uint64_t> work(n * ncand, none);
  for (uint32_t c = 0; c < ncand; ++c) {
    std::vector<uint32_t> nc(n, cands[c]);
    std::vector<bbtc_layout*> L(n, nullptr);
    bbtc_status st = bbtc_layouts_build(sample, n, nc.data(), kind, L.data());
    if (st != BBTC_OK) return st;
    for (uint64_t x = 0; x < n && st == BBTC_OK; ++x) {
      bbtc_dgraph* g = nullptr;
      const auto t_up = std::chrono::steady_clock::now();
      st = bbtc_graph_upload(ctx, L[x], &g);
      if (st != BBTC_OK) break;
      const double up_ns = std::chrono::duration<double, std::nano>(std::chrono::steady_clock::now() - t_up).count();
      const uint64_t nt = bbtc_task_count(cands[c]);
      std::vector<bbtc_task> t(nt);
      uint64_t got = 0;
      st = bbtc_tasks_build(L[x], BBTC_ORDER_DESC, t.data(), nt, &got);
      std::vector<uint8_t> choice(nt);
      const bool forms = (tune_forms & BBTC_TUNE_FORMS) != 0;
      if (st == BBTC_OK && forms) st = bbtc_autotune(ctx, g, t.data(), got, choice.data(), nullptr);
      bbtc_plan* p = nullptr;
      if (st == BBTC_OK)
        st = bbtc_plan_create_tuned(ctx, g, t.data(), got, BBTC_KERNEL_AUTO, forms ? choice.data() : nullptr, &p);
      if (tune_forms & BBTC_TUNE_BY_TIME) {
        // the upload, and the fastest of three launches after one that warms
        // the caches, in ns
        float ms[4] = {0.f, 0.f, 0.f, 0.f};
        if (st == BBTC_OK) st = bbtc_plan_time(p, 4, ms);
        if (st == BBTC_OK)
          work[x * ncand + c] = uint64_t(up_ns + double(std::min(ms[1], std::min(ms[2], ms[3]))) * 1e6);
      } else {
        std::vector<uint64_t> cyc(got);
        if (st == BBTC_OK) st = bbtc_plan_profile(p, cyc.data());
        if (st == BBTC_OK) {
          uint64_t s = 0;
          for (uint64_t v : cyc) s += v;
          work[x * ncand + c] = s;
        }
      }
      bbtc_plan_destroy(p);
      bbtc_graph_release(g);
    }
    for (bbtc_layout* l : L) bbtc_layout_free(l);
    if (st != BBTC_OK) return st;
  }
  // rule[b]: the candidate with the least total work over samples in bucket b;
  // buckets without samples take the nearest sampled bucket's choice (below
  // first); then no larger bucket takes fewer cuts than a smaller one
  std::vector<int> best(64, -1);
  for (uint32_t b = 0; b < 64; ++b) {
    uint64_t bestw = none;
    for (uint32_t c = 0; c < ncand; ++c) {
      uint64_t s = 0;
      bool any = false;
      for (uint64_t x = 0; x < n; ++x) {
        uint32_t nn = 0, nc = 0;
        uint64_t m = 0;
        bbtc_hgraph_info(sample[x], &nn, &nc, &m);
        if (bbtc::bucket_of(m) != b) continue;
        any = true;
        s += work[x * ncand + c] == none ? 0 : work[x * ncand + c];
      }
      if (any && s < bestw) {
        bestw = s;
        best[b] = int(c);
      }
    }
  }
  for (int b = 0; b < 64; ++b) {
    int c = best[b];
    for (int d = 1; c < 0 && d < 64; ++d) {
      if (b - d >= 0 && best[b - d] >= 0) c = best[b - d];
      else if (b + d < 64 && best[b + d] >= 0) c = best[b + d];
    }
    rule[b] = cands[c < 0 ? 0 : c];
    if (b > 0 && rule[b] < rule[b - 1]) rule[b] = rule[b - 1];
  }
  if (work_out)
    for (uint64_t x = 0; x < n * ncand; ++x) work_out[x] = work[x];
  return BBTC_OK;
}

uint32_t bbtc_rule_ncut(const uint32_t* rule, uint64_t m) {
  return rule ? rule[bbtc::bucket_of(m)] : 0;
}

}  // extern "C"
