// Autotuner: per-task choice of the intersection form from measured cost.
//
// The reference picks its GPU kernel per task from two static thresholds
// (bb_tc.cpp:98: bitmap iff width(ac) < 46336 and h(ab) < 46336). Here every
// form is measured instead: the whole task list runs once per form (merge,
// bitmap, dense, rdense - each forced wherever it applies), with per-item shader
// clock counters, while all tasks share the device as in a real count. Each
// task keeps the form whose items took the fewest clocks.
#include <cstdint>
#include <limits>
#include <vector>

#include "handles.hpp"

namespace bbtc {
void set_last_error(const std::string& s);
}

extern "C" bbtc_status bbtc_autotune(bbtc_ctx* ctx, bbtc_dgraph* g, const bbtc_task* tasks,
                                     uint64_t n, uint8_t* choice_out, uint64_t* cycles_out) {
  if (!ctx || !g || (!tasks && n) || (!choice_out && n)) {
    bbtc::set_last_error("null argument");
    return BBTC_ERR_INVALID;
  }
  const uint64_t kNone = std::numeric_limits<uint64_t>::max();
  // form codes as in bbtc_plan_create_tuned / bbtc_plan_task_kinds
  constexpr int kForms = 4;
  const bbtc_kernel forced[kForms] = {BBTC_KERNEL_MERGE, BBTC_KERNEL_BITMAP, BBTC_KERNEL_DENSE,
                                      BBTC_KERNEL_RDENSE};
  std::vector<uint64_t> best(n, kNone);
  std::vector<uint8_t> choice(n, 1);
  std::vector<uint64_t> cyc(n);
  std::vector<uint8_t> kinds(n);
  if (cycles_out)
    for (uint64_t x = 0; x < kForms * n; ++x) cycles_out[x] = kNone;
  for (int f = 0; f < kForms; ++f) {
    // the tasks this form applies to: a forced form that does not apply
    // falls back to another one, which the plan reports
    bbtc_plan* p = nullptr;
    bbtc_status st = bbtc_plan_create(ctx, g, tasks, n, forced[f], &p);
    if (st != BBTC_OK) return st;
    st = bbtc_plan_task_kinds(p, kinds.data());
    if (st == BBTC_OK && f == 0) {  // warm caches and code once
      st = bbtc_plan_launch(p, nullptr);
      if (st == BBTC_OK) st = bbtc_plan_results(p, nullptr, nullptr);
    }
    bbtc_plan_destroy(p);
    if (st != BBTC_OK) return st;
    std::vector<uint64_t> idx;
    std::vector<bbtc_task> sub;
    for (uint64_t q = 0; q < n; ++q)
      if (kinds[q] == uint8_t(f + 1)) {
        idx.push_back(q);
        sub.push_back(tasks[q]);
      }
    if (sub.empty()) continue;
    st = bbtc_plan_create(ctx, g, sub.data(), sub.size(), forced[f], &p);
    if (st != BBTC_OK) return st;
    st = bbtc_plan_profile(p, cyc.data());
    bbtc_plan_destroy(p);
    if (st != BBTC_OK) return st;
    for (size_t x = 0; x < idx.size(); ++x) {
      const uint64_t q = idx[x];
      if (cycles_out) cycles_out[kForms * q + f] = cyc[x];
      if (cyc[x] < best[q]) {
        best[q] = cyc[x];
        choice[q] = uint8_t(f + 1);
      }
    }
  }
  for (uint64_t q = 0; q < n; ++q) choice_out[q] = choice[q];
  return BBTC_OK;
}
