// The kernel-form rule of BBTC_KERNEL_AUTO: a cost per form from a task's
// work statistics (bbtc_tasks_stats), the cheapest form where it applies.
//
// Each cost is a linear model of the shader clocks the form's work items
// take, fitted (non-negative least squares, relative error) to the per-task
// clocks that bbtc_autotune measured on one MI355X over 9,772 tasks of twelve
// graphs: RMAT scale 16-26 with 2-32 pal cuts, RMAT scale 23 with 16 uni
// cuts, and G(n,m) graphs of 2^18-2^22 vertices and average degree 16-64.
// Fitted on eleven of the graphs, the rule's choice on the twelfth costs
// 1.00-1.66x the clocks of the cheapest measured form (RMAT scale 26, 32 cuts:
// 1.20x), against 2.9x for the rule without statistics on scale 26.
//
// Terms: R rows with non-empty ab(u) and ac(u); W wedges (sum d_ab); A sum
// d_ac; Pbc probes of bc lists; Pac probes of ac lists; Lac bitmap-row lines
// met by ac probes; dwpr bitmap words per row of band k.
#pragma once

#include <cstdint>

namespace bbtc {

struct TaskWork {
  float R, W, A, Pbc, Pac, Lac;
};

// Estimated shader clocks of a task's work items in each form.
struct FormCosts {
  double merge, bitmap, dense, rdense;
};

inline FormCosts form_costs(const TaskWork& s, uint32_t dwpr) {
  const double R = s.R, W = s.W, A = s.A, Pbc = s.Pbc, Pac = s.Pac, Lac = s.Lac, d = dwpr;
  FormCosts c;
  // sorted ac(u) in LDS, binary search per bc(v) entry
  c.merge = 8938.2 + 6.0413 * W + 27.155 * Pbc + 11.110 * Pac + 341.79 * R;
  // ac(u) bitmaps in LDS, one bit test per bc(v) entry (set-up and clearing per row)
  c.bitmap = 2655180.8 + 43.492 * W + 19.116 * Pbc + 284.13 * R + 7.9506 * R * d;
  // ac(u) bitmap AND the bc(v) bitmap row, 16 bytes per lane step
  c.dense = 2.1755 * W + 4.3492 * W * d + 128.28 * R + 13.342 * R * d + 36.203 * Lac;
  // ac(u) entries tested in the bc(v) bitmap row
  c.rdense = 35.575 * W + 22.122 * Pac + 307.49 * R + 4.5224 * A + 26.275 * Lac;
  return c;
}

// Units of 64 consecutive non-empty ab rows read per task when the statistics
// are estimated (bbtc_tasks_stats sample_units): about 1-3% error on the probe
// counts of RMAT scale 24, with the same form choices as exact statistics.
constexpr uint32_t kRuleSampleUnits = 16;
// Above this many cuts the statistics of every task are not gathered at
// upload (C(130,3) = 357,760 tasks at 128 cuts) and AUTO keeps the rule
// without statistics.
constexpr uint32_t kRuleMaxCuts = 128;

}  // namespace bbtc
