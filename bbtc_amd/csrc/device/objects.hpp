// Host-side objects of the device library (context, resident graph, plan),
// shared by the HIP translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../host/handles.hpp"
#include "descriptors.hpp"
#include "form_rule.hpp"

#define HIP_OK(expr)                                                                   \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess)                                                              \
      ::bbtc::fail(BBTC_ERR_HIP, std::string(#expr) + " failed: " + hipGetErrorString(e_));    \
  } while (0)

// ---------------------------------------------------------------------------
// host objects
// ---------------------------------------------------------------------------
namespace bbtc {
struct Staging;  // pinned host-to-device staging buffers (device.hip)
}

struct bbtc_ctx {
  int device = 0;
  int num_cu = 0;
  int blocks_per_cu = 1;  // resident blocks of count_probe per CU
  int blocks_per_cu_lane = 1;  // the same for its lane-form instantiation
  hipStream_t stream = nullptr;          // slot 0
  std::vector<hipStream_t> slots;         // slot i > 0, created on first use
  uint64_t budget = 0;
  uint64_t in_use = 0;
  unsigned char* scratch = nullptr;  // this device's address of the scratch array (bbtc_count_noalloc)
  int host_atomics = -1;             // device atomics on host memory supported (-1: not asked yet)
  bbtc::Staging* h2d = nullptr;      // created on the first large upload
};

struct bbtc_dgraph {
  bbtc_ctx* ctx = nullptr;
  uint32_t P = 0;
  uint64_t bytes = 0;
  uint32_t* rowptr = nullptr;
  uint32_t* cols = nullptr;
  uint32_t* nzrows = nullptr;
  uint32_t* dense = nullptr;
  bbtc::TileDesc* tiles = nullptr;
  std::vector<bbtc::TileDesc> htiles;
  std::vector<float> deg;
  std::vector<uint32_t> dwpr;  // bitmap words per row of each band
  // Work statistics of every task (i, j, k), i <= j <= k, in lexicographic
  // order (empty when there are more than kRuleMaxCuts cuts); task (i, j, k)
  // is lex_work[lex_ij[i * P + j] + k - j].
  std::vector<bbtc::TaskWork> lex_work;
  std::vector<uint64_t> lex_ij;
  uint64_t rowptr_words = 0;
  uint32_t* slots = nullptr;  // row slots of the lane form (4 words per rowptr word), built on first need
  // Row table of the row form (rows_S words per global row), built on first
  // need; rows_tried: the build was attempted (rows stays NULL when a row is
  // longer than kRowSlotMax or the table would be too large).
  uint32_t* rows = nullptr;
  uint32_t* rowlen = nullptr;  // entries per global row
  uint32_t rows_S = 0;
  bool rows_tried = false;
  std::vector<uint32_t> cuts;  // band starts (P + 1), from the band heights
};

struct bbtc_batch;

struct bbtc_plan {
  bbtc_ctx* ctx = nullptr;
  bbtc_dgraph* g = nullptr;
  uint64_t ntasks = 0;
  uint64_t nitems = 0;
  uint64_t bitmap_tasks = 0;
  uint64_t bytes = 0;
  bbtc::TaskDesc* tasks = nullptr;
  uint64_t* item_prefix = nullptr;
  bbtc::SegDesc* segs = nullptr;        // the whole plan: [0] plain, [1] with clock counters
  std::vector<uint64_t> hprefix;        // host copy of item_prefix
  std::vector<uint32_t> hkind;          // form per task (1 search, 2 bitmap, 3 dense, 4 rdense, 5 lane)
  // The plan's own task order: its task q is the caller's task order[q]
  // (empty: the caller's order). Tasks [mixed, ntasks) keep their places.
  std::vector<uint64_t> order;
  uint64_t mixed = 0;
  bool lane_only = false;               // every task with items uses the lane form
  // Row form (a lane-only plan of a graph with a row table): launch() counts
  // every task in one count_rows launch over the graph's rows; task ranges,
  // profiling and the shared queue keep the lane form.
  bool rows_form = false;
  uint32_t* lex2plan = nullptr;         // lexicographic task -> plan task
  uint32_t* cuts = nullptr;             // the graph's band starts (P + 1)
  uint64_t chunk = 0;                   // first-phase chunk (0: one per part; batches: list order)
  unsigned long long* per_task = nullptr;
  bbtc::SlotCtl* ctl = nullptr;         // one control block per launch slot
  struct Spill {
    bbtc::SpillRec* rec = nullptr;
    uint64_t* base = nullptr;
  };
  std::vector<Spill> spill;             // per launch slot; slot 0 made with the plan
  uint32_t spill_cap = 0;               // records per slot
  unsigned long long* cycles = nullptr; // per-task clock counters while profiling
  bool profiling = false;
  unsigned long long* track_left = nullptr;  // per-graph completion (batch-owned)
  unsigned long long* track_done = nullptr;
  hipStream_t last_stream = nullptr;
};

