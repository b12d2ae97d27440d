// Constants and descriptors shared by the device kernels and the host-side
// objects that prepare their launches.
#pragma once

#include <cstdint>

namespace bbtc {

constexpr int kWave = 64;
constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / kWave;
constexpr int kLdsWords = 1792;  // bitmap / sorted-list slice of a wave
constexpr int kMarkWords = 4 * 64;  // per-step wedge marks of a wave
constexpr int kWaveLds = kLdsWords + kMarkWords;  // 8 KiB of LDS per wave
constexpr int kMaxSlots = 64;    // concurrent range launches of one plan
constexpr int kItemBatch = 8;    // work items a wave claims per atomic (even: the lane
                                 // kernel claims half as many for one graph)
// First-phase items are dealt to kQueues parts in chunks of consecutive
// items (chunk c to part c % kQueues); each part has its own head counter in
// its own 128-byte line and is claimed in its own order. A block starts on
// part blockIdx % kQueues (its XCD when blocks are dealt to the 8 XCDs in
// turn) and moves to the next part when its own is drained. Eight counters
// take the claims of 256 CUs instead of one, and each XCD works through
// chunks of its own (the tiles its L2 holds). One graph's launch uses one
// chunk per part (each XCD on its own eighth of the task list: the least
// sharing of tiles between L2s); batches and the server use chunks of
// kListChunkItems, so the parts advance through the item list together and
// items early in the list (small graphs, old requests) finish early.
constexpr int kQueues = 8;
constexpr uint64_t kListChunkItems = 4096;
constexpr int kCtlStride = 16;   // 64-bit words per counter line
// Splitting of heavy items. A first-phase item that has spent more than
// kDeferSteps wave steps (one step: 64 probes, one 64-wedge batch, or one
// 16-byte bitmap word per lane) hands its remaining wedges to the second phase
// in chunks of about kChunkSteps steps each; a launch slot holds at most
// kSpillCap such records. A launch over one graph with at least
// kManyItemsPerWave items per resident wave splits at kDeferStepsMany steps
// instead: the machine is kept busy by other items, so fewer splits pay
// (RMAT s26: 614 vs 646 ms), while batches of small graphs and the server
// split early, where a heavy item is a graph's or request's critical path.
constexpr uint32_t kDeferSteps = 512;
constexpr uint32_t kDeferStepsMany = 1024;
constexpr uint64_t kManyItemsPerWave = 64;
constexpr uint32_t kChunkSteps = 256;
constexpr uint32_t kSpillCap = 1u << 20;
constexpr int kSpillShift = 36;  // SlotCtl::spill = (records << 36) | chunks
// bbtc_count_noalloc: the most tasks its fixed scratch array holds, the spill
// records it keeps (items past them are not split, only slower), and the
// array's size.
constexpr uint64_t kNoAllocMaxTasks = 1u << 16;
constexpr uint32_t kNoAllocSpillCap = 1u << 18;
constexpr uint64_t kScratchBytes = 12ull << 20;
// BBTC_KERNEL_AUTO uses the bitmap form when at least this many rows' bitmaps
// of band k fit the wave slice (band width <= kLdsWords * 32 / this).
constexpr int kAutoBitmapRows = 1;
constexpr uint32_t kKindSearch = 0;
constexpr uint32_t kKindBitmap = 1;
constexpr uint32_t kKindDense = 2;
constexpr uint32_t kKindRdense = 3;
constexpr uint32_t kKindLane = 4;
constexpr uint32_t kKindRows = 5;  // reported per task only: the row form counts a whole plan
// A per-task choice byte (bbtc_plan_create_tuned): the form in the low bits
// (0 the kernel's rule, 1-5 a form), and in bits 4-5 the rows per work item
// as a right shift of the form's default (64, or as many bitmaps as fit).
constexpr uint32_t kChoiceFormMask = 7;
constexpr uint32_t kChoiceRowsShift = 4;
constexpr uint32_t kRowsShifts = 3;  // the autotuner tries 64, 32 and 16 rows
// The lane form (one lane per row u, lists merged in HBM) applies to tasks
// whose three tiles hold at most this many entries in any row, which bounds
// a lane's work to kLaneMaxRow * 2 * kLaneMaxRow merge steps.
constexpr uint32_t kLaneMaxRow = 32;

struct TileDesc {
  uint64_t rp;     // offset of the tile's row offsets in rowptr
  uint64_t col;    // offset of the tile's columns in cols
  uint64_t nz;     // offset of the tile's non-empty row list in nzrows
  uint64_t dense;  // offset of the tile's bitmap rows in dense (~0: none)
  uint32_t nzr;    // number of non-empty rows
  uint32_t h;      // band height
  uint32_t dwpr;   // bitmap words per row (multiple of 4)
  uint32_t maxd;   // entries in the tile's longest row
};
static_assert(sizeof(TileDesc) == 48, "TileDesc layout");

struct TaskDesc {
  uint32_t ab, bc, ac;  // tile ids i*P+j, j*P+k, i*P+k
  uint32_t kind;        // kKindSearch, kKindBitmap, kKindDense, kKindRdense or kKindLane
  uint32_t rows;        // rows of the ab non-empty list per item (1..64)
  uint32_t wpr;         // bitmap words per row (bitmap and dense kinds, multiple of 4)
  uint32_t graph;        // index of the task's graph in a batch (0 otherwise)
  uint32_t pad1;
};
static_assert(sizeof(TaskDesc) == 32, "TaskDesc layout");

// Control words of one launch slot, reset before each launch.
struct SlotCtl {
  // part q: next position at head[q * kCtlStride], positions [0, end[q]);
  // position x of part q is item base + ((x / C) * kQueues + q) * C + x % C
  // with C = kChunkItems
  unsigned long long head[kQueues * kCtlStride];
  // second phase: the spilled chunks are split into kQueues contiguous
  // parts too; next chunk of part q (counted from the part's start)
  unsigned long long head2[kQueues * kCtlStride];
  unsigned long long end[(kQueues + kCtlStride - 1) / kCtlStride * kCtlStride];
  unsigned long long base;   // first item of the launch
  unsigned long long chunk;  // items per chunk (a multiple of kItemBatch)
  unsigned long long shift;  // log2(chunk), or 63 with one chunk per part
  unsigned long long spill;  // (records << kSpillShift) | chunks reserved
  unsigned long long front;  // largest front of the shared task queue seen (CountArgs::queue)
  unsigned long long pad[kCtlStride - 5];
};
static_assert(sizeof(SlotCtl) % 128 == 0, "SlotCtl layout");
static_assert(kListChunkItems % kItemBatch == 0, "claims must not straddle chunks");
static_assert((kListChunkItems & (kListChunkItems - 1)) == 0, "list chunks are a power of two");

// Chunk size of a launch of len items: `chunk` (a power of two and a
// multiple of kItemBatch), or for chunk == 0 one chunk per part, rounded up
// to whole claims; the shift that splits a part position into chunk and
// offset (63 with one chunk per part, where every position is below the
// chunk). Then the positions of part q, and the item at position x of part q
// (constexpr: used by the device and checked on the host).
constexpr unsigned long long part_chunk(unsigned long long len, unsigned long long chunk) {
  if (chunk) return chunk;
  const unsigned long long per = (len + kQueues - 1) / kQueues;
  const unsigned long long c = (per + kItemBatch - 1) / kItemBatch * kItemBatch;
  return c ? c : kItemBatch;
}
constexpr unsigned part_shift(unsigned long long chunk) {
  unsigned s = 0;
  while (chunk > 1) {
    chunk >>= 1;
    ++s;
  }
  return s;
}
constexpr unsigned part_shift(unsigned long long len, unsigned long long chunk) {
  return chunk ? part_shift(chunk) : 63u;
}
constexpr unsigned long long part_positions(unsigned long long len, unsigned q, unsigned long long C) {
  const unsigned long long nch = (len + C - 1) / C;
  const unsigned long long mine = nch > q ? (nch - 1 - q) / kQueues + 1 : 0;
  const unsigned long long last = len - (nch ? nch - 1 : 0) * C;
  return mine == 0 ? 0 : (nch - 1) % kQueues == q ? (mine - 1) * C + last : mine * C;
}
constexpr unsigned long long part_item(unsigned long long base, unsigned q, unsigned long long x,
                                       unsigned long long C, unsigned shift) {
  return base + ((x >> shift) * kQueues + q) * C + (x & ((1ull << shift) - 1));
}

// One tile whose bitmap rows are built on the device: its row offsets (h + 1
// tile-local words at rp), columns (at col), bitmap rows (h rows of dwpr words
// at dense), and row0, the number of bitmap rows of the jobs before it.
struct DenseJob {
  uint64_t rp, col, dense, row0;
  uint32_t h, dwpr;
};
static_assert(sizeof(DenseJob) == 40, "DenseJob layout");

// Wedges [w0, w1) of launch-local item `item` (in its flattened wedge order),
// counted by the second phase in chunks of wc wedges.
struct SpillRec {
  uint64_t item;
  uint32_t w0, w1, wc;
  uint32_t pad;
};
static_assert(sizeof(SpillRec) == 24, "SpillRec layout");

// One plan's item range in a launch. A launch counts the items of one or
// more segments, numbered consecutively from 0 (launch-local items).
struct SegDesc {
  const uint32_t* rowptr;
  const uint32_t* cols;
  const uint32_t* nzrows;
  const uint32_t* dense;
  const TileDesc* tiles;
  const TaskDesc* tasks;
  const uint64_t* item_prefix;  // ntasks + 1 entries
  unsigned long long* per_task;
  unsigned long long* cycles;   // per task: shader clocks its units took (NULL: off)
  uint64_t item0, item1;        // plan items [item0, item1)
  uint64_t launch0;             // launch-local index of item0
  uint32_t ntasks;
  uint32_t pad;
  // Row slots of the lane form (NULL: none): for row v of tile t, 16 bytes at
  // slots[4 * (rp(t) + v)]: the row's entry count, then its entries when it
  // has at most kSlotEntries, else its first entry's offset in the tile.
  const uint32_t* slots;
};
static_assert(sizeof(SegDesc) == 112, "SegDesc layout");
constexpr uint32_t kSlotEntries = 3;
// Row slots are built for a graph's lane-form plans only up to this size.
constexpr uint64_t kSlotMaxBytes = 8ull << 30;

// The row form (a single graph whose tasks all admit the lane form): every
// global row v gets a slot of S words (S a power of two, 4..kRowSlotMax)
// holding N+(v) over all bands, global ids ascending, padded with kRowPad,
// and its length in a separate array. Each row u is enumerated once: a
// wedge (u, v), v in N+(u), merges the part of u's slot after v with v's
// slot, and each common entry w adds one to task (band(u), band(v),
// band(w)). So N+(v) is read once per wedge for every k, instead of once
// per bc tile, and N+(u) once per row instead of once per band pair.
constexpr uint32_t kRowSlotMax = 64;
constexpr uint32_t kRowPad = 0xFFFFFFFFu;
constexpr uint64_t kRowTableMaxBytes = 8ull << 30;
constexpr uint32_t kRowMaxCuts = 256;  // band starts are kept in LDS
constexpr uint32_t kRowItemBatch = 1;  // items (64 rows each) a wave claims per atomic

// Index of task (i, j, k), i <= j <= k < P, in lexicographic order:
// sum over i' < i of (P - i')(P - i' + 1)/2, then sum over j' in [i, j) of
// P - j', then k - j.
constexpr uint64_t lex_task(uint32_t P, uint32_t i, uint32_t j, uint32_t k) {
  const uint64_t p = P, a = uint64_t(P) - i, d = uint64_t(j) - i;
  return (p * (p + 1) * (p + 2) - a * (a + 1) * (a + 2)) / 6 + d * p - (uint64_t(i) * 2 + d - 1) * d / 2 +
         (uint64_t(k) - j);
}

// lex_task against the enumeration itself (i, then j >= i, then k >= j).
constexpr bool lex_task_enumerates(uint32_t P) {
  uint64_t n = 0;
  for (uint32_t i = 0; i < P; ++i)
    for (uint32_t j = i; j < P; ++j)
      for (uint32_t k = j; k < P; ++k)
        if (lex_task(P, i, j, k) != n++) return false;
  return true;
}
static_assert(lex_task_enumerates(1) && lex_task_enumerates(2) && lex_task_enumerates(7) &&
                  lex_task_enumerates(32),
              "lex_task order");

struct RowArgs {
  const uint32_t* rows;         // row table: S words per global row
  const uint32_t* rowlen;       // entries per global row
  const uint32_t* cuts;         // P + 1 band starts
  const uint32_t* lex2plan;     // lexicographic task -> plan task (~0: not in the plan)
  unsigned long long* per_task;
  SlotCtl* ctl;                 // claim heads, set by reset_slot
  uint32_t nrows;
  uint32_t P;
  uint32_t s_shift;             // log2(S)
  uint32_t pad;
};

struct CountArgs {
  const SegDesc* __restrict__ segs;  // launch0 increasing
  uint32_t nseg;
  uint32_t phase;             // 0: items from ctl->head; 1: chunks from ctl->head2
  uint64_t item_end;          // launch-local items [first item, item_end)
  SlotCtl* ctl;               // heads and part ends set by reset_slot beforehand
  SpillRec* spill;            // spill_cap records (first phase writes, second reads)
  uint64_t* spill_base;       // first chunk of each record (increasing)
  uint32_t spill_cap;         // 0: no splitting
  uint32_t defer_steps;       // wave steps after which a first-phase unit spills
  // Per-graph completion (batches; NULL: off): units still to finish, and
  // the device wall clock when a graph's last unit finished.
  unsigned long long* left;
  unsigned long long* done;
  // Task queue shared with host workers (NULL: none; one segment): the
  // word (front << 32) | back in host-coherent memory. Tasks [front, back)
  // are unclaimed; the device claims from the front with system-scope
  // atomics, host workers from the back. Tasks below queue_lo are the
  // device's alone; a claim reaches about queue_ahead items past the start
  // of the task that needs it.
  unsigned long long* queue;
  uint32_t queue_lo;
  uint32_t queue_ahead;
};

}  // namespace bbtc
