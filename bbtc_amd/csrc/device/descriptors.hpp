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
constexpr int kItemBatch = 16;   // work items a wave claims per atomic
// Splitting of heavy items. A first-phase item that has spent more than
// kDeferSteps wave steps (one step: 64 probes, one 64-wedge batch, or one
// 16-byte bitmap word per lane) hands its remaining wedges to the second phase
// in chunks of about kChunkSteps steps each; a launch slot holds at most
// kSpillCap such records.
constexpr uint32_t kDeferSteps = 512;
constexpr uint32_t kChunkSteps = 256;
constexpr uint32_t kSpillCap = 1u << 20;
constexpr int kSpillShift = 36;  // SlotCtl::spill = (records << 36) | chunks
// BBTC_KERNEL_AUTO uses the bitmap form when at least this many rows' bitmaps
// of band k fit the wave slice (band width <= kLdsWords * 32 / this).
constexpr int kAutoBitmapRows = 1;
constexpr uint32_t kKindSearch = 0;
constexpr uint32_t kKindBitmap = 1;
constexpr uint32_t kKindDense = 2;
constexpr uint32_t kKindRdense = 3;

struct TileDesc {
  uint64_t rp;     // offset of the tile's row offsets in rowptr
  uint64_t col;    // offset of the tile's columns in cols
  uint64_t nz;     // offset of the tile's non-empty row list in nzrows
  uint64_t dense;  // offset of the tile's bitmap rows in dense (~0: none)
  uint32_t nzr;    // number of non-empty rows
  uint32_t h;      // band height
  uint32_t dwpr;   // bitmap words per row (multiple of 4)
  uint32_t pad;
};
static_assert(sizeof(TileDesc) == 48, "TileDesc layout");

struct TaskDesc {
  uint32_t ab, bc, ac;  // tile ids i*P+j, j*P+k, i*P+k
  uint32_t kind;        // kKindSearch, kKindBitmap, kKindDense or kKindRdense
  uint32_t rows;        // rows of the ab non-empty list per item (1..64)
  uint32_t wpr;         // bitmap words per row (bitmap and dense kinds, multiple of 4)
  uint32_t graph;        // index of the task's graph in a batch (0 otherwise)
  uint32_t pad1;
};
static_assert(sizeof(TaskDesc) == 32, "TaskDesc layout");

// Control words of one launch slot, reset before each launch.
struct SlotCtl {
  unsigned long long head;   // next first-phase item
  unsigned long long spill;  // (records << kSpillShift) | chunks reserved
  unsigned long long head2;  // next second-phase chunk
  unsigned long long pad;
};

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
};
static_assert(sizeof(SegDesc) == 104, "SegDesc layout");

struct CountArgs {
  const SegDesc* __restrict__ segs;  // launch0 increasing
  uint32_t nseg;
  uint32_t phase;             // 0: items from ctl->head; 1: chunks from ctl->head2
  uint64_t item_end;          // launch-local items [ctl->head at launch, item_end)
  SlotCtl* ctl;               // head set to the first item of the launch beforehand
  SpillRec* spill;            // spill_cap records (first phase writes, second reads)
  uint64_t* spill_base;       // first chunk of each record (increasing)
  uint32_t spill_cap;         // 0: no splitting
  uint32_t pad;
  // Per-graph completion (batches; NULL: off): units still to finish, and
  // the device wall clock when a graph's last unit finished.
  unsigned long long* left;
  unsigned long long* done;
};

}  // namespace bbtc
