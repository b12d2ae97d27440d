// Internal interface between the OpenMP host code (g++) and the HIP device
// code (hipcc) of libbbtc.so. No HIP types appear here; a "slot" names one of
// the context's streams and the plan's work counter of the same index.
#pragma once

#include <cstdint>

#include "bbtc.h"

namespace bbtc {
namespace dev {

// bbtc_plan_create_tuned for the --border split (throws on error): the plan
// may list tasks [0, upto) in an order of its own (with `contested`, up to
// the first task of a whole-plan launch's last claim part if that is
// further); tasks from plan_mixed on keep the caller's indices. Per-task
// outputs (plan_copy_counts, bbtc_plan_results) are in the caller's order.
bbtc_plan* plan_create_split(bbtc_ctx* ctx, bbtc_dgraph* g, const bbtc_task* tasks, uint64_t n, bbtc_kernel kernel,
                             const uint8_t* choice, uint64_t upto, bool contested);
uint64_t plan_mixed(const bbtc_plan* p);
// Zeroes every per-task counter of the plan on slot 0 and waits.
void plan_zero_sync(bbtc_plan* p);
// Enqueues the count of tasks [t0, t1) of the plan on slot `slot`.
void plan_launch_range(bbtc_plan* p, uint64_t t0, uint64_t t1, int slot);
// Work items of tasks [t0, t1) (prefix difference).
uint64_t plan_items(const bbtc_plan* p, uint64_t t0, uint64_t t1);
// Items the device keeps busy with: resident waves times a small factor.
uint64_t plan_batch_items(const bbtc_plan* p);
// True when every launch enqueued on the slot has finished.
bool plan_slot_idle(bbtc_plan* p, int slot);
void plan_slot_wait(bbtc_plan* p, int slot);
// Waits for all slots, copies all per-task counters to host (caller's order).
void plan_copy_counts(bbtc_plan* p, uint64_t* out);
int max_slots();
// Task queue shared by the device and host workers (CountArgs::queue):
// true when the device's atomics reach host memory over its link.
bool host_atomics(bbtc_ctx* ctx);
// A host-coherent word for the queue (128 bytes), and its release.
uint64_t* queue_alloc(bbtc_ctx* ctx);
void queue_free(uint64_t* q);
// The first task the host may share: every item of it and of the tasks
// after it lies in the last part of a whole-plan launch (ntasks: none).
uint64_t plan_contested_from(const bbtc_plan* p);
// Enqueues the count of the whole plan on slot 0 with the device claiming
// tasks [queue_lo, n) from the queue word's front (queue: host pointer).
void plan_launch_shared(bbtc_plan* p, uint64_t* queue, uint32_t queue_lo);
// Estimated work of each task (algorithmic bytes with 4-byte offsets) from the
// work statistics the graph gathered at upload; false when it has none.
bool graph_task_work(const bbtc_dgraph* g, const bbtc_task* tasks, uint64_t n, double* out);

}  // namespace dev
}  // namespace bbtc
