// Device code: the counting kernels for gfx950.
//
// Work decomposition. A plan lists work items; item x of task t covers
// rows [x*R, (x+1)*R) of the ab tile's compact non-empty row list (R <= 64,
// per task). A launch counts one or more segments (item ranges of plans,
// possibly of different graphs) numbered consecutively. One wave owns one
// unit at a time and draws the next from one of eight per-XCD head counters
// (the launch's items are dealt to eight parts), so a single persistent
// launch drains every task of every segment.
//
// Inside a unit the wave flattens its work twice with wave prefix sums:
//   rows  -> wedges (u, v), v in ab(u), for rows whose ac(u) is non-empty
//   wedges -> probes (u, v, w), w in bc(v)
// so each lane handles one probe per step, lanes read bc(v) lists
// contiguously, and a hub row's wedges spread over all 64 lanes instead of
// serialising one thread as the reference kernels do
// (gpu_sparse_kernel.cu:25-46, gpu_dense_kernel.cu:74-93).
//
// Two phases. The first takes whole items; an item that runs past a step
// budget spills its remaining wedges as chunks, which the second phase
// spreads over the whole machine, so a hub-core item never leaves one wave
// running alone at the end of a launch.
//
// Membership of w in ac(u) is tested in one of five forms, chosen per task:
//   search: ac(u) as a sorted list in the wave's LDS slice (a window of rows
//           at a time); a probe is a binary search.
//   bitmap: ac(u) as one bit per column of band k in LDS; a probe is one bit
//           test.
//   dense:  ac(u) bitmap in LDS AND the bc(v) bitmap row from HBM, popcount.
//   rdense: ac(u) entries tested against the bc(v) bitmap row in HBM.
//   lane:   one lane per wedge intersects bc(v) and ac(u) in HBM, four
//           entries of each at a time (rows of at most 32 entries; no LDS,
//           no second phase; a launch of lane-form tasks only runs the
//           count_probe_lanes instantiation at 8 waves per SIMD, which reads
//           a bc(v) of up to 3 entries from its 16-byte row slot: the count
//           and the entries in one load).
// Matches are counted with ballot + popcount (or a wave sum) into a
// wave-uniform 64-bit count and added to the task's slot with one atomic per
// task change.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "descriptors.hpp"

namespace bbtc {

// Inclusive prefix sum over the wave's 64 lanes with DPP moves (no LDS
// round trips): shifts by 1, 2, 4 and 8 lanes within each row of 16 (a lane
// whose source lies before its row adds 0), then lane 15 of rows 0 and 2
// added to rows 1 and 3 (row_bcast:15), then lane 31 to rows 2 and 3
// (row_bcast:31).
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x, int lane) {
  (void)lane;
  x += uint32_t(__builtin_amdgcn_update_dpp(0, int(x), 0x111, 0xf, 0xf, false));  // row_shr:1
  x += uint32_t(__builtin_amdgcn_update_dpp(0, int(x), 0x112, 0xf, 0xf, false));  // row_shr:2
  x += uint32_t(__builtin_amdgcn_update_dpp(0, int(x), 0x114, 0xf, 0xf, false));  // row_shr:4
  x += uint32_t(__builtin_amdgcn_update_dpp(0, int(x), 0x118, 0xf, 0xf, false));  // row_shr:8
  x += uint32_t(__builtin_amdgcn_update_dpp(0, int(x), 0x142, 0xa, 0xf, false));  // row_bcast:15
  x += uint32_t(__builtin_amdgcn_update_dpp(0, int(x), 0x143, 0xc, 0xf, false));  // row_bcast:31
  return x;
}

__device__ __forceinline__ uint32_t wave_sum32(uint32_t x) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) x += __shfl_xor(x, off, kWave);
  return x;
}

__device__ __forceinline__ unsigned long long wave_sum64(unsigned long long x) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) x += __shfl_xor(x, off, kWave);
  return x;
}

// Lane owning flattened index x: the number of lanes whose inclusive prefix
// is <= x (prefixes never decrease across lanes), clamped to the last lane.
__device__ __forceinline__ int owner_of(uint32_t incl, uint32_t x) {
  int q = 0;
#pragma unroll
  for (int s = kWave / 2; s >= 1; s >>= 1) {
    const uint32_t pv = __shfl(incl, q + s - 1, kWave);
    if (pv <= x) q += s;
  }
  return q > kWave - 1 ? kWave - 1 : q;
}

// Pointers into global memory (address space 1). The graph arrays are
// reached through descriptors loaded from memory, so without the cast the
// compiler cannot tell them from LDS and emits flat loads, which wait on
// both the vector-memory and the LDS counters.
template <class T>
using gptr = const __attribute__((address_space(1))) T*;
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
template <class T>
__device__ __forceinline__ gptr<T> as_global(const T* p) {
  return (gptr<T>)p;
}
// A descriptor struct (a multiple of 8 bytes, 8-byte aligned) read from
// global memory word by word.
template <class T>
__device__ __forceinline__ T gload(const T* p) {
  static_assert(sizeof(T) % 8 == 0, "descriptor size");
  T v;
  const gptr<uint64_t> q = (gptr<uint64_t>)p;
  uint64_t* d = reinterpret_cast<uint64_t*>(&v);
#pragma unroll
  for (int i = 0; i < int(sizeof(T) / 8); ++i) d[i] = q[i];
  return v;
}

template <class Ptr>
__device__ __forceinline__ bool sorted_contains(Ptr s, uint32_t n, uint32_t x) {
  uint32_t lo = 0, len = n;
  while (len) {
    const uint32_t half = len >> 1;
    if (s[lo + half] < x) {
      lo += half + 1;
      len -= half + 1;
    } else {
      len = half;
    }
  }
  return lo < n && s[lo] == x;
}

// |A[a, ae) ∩ B[b, be)| for sorted lists of distinct entries, four entries
// of each list at a time: all 16 pairs of the two current blocks are
// compared, and the block with the smaller last entry moves on (both when
// equal). A block past its list's end is padded with values no entry takes
// (ids are below 2^32 - 2), distinct between the lists.
//
// Block i..i+4 of P (entries at or past e read as pad).
__device__ __forceinline__ void load4(gptr<uint32_t> P, uint32_t i, uint32_t e, uint32_t pad, uint32_t (&x)[4]) {
#pragma unroll
  for (int k = 0; k < 4; ++k) x[k] = i + k < e ? P[i + k] : pad;
}

// The intersection from blocks x = A[a, a+4) and y = B[b, b+4) already
// loaded (a < ae, b < be); only the block that moves on is loaded again.
__device__ __forceinline__ uint32_t intersect_from(gptr<uint32_t> A, uint32_t a, uint32_t ae, gptr<uint32_t> B,
                                                   uint32_t b, uint32_t be, uint32_t (&x)[4], uint32_t (&y)[4]) {
  uint32_t n = 0;
  for (;;) {
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) n += x[i] == y[j] ? 1u : 0u;
    const uint32_t xl = x[3], yl = y[3];
    const bool na = xl <= yl, nb = yl <= xl;
    if (na) a += 4;
    if (nb) b += 4;
    if (a >= ae || b >= be) break;
    if (na) load4(A, a, ae, 0xFFFFFFFFu, x);
    if (nb) load4(B, b, be, 0xFFFFFFFEu, y);
  }
  return n;
}

__device__ __forceinline__ uint32_t and_popc(u32x4 a, u32x4 b) {
  return __popc(a.x & b.x) + __popc(a.y & b.y) + __popc(a.z & b.z) + __popc(a.w & b.w);
}

// Largest t with prefix[t] <= item (prefix[n] > item).
__device__ __forceinline__ uint32_t find_task(gptr<uint64_t> prefix, uint32_t n,
                                              uint64_t item) {
  uint32_t lo = 0, hi = n;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (prefix[mid] <= item) lo = mid;
    else hi = mid;
  }
  return lo;
}

// x as a wave-uniform (scalar) value; x must be equal in every lane.
__device__ __forceinline__ uint64_t uniform64(uint64_t x) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane(uint32_t(x));
  const uint32_t hi = __builtin_amdgcn_readfirstlane(uint32_t(x >> 32));
  return (uint64_t(hi) << 32) | lo;
}

__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// Zeroes the per-task counters (a fill kernel rather than a memset, so the
// launch sequence stays valid inside a captured graph).
__global__ void zero_counts(unsigned long long* __restrict__ p, uint64_t n) {
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// Resets a launch slot: first-phase items from i0, no spilled work.
__global__ void reset_slot(SlotCtl* __restrict__ c, unsigned long long i0, unsigned long long i1,
                           unsigned long long chunk) {
  const unsigned long long len = i1 > i0 ? i1 - i0 : 0;
  const unsigned long long C = part_chunk(len, chunk);
  for (unsigned q = 0; q < unsigned(kQueues); ++q) {
    c->head[q * kCtlStride] = 0;
    c->head2[q * kCtlStride] = 0;
    c->end[q] = part_positions(len, q, C);
  }
  c->base = i0;
  c->chunk = C;
  c->shift = part_shift(len, chunk);
  c->spill = 0;
  c->front = 0;
}

// True when task t (t >= queue_lo) is the device's: below the queue's front,
// or claimed now by moving the front past it (through about queue_ahead
// more items) while it is below the back. False once a host worker has it.
// Lane 0 does the atomics; every CAS that fails means another claimant
// moved the front or the back, so the loop ends.
__device__ __forceinline__ uint32_t claim_shared(const CountArgs& a, const uint64_t* prefix, uint32_t ntasks,
                                              uint32_t t) {
  const gptr<uint64_t> ip = as_global(prefix);
  unsigned long long cur = __hip_atomic_load(a.queue, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  for (;;) {
    const uint32_t f = uint32_t(cur >> 32), b = uint32_t(cur & 0xFFFFFFFFull);
    if (t < f) {
      atomicMax(&a.ctl->front, (unsigned long long)f);
      return 1;
    }
    if (t >= b) return 0;
    const uint64_t want = ip[t] + a.queue_ahead;
    uint32_t t1 = want >= ip[b] ? b : find_task(ip, ntasks, want) + 1;
    t1 = t1 < t + 1 ? t + 1 : (t1 > b ? b : t1);
    const unsigned long long next = (uint64_t(t1) << 32) | b;
    if (__hip_atomic_compare_exchange_strong(a.queue, &cur, next, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_SYSTEM)) {
      atomicMax(&a.ctl->front, (unsigned long long)t1);
      return 1;
    }
  }
}

// One persistent launch per phase. Phase 0 drains the first-phase items
// (kQueues parts with their own heads); a unit is an item's whole wedge range. Phase 1 drains the
// chunks that phase 0 spilled; a unit is one chunk of an item's wedges.
// kAll: every form (count_probe); otherwise the lane form only
// (count_probe_lanes: no LDS, fewer registers, more resident waves).
template <bool kAll>
__device__ __forceinline__ void count_units(const CountArgs& a, uint32_t* lds_all) {
  const int lane = threadIdx.x & (kWave - 1);
  uint32_t* lds = kAll ? lds_all + (threadIdx.x / kWave) * kWaveLds : nullptr;
  uint32_t* mk_d = lds + kLdsWords;  // x-index base of the wedge starting here
  uint32_t* mk_s = mk_d + kWave;     // its row's membership slot
  uint32_t* mk_n = mk_s + kWave;     // its row's list length (search form)
  uint32_t* mk_f = mk_n + kWave;     // a wedge starts here
  if constexpr (kAll) {
    // bitmap slices are all-zero between units
    for (int x = lane; x < kLdsWords; x += kWave) lds[x] = 0;
    wave_lds_sync();
  }

  // current segment and task; [t_lo, t_hi) are t_cur's launch-local items
  uint32_t s_cur = 0;
  uint32_t t_cur = 0;
  uint64_t t_lo = 1, t_hi = 0;   // (empty at start)
  uint64_t t_first = 0;          // plan item of t_cur's first item
  unsigned long long acc = 0;    // count of t_cur not yet added to its slot
  unsigned long long acc_cyc = 0;
  unsigned long long run = 0;    // units of t_cur finished since the last flush
  auto flush = [&]() {
    if (lane == 0) {
      const SegDesc& S = a.segs[s_cur];
      if (acc) atomicAdd(S.per_task + t_cur, acc);
      if (acc_cyc) atomicAdd(S.cycles + t_cur, acc_cyc);
      if (a.left && run) {
        const uint32_t gi = gload(S.tasks + t_cur).graph;
        if (atomicSub(a.left + gi, run) == run) a.done[gi] = wall_clock64();
      }
    }
    acc = 0;
    acc_cyc = 0;
    run = 0;
  };

  // second phase: spilled records and chunks, final once the first phase ended
  uint32_t nrec = 0;
  unsigned long long nchunks = 0;
  if (a.phase) {
    const unsigned long long sp = a.ctl->spill;
    nrec = uint32_t(sp >> kSpillShift);
    if (nrec > a.spill_cap) nrec = a.spill_cap;
    nchunks = sp & ((1ull << kSpillShift) - 1);
  }

  unsigned long long item = 0, batch_end = 0;
  uint32_t part = blockIdx.x % kQueues, parts_left = kQueues;  // current part; parts not yet drained
  const unsigned long long base = a.phase == 0 ? a.ctl->base : 0;
  const unsigned long long chunk = a.phase == 0 ? a.ctl->chunk : 1;
  const unsigned shift = a.phase == 0 ? unsigned(a.ctl->shift) : 0u;
  // The lane kernel has twice the waves; counting one graph, its claims are
  // halved so each XCD keeps about one task's tiles in flight (G(n=2^22,
  // m=2^26): 6.9 vs 7.2 ms). Batches and the server keep whole claims.
  const unsigned long long batch = (!kAll && shift == 63u) ? kItemBatch / 2 : kItemBatch;
  for (;;) {
    uint64_t it;
    uint32_t wlo, whi;
    bool may_defer;
    if (a.phase == 0) {
      // Items are claimed kItemBatch at a time from the current part (one
      // device-scope atomic per batch); a drained part sends the wave on to
      // the next one.
      if (item >= batch_end) {
        // wave-uniform loop; lane 0 does the atomic
        unsigned long long old = 0, pe = 0;
        while (parts_left) {
          pe = a.ctl->end[part];
          unsigned long long got = 0;
          if (lane == 0) got = atomicAdd(&a.ctl->head[part * kCtlStride], batch);
          old = uniform64(__shfl(got, 0, kWave));
          if (old < pe) break;
          part = part + 1 == kQueues ? 0 : part + 1;
          --parts_left;
        }
        if (!parts_left) break;
        // a claim stays inside one chunk (the batch divides the chunk)
        item = part_item(base, part, old, chunk, shift);
        batch_end = item + ((old + batch < pe ? old + batch : pe) - old);
      }
      it = item++;
      wlo = 0;
      whi = 0xffffffffu;
      may_defer = a.spill_cap != 0;
    } else {
      // one chunk per claim from the current part of the chunk list (part q:
      // chunks [nchunks * q / kQueues, nchunks * (q + 1) / kQueues)), so an
      // item's chunks mostly stay on one XCD
      unsigned long long c = 0;
      while (parts_left) {  // wave-uniform loop; lane 0 does the atomic
        const unsigned long long lo = nchunks * part / kQueues, hi = nchunks * (part + 1) / kQueues;
        unsigned long long got = 0;
        if (lane == 0) got = atomicAdd(&a.ctl->head2[part * kCtlStride], 1ull);
        c = lo + uniform64(__shfl(got, 0, kWave));
        if (c < hi) break;
        part = part + 1 == kQueues ? 0 : part + 1;
        --parts_left;
      }
      if (!parts_left || nrec == 0) break;
      uint32_t lo = 0, hi = nrec;  // record holding chunk c: last base <= c
      while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (as_global(a.spill_base)[mid] <= c) lo = mid;
        else hi = mid;
      }
      const SpillRec R = gload(a.spill + lo);
      it = R.item;
      wlo = R.w0 + uint32_t(c - as_global(a.spill_base)[lo]) * R.wc;
      whi = R.w1 - wlo > R.wc ? wlo + R.wc : R.w1;
      may_defer = false;
    }
    if (!(it >= t_lo && it < t_hi)) {
      flush();  // the finished task's count and progress, before moving on
      uint32_t sg = 0;
      if (a.nseg > 1) {  // last segment starting at or before it
        uint32_t lo = 0, hi = a.nseg;
        while (hi - lo > 1) {
          const uint32_t mid = (lo + hi) >> 1;
          if (a.segs[mid].launch0 <= it) lo = mid;
          else hi = mid;
        }
        sg = lo;
      }
      s_cur = __builtin_amdgcn_readfirstlane(sg);
      const SegDesc& S = a.segs[s_cur];
      const uint64_t pi = S.item0 + (it - S.launch0);
      const gptr<uint64_t> ip = as_global(S.item_prefix);
      t_cur = find_task(ip, S.ntasks, pi);
      t_first = ip[t_cur];
      const uint64_t e = ip[t_cur + 1];
      t_lo = S.launch0 + ((t_first > S.item0 ? t_first : S.item0) - S.item0);
      t_hi = S.launch0 + ((e < S.item1 ? e : S.item1) - S.item0);
      if (a.queue && a.phase == 0 && t_cur >= a.queue_lo) {
        uint32_t own = 0;
        if (lane == 0) own = t_cur < a.ctl->front ? 1u : claim_shared(a, S.item_prefix, S.ntasks, t_cur);
        if (!__builtin_amdgcn_readfirstlane(own)) {
          // a host worker has this task and every later one: the rest of
          // this part is theirs
          t_lo = 1;
          t_hi = 0;
          item = batch_end;
          part = part + 1 == kQueues ? 0 : part + 1;
          --parts_left;
          continue;
        }
      }
    }
    const SegDesc& S = a.segs[s_cur];
    const unsigned long long t_start = S.cycles ? clock64() : 0;
    const TaskDesc T = gload(S.tasks + t_cur);
    const TileDesc AB = gload(S.tiles + T.ab);
    const TileDesc BC = gload(S.tiles + T.bc);
    const TileDesc AC = gload(S.tiles + T.ac);
    const gptr<uint32_t> ab_rp = as_global(S.rowptr + AB.rp);
    const gptr<uint32_t> ac_rp = as_global(S.rowptr + AC.rp);
    const gptr<uint32_t> bc_rp = as_global(S.rowptr + BC.rp);
    const gptr<uint32_t> ab_c = as_global(S.cols + AB.col);
    const gptr<uint32_t> ac_c = as_global(S.cols + AC.col);
    const gptr<uint32_t> bc_c = as_global(S.cols + BC.col);
    const bool dense = T.kind == kKindDense;
    const bool rdense = T.kind == kKindRdense;
    const bool bitmap = T.kind == kKindBitmap || dense;  // ac(u) staged as bitmaps
    const gptr<uint32_t> bc_d = as_global(S.dense + BC.dense);  // bc bitmap rows (dense, rdense)

    // one row per lane
    const uint64_t z = (S.item0 + (it - S.launch0) - t_first) * T.rows + uint64_t(lane);
    uint32_t ab0 = 0, dab = 0, ac0 = 0, dac = 0;
    if (uint32_t(lane) < T.rows && z < AB.nzr) {
      const uint32_t r = as_global(S.nzrows)[AB.nz + z];
      ab0 = ab_rp[r];
      dab = ab_rp[r + 1] - ab0;
      ac0 = ac_rp[r];
      dac = ac_rp[r + 1] - ac0;
      if (dac == 0) dab = 0;
      if (dab == 0) dac = 0;
    }
    if (!kAll || T.kind == kKindLane) {
      // Lane form: one lane per wedge (u, v), v in ab(u), intersecting bc(v)
      // with ac(u) straight from HBM (rows of at most kLaneMaxRow entries, so
      // an item is bounded work and is never split).
      const uint32_t inclW = wave_incl_scan(dab, lane);
      const uint32_t exclW = inclW - dab;
      const uint32_t W = __builtin_amdgcn_readlane(inclW, kWave - 1);
      uint32_t c = 0;
      for (uint32_t wb = 0; wb < W; wb += kWave) {
        const uint32_t w = wb + lane;
        const int q = owner_of(inclW, w);
        const uint32_t q_ab0 = __shfl(ab0, q, kWave);
        const uint32_t q_ex = __shfl(exclW, q, kWave);
        const uint32_t q_ac0 = __shfl(ac0, q, kWave);
        const uint32_t q_dac = __shfl(dac, q, kWave);
        if (w < W) {
          const uint32_t v = ab_c[q_ab0 + (w - q_ex)];
          uint32_t b0, b1, y[4];
          bool in_slot = false;
          if (!kAll && S.slots) {
            // one 16-byte read: the count and the entries of a short bc(v)
            const u32x4 sl = ((gptr<u32x4>)S.slots)[BC.rp + v];
            if (sl.x <= kSlotEntries) {
              in_slot = true;
              b0 = 0;
              b1 = sl.x;
              y[0] = sl.x > 0 ? sl.y : 0xFFFFFFFEu;
              y[1] = sl.x > 1 ? sl.z : 0xFFFFFFFEu;
              y[2] = sl.x > 2 ? sl.w : 0xFFFFFFFEu;
              y[3] = 0xFFFFFFFEu;
            } else {
              b0 = sl.y;
              b1 = sl.y + sl.x;
            }
          } else {
            b0 = bc_rp[v];
            b1 = bc_rp[v + 1];
          }
          if (b0 < b1) {
            uint32_t x[4];
            load4(ac_c, q_ac0, q_ac0 + q_dac, 0xFFFFFFFFu, x);
            if (!in_slot) load4(bc_c, b0, b1, 0xFFFFFFFEu, y);
            // a list held in the slot is one block: it never loads from bc_c
            c += intersect_from(ac_c, q_ac0, q_ac0 + q_dac, bc_c, b0, b1, x, y);
          }
        }
      }
      acc += wave_sum32(c);
      if (S.cycles) acc_cyc += clock64() - t_start;
      ++run;
      continue;
    }
    // rows -> wedges (u, v), v in ab(u); the unit covers wedges [wlo, hi)
    const uint32_t inclW = wave_incl_scan(dab, lane);
    const uint32_t exclW = inclW - dab;
    const uint32_t W = __shfl(inclW, kWave - 1, kWave);
    const uint32_t hi = whi < W ? whi : W;
    const bool active = dab != 0 && exclW < hi && inclW > wlo;
    const uint32_t dact = active ? dac : 0;
    // ac(u) entries of the rows the unit touches
    const uint32_t inclA = wave_incl_scan(dact, lane);
    const uint32_t exclA = inclA - dact;
    const uint32_t A = __shfl(inclA, kWave - 1, kWave);

    unsigned long long cnt = 0;   // wave-uniform (ballot counts)
    unsigned long long lcnt = 0;  // per lane (dense kind)
    uint32_t steps = 0, iters = 0;
    bool stop = false;  // the rest of the unit was spilled

    // Wedges [w_from, w_to) against ac(u) staged at seg (per row) in base:
    // the wave's LDS slice, or the ac tile's columns in HBM (one
    // instantiation each, so every load has a known address space).
    auto wedges = [&](uint32_t w_from, uint32_t w_to, uint32_t seg, auto base) {
      for (uint32_t wb = w_from; wb < w_to; wb += kWave) {
        const uint32_t w = wb + lane;
        const bool valid = w < w_to;
        const int q = owner_of(inclW, w);
        const uint32_t q_ab0 = __shfl(ab0, q, kWave);
        const uint32_t q_ex = __shfl(exclW, q, kWave);
        const uint32_t q_seg = __shfl(seg, q, kWave);
        const uint32_t q_dac = __shfl(dac, q, kWave);
        if (dense) {
          // |ac(u) ∩ bc(v)| = popcount of the AND of the two bitmap rows: one
          // lane per wedge, 16 bytes per step from LDS and from HBM.
          if (valid) {
            const uint32_t v = ab_c[q_ab0 + (w - q_ex)];
            const gptr<u32x4> brow = (gptr<u32x4>)(bc_d + uint64_t(v) * T.wpr);
            const u32x4* arow = reinterpret_cast<const u32x4*>(lds + q_seg);
            // four 16-byte loads of the bitmap row in flight per lane (half a
            // line: a line read a quarter at a time is often evicted by the
            // other lanes' rows before its next quarter is read); the LDS
            // words are read two at a time, so the registers of the four
            // loads stay within the kernel's budget for 4 waves per SIMD
            uint32_t c = 0;
            const uint32_t n4 = T.wpr / 4;
            uint32_t j = 0;
            for (; j + 4 <= n4; j += 4) {
              const u32x4 b0 = brow[j], b1 = brow[j + 1], b2 = brow[j + 2], b3 = brow[j + 3];
              __builtin_amdgcn_sched_barrier(0);
              c += and_popc(arow[j], b0);
              c += and_popc(arow[j + 1], b1);
              __builtin_amdgcn_sched_barrier(0);
              c += and_popc(arow[j + 2], b2);
              c += and_popc(arow[j + 3], b3);
            }
            for (; j < n4; ++j) c += and_popc(arow[j], brow[j]);
            lcnt += c;
          }
          steps += 1 + T.wpr / 4;
        } else {
          // the list walked per wedge: bc(v) (search, bitmap) or ac(u) (rdense)
          uint32_t b0 = 0, dlen = 0, v = 0;
          if (valid) {
            v = ab_c[q_ab0 + (w - q_ex)];
            if (rdense) {
              dlen = q_dac;
            } else {
              b0 = bc_rp[v];
              dlen = bc_rp[v + 1] - b0;
            }
          }
          const uint32_t inclP = wave_incl_scan(dlen, lane);
          const uint32_t exclP = inclP - dlen;
          const uint32_t Pn = __builtin_amdgcn_readlane(inclP, kWave - 1);
          // entry p of this lane's wedge is list[my_d + p]; my_s locates the
          // structure it is tested against
          const uint32_t my_d = (rdense ? q_seg : b0) - exclP;
          const uint32_t my_s = rdense ? v * T.wpr : q_seg;
          // Each 64-entry step: wedges starting inside the step post their
          // data at their start position; a lane's wedge is the last start at
          // or before it (ballot + count of leading zeros), else the wedge
          // carried over from the previous step.
          uint32_t c_d = 0, c_s = 0, c_n = 0;
          for (uint32_t pb = 0; pb < Pn; pb += kWave) {
            mk_f[lane] = 0;
            wave_lds_sync();
            if (dlen != 0 && exclP >= pb && exclP < pb + kWave) {
              const uint32_t st = exclP - pb;
              mk_f[st] = 1;
              mk_d[st] = my_d;
              mk_s[st] = my_s;
              mk_n[st] = q_dac;
            }
            wave_lds_sync();
            const uint64_t upto = __ballot(mk_f[lane] != 0) & ((2ull << lane) - 1);
            uint32_t d = c_d, sg = c_s, nn = c_n;
            if (upto) {
              const int hb = 63 - __clzll(upto);
              d = mk_d[hb];
              sg = mk_s[hb];
              nn = mk_n[hb];
            }
            c_d = __builtin_amdgcn_readlane(d, kWave - 1);
            c_s = __builtin_amdgcn_readlane(sg, kWave - 1);
            c_n = __builtin_amdgcn_readlane(nn, kWave - 1);
            const uint32_t p = pb + lane;
            bool found = false;
            if (p < Pn) {
              if (rdense) {
                const uint32_t x = base[d + p];
                found = (bc_d[sg + (x >> 5)] >> (x & 31)) & 1u;
              } else {
                const uint32_t x = bc_c[d + p];
                if (bitmap) found = (lds[sg + (x >> 5)] >> (x & 31)) & 1u;
                else found = sorted_contains(base + sg, nn, x);
              }
            }
            cnt += __popcll(__ballot(found));
            wave_lds_sync();
          }
          steps += 1 + (Pn + kWave - 1) / kWave;
        }
        ++iters;
        // Spill the rest of a heavy first-phase unit as second-phase chunks
        // of about kChunkSteps steps each, sized from the steps per 64-wedge
        // batch seen so far.
        // first wedge not yet counted (a window may end inside this batch)
        const uint32_t wn = wb + kWave < w_to ? wb + kWave : w_to;
        if (may_defer && steps > a.defer_steps && wn < hi) {
          may_defer = false;
          const uint32_t per = steps / iters;
          uint32_t k = kChunkSteps / (per ? per : 1);
          k = k < 1 ? 1 : (k > 64 ? 64 : k);
          const uint32_t wc = k * kWave;
          const uint32_t nch = (hi - wn + wc - 1) / wc;
          int ok = 0;
          if (lane == 0) {
            unsigned long long old = __hip_atomic_load(&a.ctl->spill, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            for (;;) {
              const unsigned long long r = old >> kSpillShift;
              if (r >= a.spill_cap) break;
              const unsigned long long prev =
                  atomicCAS(&a.ctl->spill, old, old + (1ull << kSpillShift) + nch);
              if (prev == old) {
                SpillRec R;
                R.item = it;
                R.w0 = wn;
                R.w1 = hi;
                R.wc = wc;
                R.pad = 0;
                a.spill[r] = R;
                a.spill_base[r] = old & ((1ull << kSpillShift) - 1);
                if (a.left) atomicAdd(a.left + T.graph, (unsigned long long)nch);
                ok = 1;
                break;
              }
              old = prev;
            }
          }
          if (__shfl(ok, 0, kWave)) {
            stop = true;
            break;
          }
        }
      }
    };

    if (bitmap) {
      // bitmaps of the touched rows (T.rows of them fit the slice)
      for (uint32_t s = 0; s < A; s += kWave) {
        const uint32_t e = s + lane;
        const int q = owner_of(inclA, e);
        const uint32_t q_ac0 = __shfl(ac0, q, kWave);
        const uint32_t q_ex = __shfl(exclA, q, kWave);
        if (e < A) {
          const uint32_t x = ac_c[q_ac0 + (e - q_ex)];
          atomicOr(&lds[uint32_t(q) * T.wpr + (x >> 5)], 1u << (x & 31));
        }
      }
      wave_lds_sync();
      wedges(wlo, hi, uint32_t(lane) * T.wpr, lds);
      // restore the all-zero slice: the bitmap rows of the staged rows (a
      // contiguous lane range), 16 bytes per store, without reading ac(u)
      // from memory again
      wave_lds_sync();
      const uint64_t staged = __ballot(dact != 0);
      if (staged) {
        const uint32_t r0 = uint32_t(__ffsll((long long)staged) - 1);
        const uint32_t r1 = uint32_t(64 - __clzll(staged));
        uint4* z = reinterpret_cast<uint4*>(lds + r0 * T.wpr);
        const uint32_t nz = (r1 - r0) * T.wpr / 4;
        for (uint32_t x = uint32_t(lane); x < nz; x += kWave) z[x] = make_uint4(0, 0, 0, 0);
      }
    } else {
      // Sorted ac(u) lists, a window of consecutive rows at a time: as many
      // rows as fit the slice together, or one row searched in HBM when its
      // list alone is larger than the slice.
      uint32_t a0 = 0, dirty = 0;
      while (a0 < A && !stop) {
        uint64_t win = __ballot(dact != 0 && exclA >= a0 && inclA - a0 <= uint32_t(kLdsWords));
        const bool in_hbm = win == 0;
        if (in_hbm) win = __ballot(dact != 0 && exclA == a0);
        const int first = __ffsll((long long)win) - 1;
        const int last = 63 - __clzll(win);
        const uint32_t a1 = __shfl(inclA, last, kWave);
        if (!in_hbm) {
          for (uint32_t s = a0; s < a1; s += kWave) {
            const uint32_t e = s + lane;
            const int q = owner_of(inclA, e);
            const uint32_t q_ac0 = __shfl(ac0, q, kWave);
            const uint32_t q_ex = __shfl(exclA, q, kWave);
            if (e < a1) lds[e - a0] = ac_c[q_ac0 + (e - q_ex)];
          }
          if (a1 - a0 > dirty) dirty = a1 - a0;
        }
        wave_lds_sync();
        const uint32_t f_ex = __shfl(exclW, first, kWave);
        const uint32_t l_in = __shfl(inclW, last, kWave);
        const uint32_t w_from = f_ex > wlo ? f_ex : wlo, w_to = l_in < hi ? l_in : hi;
        if (in_hbm) wedges(w_from, w_to, ac0, ac_c);
        else wedges(w_from, w_to, exclA - a0, lds);
        wave_lds_sync();
        a0 = a1;
      }
      for (uint32_t e = lane; e < dirty; e += kWave) lds[e] = 0;
    }
    wave_lds_sync();
    if (dense) cnt += wave_sum64(lcnt);
    acc += cnt;
    if (S.cycles) acc_cyc += clock64() - t_start;
    ++run;
  }
  flush();
}

__global__ __launch_bounds__(kBlock) void count_probe(CountArgs a) {
  __shared__ __attribute__((aligned(16))) uint32_t lds_all[kWavesPerBlock * kWaveLds];
  count_units<true>(a, lds_all);
}

// 8 waves per SIMD: the lane form has no LDS and few registers, and it is
// latency-bound, so it keeps the full occupancy.
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(8, 8))) void count_probe_lanes(CountArgs a) {
  count_units<false>(a, nullptr);
}

// Zeroes n words (grid-stride; a fill kernel rather than a memset).
__global__ void fill_zero_u32(uint32_t* __restrict__ p, uint64_t n) {
  const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = 0;
}

// Bitmap rows of the tiles listed in jobs (all-zero beforehand): one wave per
// 64 consecutive bitmap rows of the concatenated jobs, a lane per row; the
// rows' entries are flattened over the lanes (a wave prefix sum), so each
// lane reads one column id and sets its bit (entries of one row may share a
// word, hence the atomic OR).
__global__ __launch_bounds__(kBlock) void build_bitmaps(const uint32_t* __restrict__ rowptr,
                                                        const uint32_t* __restrict__ cols,
                                                        uint32_t* __restrict__ dense,
                                                        const DenseJob* __restrict__ jobs, uint32_t njobs,
                                                        uint64_t total_rows) {
  const int lane = threadIdx.x & (kWave - 1);
  const uint64_t wave = (uint64_t(blockIdx.x) * blockDim.x + threadIdx.x) / kWave;
  const uint64_t nwaves = uint64_t(gridDim.x) * blockDim.x / kWave;
  for (uint64_t r0 = wave * kWave; r0 < total_rows; r0 += nwaves * kWave) {
    const uint64_t row = r0 + uint64_t(lane);
    uint32_t d = 0;
    uint64_t cb = 0, db = 0;  // first column entry, first bitmap word of the row
    if (row < total_rows) {
      uint32_t lo = 0, hi = njobs;  // last job with row0 <= row
      while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (jobs[mid].row0 <= row) lo = mid;
        else hi = mid;
      }
      const DenseJob J = gload(jobs + lo);
      const uint64_t r = row - J.row0;
      const uint32_t b = rowptr[J.rp + r], e = rowptr[J.rp + r + 1];
      d = e - b;
      cb = J.col + b;
      db = J.dense + r * J.dwpr;
    }
    const uint32_t incl = wave_incl_scan(d, lane);
    const uint32_t excl = incl - d;
    const uint32_t D = __shfl(incl, kWave - 1, kWave);
    for (uint32_t x = 0; x < D; x += kWave) {
      const uint32_t e = x + uint32_t(lane);
      const int q = owner_of(incl, e);
      const uint64_t qcb = __shfl(cb, q, kWave);
      const uint64_t qdb = __shfl(db, q, kWave);
      const uint32_t qex = __shfl(excl, q, kWave);
      if (e < D) {
        const uint32_t c = cols[qcb + (e - qex)];
        atomicOr(&dense[qdb + (c >> 5)], 1u << (c & 31));
      }
    }
  }
}

// Row slots (see SegDesc::slots) of the tiles listed in jobs (the dense and
// dwpr fields unused): one thread per row of the concatenated jobs.
__global__ __launch_bounds__(kBlock) void build_slots(const uint32_t* __restrict__ rowptr,
                                                      const uint32_t* __restrict__ cols, u32x4* __restrict__ slots,
                                                      const DenseJob* __restrict__ jobs, uint32_t njobs,
                                                      uint64_t total_rows) {
  const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
  for (uint64_t row = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; row < total_rows; row += stride) {
    uint32_t lo = 0, hi = njobs;  // last job with row0 <= row
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (jobs[mid].row0 <= row) lo = mid;
      else hi = mid;
    }
    const DenseJob J = gload(jobs + lo);
    const uint64_t r = row - J.row0;
    const uint32_t b = rowptr[J.rp + r], e = rowptr[J.rp + r + 1];
    const uint32_t cnt = e - b;
    u32x4 sl = {cnt, b, 0u, 0u};
    if (cnt <= kSlotEntries) {
      const uint32_t* c = cols + J.col + b;
      sl.y = cnt > 0 ? c[0] : 0u;
      sl.z = cnt > 1 ? c[1] : 0u;
      sl.w = cnt > 2 ? c[2] : 0u;
    }
    slots[J.rp + r] = sl;
  }
}

// ---------------------------------------------------------------------------
// Row form. The row table holds N+(v) of every global row v in a slot of
// S = 2^s_shift words (global ids ascending, then kRowPad), gathered from
// the tiles (j, k), k >= j, of v's band j.
// ---------------------------------------------------------------------------

// Band of global row or column id x: the last b < P with cuts[b] <= x.
__device__ __forceinline__ uint32_t band_of(const uint32_t* cuts, uint32_t P, uint32_t x) {
  uint32_t lo = 0, hi = P;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (cuts[mid] <= x) lo = mid;
    else hi = mid;
  }
  return lo;
}

// Last entry of a block of a row slot (its first entry is an entry).
__device__ __forceinline__ uint32_t last_entry(const u32x4& b) {
  return b.w != kRowPad ? b.w : b.z != kRowPad ? b.z : b.y != kRowPad ? b.y : b.x;
}

// Entries of each global row v < nrows over all its tiles; the longest goes
// to *maxlen (atomic max).
__global__ __launch_bounds__(kBlock) void row_lengths(const uint32_t* __restrict__ rowptr,
                                                      const TileDesc* __restrict__ tiles,
                                                      const uint32_t* __restrict__ cuts_g, uint32_t P,
                                                      uint32_t* __restrict__ maxlen) {
  __shared__ uint32_t cuts[kRowMaxCuts + 1];
  for (uint32_t x = threadIdx.x; x <= P; x += kBlock) cuts[x] = cuts_g[x];
  __syncthreads();
  const uint32_t nrows = cuts[P];
  const uint64_t v = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
  uint32_t len = 0;
  if (v < nrows) {
    const uint32_t j = band_of(cuts, P, uint32_t(v));
    const uint32_t r = uint32_t(v) - cuts[j];
    for (uint32_t k = j; k < P; ++k) {
      const TileDesc T = gload(tiles + uint64_t(j) * P + k);
      if (T.nzr) len += rowptr[T.rp + r + 1] - rowptr[T.rp + r];
    }
  }
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) {
    const uint32_t o = __shfl_xor(len, off, kWave);
    len = o > len ? o : len;
  }
  if ((threadIdx.x & (kWave - 1)) == 0 && len) atomicMax(maxlen, len);
}

// Writes the row table (S = 2^s_shift words per row, S >= every row's
// length): one thread per global row, 16 bytes per store.
__global__ __launch_bounds__(kBlock) void build_rows(const uint32_t* __restrict__ rowptr,
                                                     const uint32_t* __restrict__ cols,
                                                     const TileDesc* __restrict__ tiles,
                                                     const uint32_t* __restrict__ cuts_g, uint32_t P,
                                                     uint32_t s_shift, u32x4* __restrict__ rows,
                                                     uint32_t* __restrict__ rowlen) {
  __shared__ uint32_t cuts[kRowMaxCuts + 1];
  for (uint32_t x = threadIdx.x; x <= P; x += kBlock) cuts[x] = cuts_g[x];
  __syncthreads();
  const uint32_t nrows = cuts[P];
  const uint64_t v = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (v >= nrows) return;
  const uint32_t j = band_of(cuts, P, uint32_t(v));
  const uint32_t r = uint32_t(v) - cuts[j];
  u32x4* out = rows + (v << (s_shift - 2));
  const uint32_t nblk = 1u << (s_shift - 2);
  uint32_t buf[4] = {kRowPad, kRowPad, kRowPad, kRowPad};
  uint32_t fill = 0, blk = 0;
  for (uint32_t k = j; k < P; ++k) {
    const TileDesc T = gload(tiles + uint64_t(j) * P + k);
    if (!T.nzr) continue;
    const uint32_t b = rowptr[T.rp + r], e = rowptr[T.rp + r + 1];
    for (uint32_t x = b; x < e; ++x) {
      const uint32_t c = cols[T.col + x] + cuts[k];
#pragma unroll
      for (int q = 0; q < 4; ++q)
        if (uint32_t(q) == fill) buf[q] = c;
      if (++fill == 4) {
        out[blk++] = u32x4{buf[0], buf[1], buf[2], buf[3]};
        fill = 0;
      }
    }
  }
  if (fill) {
#pragma unroll
    for (int q = 0; q < 4; ++q)
      if (uint32_t(q) >= fill) buf[q] = kRowPad;
    out[blk++] = u32x4{buf[0], buf[1], buf[2], buf[3]};
  }
  rowlen[v] = blk * 4 - (fill ? 4 - fill : 0);
  for (; blk < nblk; ++blk) out[blk] = u32x4{kRowPad, kRowPad, kRowPad, kRowPad};
}

// The row form's count. An item is 64 consecutive global rows u, claimed
// like the first phase of count_probe from the launch's eight parts
// (contiguous eighths of the rows), kRowItemBatch items per atomic. The wave flattens the rows' entries into wedges (u, v)
// with a prefix sum, so a lane takes one wedge: v is entry x of u's slot,
// and the entries after it are merged with v's slot 16 bytes at a time (all
// 16 pairs of the two current blocks compared, the block with the smaller
// last entry moving on). Each common entry w adds one to task (band(u),
// band(v), band(w)); a lane keeps one running count per band (w ascends, so
// bands do not decrease) and adds it to the task's counter when the band
// changes and at the end of the wedge.
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(8, 8))) void count_rows(RowArgs a) {
  __shared__ uint32_t cuts[kRowMaxCuts + 1];
  for (uint32_t x = threadIdx.x; x <= a.P; x += kBlock) cuts[x] = a.cuts[x];
  __syncthreads();
  const int lane = threadIdx.x & (kWave - 1);
  const uint32_t nblk = 1u << (a.s_shift - 2);
  const gptr<u32x4> rows = (gptr<u32x4>)a.rows;
  const gptr<uint32_t> words = as_global(a.rows);

  unsigned long long item = 0, batch_end = 0;
  uint32_t part = blockIdx.x % kQueues, parts_left = kQueues;
  const unsigned long long base = a.ctl->base, chunk = a.ctl->chunk;
  const unsigned shift = unsigned(a.ctl->shift);
  for (;;) {
    if (item >= batch_end) {
      unsigned long long old = 0, pe = 0;
      while (parts_left) {  // wave-uniform loop; lane 0 does the atomic
        pe = a.ctl->end[part];
        unsigned long long got = 0;
        if (lane == 0) got = atomicAdd(&a.ctl->head[part * kCtlStride], (unsigned long long)kRowItemBatch);
        old = uniform64(__shfl(got, 0, kWave));
        if (old < pe) break;
        part = part + 1 == kQueues ? 0 : part + 1;
        --parts_left;
      }
      if (!parts_left) break;
      item = part_item(base, part, old, chunk, shift);
      batch_end = item + ((old + kRowItemBatch < pe ? old + kRowItemBatch : pe) - old);
    }
    const uint64_t u = (item++) * kWave + uint64_t(lane);
    uint32_t d = 0, bu = 0;
    if (u < a.nrows) {
      d = as_global(a.rowlen)[u];
      if (d) bu = band_of(cuts, a.P, uint32_t(u));
    }
    const uint32_t incl = wave_incl_scan(d, lane);
    const uint32_t excl = incl - d;
    const uint32_t W = __builtin_amdgcn_readlane(incl, kWave - 1);
    for (uint32_t wb = 0; wb < W; wb += kWave) {
      const uint32_t w = wb + lane;
      const int q = owner_of(incl, w);
      const uint32_t q_ex = __shfl(excl, q, kWave);
      const uint32_t q_d = __shfl(d, q, kWave);
      const uint32_t q_bu = __shfl(bu, q, kWave);
      const uint64_t q_u = uint64_t(__shfl(uint32_t(u), q, kWave));
      // v is entry x0 - 1 of u's slot; a v that is u's last entry has no w
      // after it
      const uint32_t x0 = w - q_ex + 1;
      if (w >= W || x0 == q_d) continue;
      const uint64_t xs = q_u << (a.s_shift - 2);
      const uint32_t v = words[(q_u << a.s_shift) + x0 - 1];
      const uint64_t ys = uint64_t(v) << (a.s_shift - 2);
      // N+(v) is read four blocks (64 bytes) at a time, all four loads in
      // flight together: a slot's blocks share one cache line, which a later
      // load of the same line could find evicted by the other lanes' lines.
      // u's entries after v are read a block at a time (u's last block is
      // known from its length).
      const u32x4 pads = {kRowPad, kRowPad, kRowPad, kRowPad};
      u32x4 Y[4];
#pragma unroll
      for (int yb = 0; yb < 4; ++yb) Y[yb] = uint32_t(yb) < nblk ? rows[ys + yb] : pads;
      uint32_t bx = x0 >> 2;
      const uint32_t ex = (q_d - 1) >> 2;
      u32x4 x = rows[xs + bx];
      if (Y[0].x == kRowPad) continue;  // N+(v) is empty
      // count cc of common entries in band kc not yet added to its task;
      // the task's slot (through band(v)) is looked up only when there is a
      // count to add, which is rare
      uint32_t kc = 0, cc = 0;
      auto add = [&]() {
        const uint32_t bv = band_of(cuts, a.P, v);
        // in an upper-triangular graph every common entry w > v >= cuts[bv],
        // so kc >= bv
        const uint32_t t = kc >= bv ? as_global(a.lex2plan)[lex_task(a.P, q_bu, bv, kc)] : ~0u;
        if (t != ~0u) atomicAdd(a.per_task + t, (unsigned long long)cc);
        cc = 0;
      };
      uint32_t lx = last_entry(x);
      bool done = false;
      for (uint32_t r = 0;;) {
#pragma unroll
        for (int yb = 0; yb < 4; ++yb) {
          const u32x4 y = Y[yb];
          if (done) break;
          if (y.x == kRowPad) {
            done = true;
            break;
          }
          const uint32_t ly = last_entry(y);
          for (;;) {
            // x entries equal to an entry of y (a pad matches only a pad,
            // which is excluded below)
            const uint32_t m = ((x.x == y.x) | (x.x == y.y) | (x.x == y.z) | (x.x == y.w)) |
                               ((x.y == y.x) | (x.y == y.y) | (x.y == y.z) | (x.y == y.w)) << 1 |
                               ((x.z == y.x) | (x.z == y.y) | (x.z == y.z) | (x.z == y.w)) << 2 |
                               ((x.w == y.x) | (x.w == y.y) | (x.w == y.z) | (x.w == y.w)) << 3;
            if (m) {
              const uint32_t xv[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
              for (int e = 0; e < 4; ++e) {
                const uint32_t c = xv[e];
                if (((m >> e) & 1u) && c != kRowPad) {
                  const uint32_t k = band_of(cuts, a.P, c);
                  if (k != kc && cc) add();
                  kc = k;
                  ++cc;
                }
              }
            }
            // the block whose last entry is smaller moves on (both when
            // equal); a list whose last block moves on is done
            const bool na = lx <= ly, nb = ly <= lx;
            if (na) {
              if (bx == ex) {
                done = true;
                break;
              }
              x = rows[xs + ++bx];
              lx = last_entry(x);
            }
            if (nb) {
              if (y.w == kRowPad) done = true;
              break;
            }
          }
        }
        r += 4;
        if (done || r >= nblk) break;
#pragma unroll
        for (int yb = 0; yb < 4; ++yb) Y[yb] = r + yb < nblk ? rows[ys + r + yb] : pads;
      }
      if (cc) add();
    }
  }
}

// Start of a batch launch: stamp the device wall clock and reset each
// graph's outstanding item count (graphs without items finish at once).
__global__ void batch_begin(const unsigned long long* __restrict__ left0, uint64_t ngraphs,
                            unsigned long long* __restrict__ left, unsigned long long* __restrict__ done,
                            unsigned long long* __restrict__ start) {
  const uint64_t g = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const unsigned long long now = wall_clock64();
  if (g == 0) *start = now;
  if (g < ngraphs) {
    left[g] = left0[g];
    done[g] = left0[g] ? 0ull : now;
  }
}

// Streaming read of n 16-byte words: the achievable HBM read bandwidth that
// the counting kernels are compared with. Each block reads one contiguous
// range, each lane four independent 16-byte nontemporal loads per pass (the
// fastest of the shapes measured on MI355X: grid-stride or blocked ranges,
// 2/4/8 loads in flight, cached or nontemporal, 1K-32K blocks). The XOR of
// everything read decides a store that the data never triggers (the buffer
// holds zeros), so no load can be dropped.
__global__ __launch_bounds__(kBlock) void stream_read(const u32x4* __restrict__ words, uint64_t n,
                                                      unsigned* __restrict__ sink) {
  const gptr<u32x4> p = as_global(words);
  const uint64_t per = (n + gridDim.x - 1) / gridDim.x;
  const uint64_t b0 = uint64_t(blockIdx.x) * per;
  const uint64_t b1 = b0 + per < n ? b0 + per : n;
  uint64_t i = b0 + threadIdx.x;
  u32x4 acc = {0u, 0u, 0u, 0u};
  for (; i + 3 * kBlock < b1; i += 4 * kBlock) {
    const u32x4 w0 = __builtin_nontemporal_load(&p[i]);
    const u32x4 w1 = __builtin_nontemporal_load(&p[i + kBlock]);
    const u32x4 w2 = __builtin_nontemporal_load(&p[i + 2 * kBlock]);
    const u32x4 w3 = __builtin_nontemporal_load(&p[i + 3 * kBlock]);
    acc ^= w0 ^ w1 ^ w2 ^ w3;
  }
  for (; i < b1; i += kBlock) acc ^= p[i];
  const unsigned x = acc.x ^ acc.y ^ acc.z ^ acc.w;
  if (x == 0x9E3779B9u) sink[blockIdx.x] = x;
}

}  // namespace bbtc
