// The block CSR layout built on the GPU from a CSR graph, and the work
// statistics of tasks computed on the GPU.
//
// Same arrays as host/layout.cpp (build_layout), bit for bit, made by the
// same four passes as data-parallel kernels over a device copy of the CSR:
//   1. row counts: one thread per row walks its sorted columns band by band
//      and writes the entry count of (tile, local row) into the tile's row
//      offsets (the reference's count pass, block_csr.cpp:297-323);
//   2. per-tile inclusive scan of the row offsets, in chunks of 4,096 offsets:
//      chunk sums, a scan of the chunk sums within each tile, then each chunk
//      rescans itself from its tile-local offset (block_csr.cpp:323-351),
//      writing the tile's non-empty row list on the way;
//   3. the host places the tiles from their entry and non-empty row counts
//      (the only device-to-host copy: three numbers per tile);
//   4. column scatter: one thread per row writes its tile-local columns
//      (block_csr.cpp:353-376).
// The layout then stays resident (the same device graph bbtc_graph_upload
// makes); its host arrays are copied back only when asked for.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "internal.hpp"

namespace bbtc {
void set_last_error(const std::string& s);

namespace {

constexpr uint32_t kScanChunk = 4096;             // row offsets per scan chunk
constexpr uint32_t kScanPerThread = kScanChunk / kBlock;
static_assert(kScanPerThread * kBlock == kScanChunk, "scan chunk is a whole number of per-thread runs");

struct ScanChunk {
  uint64_t start;  // first row offset of the chunk (index into rowptr)
  uint32_t len;    // offsets in the chunk (1..kScanChunk)
  uint32_t tile;   // tile id i * P + k
};
static_assert(sizeof(ScanChunk) == 16, "ScanChunk layout");

// Work statistics job: the three tile ids of a task that the reference would
// not skip, 128-byte lines of a bitmap row of its band k, and its output row.
struct StatJob {
  uint32_t ab, bc, ac, lines;
  uint64_t out;
};
static_assert(sizeof(StatJob) == 24, "StatJob layout");

// Band of a vertex: the last b with cuts[b] <= v (the non-empty band that
// holds it); v < cuts[P].
__device__ inline uint32_t band_of(const uint32_t* __restrict__ cuts, uint32_t P, uint32_t v) {
  uint32_t lo = 0, hi = P + 1;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) / 2;
    if (cuts[mid] <= v) lo = mid + 1;
    else hi = mid;
  }
  return lo - 1;
}

__global__ void fill_u32(uint32_t* __restrict__ p, uint64_t n) {
  for (uint64_t x = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; x < n; x += uint64_t(gridDim.x) * blockDim.x)
    p[x] = 0;
}

// Pass 1: rp[tile.rp + r + 1] = entries of local row r in tile (i, k).
__global__ __launch_bounds__(kBlock) void row_counts(const uint64_t* __restrict__ off, const uint32_t* __restrict__ col,
                                                     uint32_t n, const uint32_t* __restrict__ cuts, uint32_t P,
                                                     const TileDesc* __restrict__ tiles, uint32_t* __restrict__ rp) {
  for (uint64_t u = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; u < n; u += uint64_t(gridDim.x) * blockDim.x) {
    const uint64_t b = off[u], e = off[u + 1];
    if (b == e) continue;
    const uint32_t i = band_of(cuts, P, uint32_t(u));
    const uint32_t r = uint32_t(u) - cuts[i];
    uint32_t k = band_of(cuts, P, col[b]);
    uint64_t x = b;
    while (x < e) {
      const uint32_t hi = cuts[k + 1];
      uint64_t y = x;
      while (y < e && col[y] < hi) ++y;
      if (y > x) rp[tiles[size_t(i) * P + k].rp + r + 1] = uint32_t(y - x);
      x = y;
      ++k;
    }
  }
}

// Block-wide exclusive scan of one value per thread (256 threads) through LDS.
template <class T>
__device__ inline T block_exclusive_scan(T v, T* lds) {
  const uint32_t t = threadIdx.x;
  lds[t] = v;
  __syncthreads();
  for (uint32_t d = 1; d < kBlock; d <<= 1) {
    const T add = t >= d ? lds[t - d] : T(0);
    __syncthreads();
    lds[t] += add;
    __syncthreads();
  }
  const T incl = lds[t];
  __syncthreads();
  return incl - v;
}

// Pass 2a: entry sum, non-zero offsets and largest offset of every chunk.
__global__ __launch_bounds__(kBlock) void chunk_sums(const uint32_t* __restrict__ rp, const ScanChunk* __restrict__ ch,
                                                     uint64_t nch, uint64_t* __restrict__ sum,
                                                     uint32_t* __restrict__ nz, uint32_t* __restrict__ mx) {
  __shared__ uint64_t s_sum[kBlock];
  __shared__ uint32_t s_nz[kBlock];
  __shared__ uint32_t s_mx[kBlock];
  for (uint64_t c = blockIdx.x; c < nch; c += gridDim.x) {
    const ScanChunk C = ch[c];
    uint64_t a = 0;
    uint32_t z = 0, m = 0;
    for (uint32_t e = threadIdx.x; e < C.len; e += kBlock) {
      const uint32_t v = rp[C.start + e];
      a += v;
      z += v > 0 ? 1u : 0u;
      m = max(m, v);
    }
    s_sum[threadIdx.x] = a;
    s_nz[threadIdx.x] = z;
    s_mx[threadIdx.x] = m;
    __syncthreads();
    for (uint32_t d = kBlock / 2; d > 0; d >>= 1) {
      if (threadIdx.x < d) {
        s_sum[threadIdx.x] += s_sum[threadIdx.x + d];
        s_nz[threadIdx.x] += s_nz[threadIdx.x + d];
        s_mx[threadIdx.x] = max(s_mx[threadIdx.x], s_mx[threadIdx.x + d]);
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      sum[c] = s_sum[0];
      nz[c] = s_nz[0];
      mx[c] = s_mx[0];
    }
    __syncthreads();
  }
}

// Pass 2b: one thread per upper tile scans its chunks' sums: each chunk's
// tile-local entry and non-empty row offsets, and the tile's totals.
__global__ void tile_scan(const ScanChunk* __restrict__ ch, const uint64_t* __restrict__ first, uint64_t ntiles,
                          const uint64_t* __restrict__ sum, const uint32_t* __restrict__ nz,
                          const uint32_t* __restrict__ mx, uint64_t* __restrict__ coff, uint32_t* __restrict__ cnz,
                          uint64_t* __restrict__ tnnz, uint32_t* __restrict__ tnzr, uint32_t* __restrict__ tmax) {
  const uint64_t q = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (q >= ntiles) return;
  uint64_t a = 0;
  uint32_t z = 0, m = 0;
  for (uint64_t c = first[q]; c < first[q + 1]; ++c) {
    coff[c] = a;
    cnz[c] = z;
    a += sum[c];
    z += nz[c];
    m = max(m, mx[c]);
  }
  const uint32_t t = ch[first[q]].tile;
  tnnz[t] = a;
  tnzr[t] = z;
  tmax[t] = m;
}

// Pass 2c: each chunk rescans its offsets from its tile-local entry offset
// (in place) and writes the local rows of its non-zero counts into the
// tile's non-empty row list. Offset x of a tile counts row x - 1.
__global__ __launch_bounds__(kBlock) void chunk_write(uint32_t* __restrict__ rp, const ScanChunk* __restrict__ ch,
                                                      uint64_t nch, const uint64_t* __restrict__ coff,
                                                      const uint32_t* __restrict__ cnz,
                                                      const TileDesc* __restrict__ tiles,
                                                      uint32_t* __restrict__ nzrows) {
  __shared__ uint64_t s_a[kBlock];
  __shared__ uint32_t s_z[kBlock];
  for (uint64_t c = blockIdx.x; c < nch; c += gridDim.x) {
    const ScanChunk C = ch[c];
    const TileDesc T = tiles[C.tile];
    const uint64_t local0 = C.start - T.rp;
    const uint32_t e0 = threadIdx.x * kScanPerThread;
    uint32_t v[kScanPerThread];
    uint64_t a = 0;
    uint32_t z = 0;
#pragma unroll
    for (uint32_t j = 0; j < kScanPerThread; ++j) {
      v[j] = e0 + j < C.len ? rp[C.start + e0 + j] : 0u;
      a += v[j];
      z += v[j] > 0 ? 1u : 0u;
    }
    uint64_t run = coff[c] + block_exclusive_scan<uint64_t>(a, s_a);
    uint32_t zr = cnz[c] + block_exclusive_scan<uint32_t>(z, s_z);
#pragma unroll
    for (uint32_t j = 0; j < kScanPerThread; ++j) {
      if (e0 + j < C.len) {
        run += v[j];
        rp[C.start + e0 + j] = uint32_t(run);
        if (v[j] > 0) nzrows[T.nz + zr++] = uint32_t(local0 + e0 + j - 1);
      }
    }
  }
}

// Pass 4: tile-local columns at their rows' offsets.
__global__ __launch_bounds__(kBlock) void scatter_cols(const uint64_t* __restrict__ off,
                                                       const uint32_t* __restrict__ col, uint32_t n,
                                                       const uint32_t* __restrict__ cuts, uint32_t P,
                                                       const TileDesc* __restrict__ tiles,
                                                       const uint32_t* __restrict__ rp, uint32_t* __restrict__ cols) {
  for (uint64_t u = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; u < n; u += uint64_t(gridDim.x) * blockDim.x) {
    const uint64_t b = off[u], e = off[u + 1];
    if (b == e) continue;
    const uint32_t i = band_of(cuts, P, uint32_t(u));
    const uint32_t r = uint32_t(u) - cuts[i];
    uint32_t k = band_of(cuts, P, col[b]);
    uint64_t x = b;
    while (x < e) {
      const uint32_t lo = cuts[k], hi = cuts[k + 1];
      const TileDesc& T = tiles[size_t(i) * P + k];
      uint32_t* dst = cols + T.col + rp[T.rp + r];
      while (x < e && col[x] < hi) *dst++ = col[x++] - lo;
      ++k;
    }
  }
}

// Work statistics of one task per block (layout.cpp task_stats): units of 64
// consecutive non-empty ab rows, at most sample_units of them spread evenly
// over the list, one row per lane; the integer sums are exact, and the
// scaling to all rows is the same double arithmetic as the host's.
__global__ __launch_bounds__(kBlock) void task_stats(const TileDesc* __restrict__ tiles,
                                                     const uint32_t* __restrict__ rowptr,
                                                     const uint32_t* __restrict__ cols,
                                                     const uint32_t* __restrict__ nzrows,
                                                     const StatJob* __restrict__ jobs, uint64_t njobs,
                                                     uint32_t sample_units, uint32_t unit_rows,
                                                     uint64_t* __restrict__ out) {
  __shared__ unsigned long long s[6];
  for (uint64_t q = blockIdx.x; q < njobs; q += gridDim.x) {
    const StatJob J = jobs[q];
    const TileDesc AB = tiles[J.ab], BC = tiles[J.bc], AC = tiles[J.ac];
    if (threadIdx.x < 6) s[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t nzr = AB.nzr;
    const uint64_t units = (uint64_t(nzr) + unit_rows - 1) / unit_rows;
    const uint64_t take = sample_units && units > sample_units ? sample_units : units;
    // rows of the ab list this task reads: unit m is rows [m * unit_rows, ...)
    // of the list (every unit, or `take` spread evenly)
    const uint64_t slots = take * unit_rows;
    uint64_t r_acc = 0, w_acc = 0, a_acc = 0, pbc_acc = 0, pac_acc = 0, lac_acc = 0;
    for (uint64_t x = threadIdx.x; x < slots; x += kBlock) {
      const uint64_t m = x / unit_rows;
      const uint64_t unit = take == units ? m : (2 * m + 1) * units / (2 * take);
      const uint64_t z = unit * unit_rows + x % unit_rows;
      if (z >= nzr) continue;
      const uint32_t r = nzrows[AB.nz + z];
      const uint64_t dac = rowptr[AC.rp + r + 1] - rowptr[AC.rp + r];
      if (!dac) continue;
      const uint32_t a0 = rowptr[AB.rp + r], a1 = rowptr[AB.rp + r + 1];
      uint64_t pbc = 0;
      for (uint32_t y = a0; y < a1; ++y) {
        const uint32_t v = cols[AB.col + y];
        pbc += rowptr[BC.rp + v + 1] - rowptr[BC.rp + v];
      }
      const uint64_t dab = a1 - a0;
      r_acc += 1;
      w_acc += dab;
      a_acc += dac;
      pbc_acc += pbc;
      pac_acc += dab * dac;
      lac_acc += dab * min<uint64_t>(dac, J.lines);
    }
    atomicAdd(&s[0], (unsigned long long)r_acc);
    atomicAdd(&s[1], (unsigned long long)w_acc);
    atomicAdd(&s[2], (unsigned long long)a_acc);
    atomicAdd(&s[3], (unsigned long long)pbc_acc);
    atomicAdd(&s[4], (unsigned long long)pac_acc);
    atomicAdd(&s[5], (unsigned long long)lac_acc);
    __syncthreads();
    if (threadIdx.x == 0) {
      uint64_t seen = 0;
      for (uint64_t m = 0; m < take; ++m) {
        const uint64_t unit = take == units ? m : (2 * m + 1) * units / (2 * take);
        const uint64_t z0 = unit * unit_rows;
        const uint64_t z1 = min<uint64_t>(nzr, z0 + unit_rows);
        seen += z1 - z0;
      }
      const double f = seen ? double(nzr) / double(seen) : 0.0;
      for (int w = 0; w < 6; ++w) out[6 * J.out + w] = uint64_t(llround(double(s[w]) * f));
    }
    __syncthreads();
  }
}

unsigned grid_for(uint64_t threads, uint64_t cap = 65536) {
  return unsigned(std::max<uint64_t>(1, std::min<uint64_t>((threads + kBlock - 1) / kBlock, cap)));
}

// Temporary device buffers of one build, freed on every exit.
struct Temps {
  bbtc_ctx* ctx;
  uint64_t bytes = 0;
  std::vector<void*> p;
  explicit Temps(bbtc_ctx* c) : ctx(c) {}
  template <class T>
  T* get(uint64_t n) {
    T* x = dev_alloc<T>(ctx, n, &bytes);
    p.push_back(x);
    return x;
  }
  ~Temps() {
    (void)hipStreamSynchronize(ctx->stream);
    for (void* x : p) dev_free(ctx, x);
    ctx->in_use -= std::min(ctx->in_use, bytes);
  }
};

// Device to host through two pinned staging buffers: the copy of one
// chunk overlaps the spread of the previous into the (pageable, possibly
// untouched) destination by all host threads.
void copy_to_host(hipStream_t s, void* dst, const void* src, uint64_t bytes) {
  constexpr uint64_t kStage = 64ull << 20;
  if (!bytes) return;
  void* stage[2] = {nullptr, nullptr};
  hipEvent_t ev[2] = {nullptr, nullptr};
  auto cleanup = [&]() {
    for (int b = 0; b < 2; ++b) {
      if (ev[b]) (void)hipEventDestroy(ev[b]);
      if (stage[b]) (void)hipHostFree(stage[b]);
    }
  };
  try {
    for (int b = 0; b < 2; ++b) {
      HIP_OK(hipHostMalloc(&stage[b], kStage, hipHostMallocDefault));
      HIP_OK(hipEventCreateWithFlags(&ev[b], hipEventDisableTiming));
    }
    const uint64_t nchunks = (bytes + kStage - 1) / kStage;
    auto issue = [&](uint64_t c) {
      const uint64_t o = c * kStage;
      HIP_OK(hipMemcpyAsync(stage[c & 1], static_cast<const char*>(src) + o, std::min(kStage, bytes - o),
                            hipMemcpyDeviceToHost, s));
      HIP_OK(hipEventRecord(ev[c & 1], s));
    };
    issue(0);
    for (uint64_t c = 0; c < nchunks; ++c) {
      // the other buffer's chunk (c - 1) was spread in the previous iteration
      if (c + 1 < nchunks) issue(c + 1);
      HIP_OK(hipEventSynchronize(ev[c & 1]));
      const uint64_t o = c * kStage;
      parallel_copy(static_cast<char*>(dst) + o, stage[c & 1], std::min(kStage, bytes - o));
    }
  } catch (...) {
    (void)hipStreamSynchronize(s);
    cleanup();
    throw;
  }
  cleanup();
}

}  // namespace

// Statistics of tasks on a device graph; out[6 q .. 6 q + 6) as task_stats.
void device_task_stats(bbtc_dgraph* g, const std::vector<float>& deg, const bbtc_task* tasks, uint64_t n,
                       uint32_t sample_units, uint64_t* out) {
  bbtc_ctx* ctx = g->ctx;
  const uint32_t P = g->P;
  std::fill(out, out + n * kTaskStats, uint64_t(0));
  std::vector<StatJob> jobs;
  jobs.reserve(n);
  for (uint64_t q = 0; q < n; ++q) {
    const bbtc_task& x = tasks[q];
    const uint32_t ab = x.i * P + x.j, bc = x.j * P + x.k, ac = x.i * P + x.k;
    // bb_tc.cpp:86-91: a tile of average degree 0 skips the task (float compare)
    if (deg[bc] == 0 || deg[ac] == 0 || deg[ab] == 0) continue;
    jobs.push_back(StatJob{ab, bc, ac, (g->dwpr[x.k] + 31) / 32, q});
  }
  if (jobs.empty()) return;
  Temps tmp(ctx);
  StatJob* jd = tmp.get<StatJob>(jobs.size());
  uint64_t* od = tmp.get<uint64_t>(n * kTaskStats);
  HIP_OK(hipMemcpyAsync(jd, jobs.data(), jobs.size() * sizeof(StatJob), hipMemcpyHostToDevice, ctx->stream));
  hipLaunchKernelGGL(task_stats, dim3(unsigned(std::min<uint64_t>(jobs.size(), 65536))), dim3(kBlock), 0,
                     ctx->stream, g->tiles, g->rowptr, g->cols, g->nzrows, jd, uint64_t(jobs.size()), sample_units,
                     kStatUnitRows, od);
  HIP_OK(hipGetLastError());
  std::vector<uint64_t> got(n * kTaskStats);
  HIP_OK(hipMemcpyAsync(got.data(), od, got.size() * 8, hipMemcpyDeviceToHost, ctx->stream));
  HIP_OK(hipStreamSynchronize(ctx->stream));
  for (const StatJob& j : jobs)
    for (int w = 0; w < kTaskStats; ++w) out[kTaskStats * j.out + w] = got[kTaskStats * j.out + w];
}

// The device graph and host layout of g cut at `cuts`, built on the device.
void build_on_device(bbtc_ctx* ctx, const Csr& g, const std::vector<uint32_t>& cuts, bool host_arrays,
                     Layout& L, bbtc_dgraph*& dg_out) {
  L = layout_frame(g.n, g.m, cuts);
  L.longest_row = longest_row(g);
  L.longest_known = true;
  const uint32_t P = L.P;
  if (column_end(g) > cuts[P]) fail(BBTC_ERR_INVALID, "a column lies at or beyond cuts[P]");
  if (uint64_t(P) * P >= (1ull << 32)) fail(BBTC_ERR_INVALID, "too many cuts");
  auto* dg = new bbtc_dgraph;
  dg->ctx = ctx;
  dg->P = P;
  try {
    hipStream_t s = ctx->stream;
    const size_t T = size_t(P) * P;
    // scan chunks of every upper tile's row offsets, tiles in (i, k) order
    std::vector<ScanChunk> chunks;
    std::vector<uint64_t> first;
    for (uint32_t i = 0; i < P; ++i)
      for (uint32_t k = i; k < P; ++k) {
        first.push_back(chunks.size());
        const uint64_t len = uint64_t(L.height(i)) + 1, base = L.rp_base[L.tile(i, k)];
        for (uint64_t o = 0; o < len; o += kScanChunk)
          chunks.push_back(ScanChunk{base + o, uint32_t(std::min<uint64_t>(kScanChunk, len - o)), i * P + k});
      }
    first.push_back(chunks.size());
    const uint64_t nch = chunks.size(), ntiles = first.size() - 1;
    {
      Temps tmp(ctx);
      uint64_t* off = tmp.get<uint64_t>(uint64_t(g.n) + 1);
      uint32_t* col = tmp.get<uint32_t>(std::max<uint64_t>(g.m, 1));
      uint32_t* cuts_d = tmp.get<uint32_t>(P + 1);
      ScanChunk* ch = tmp.get<ScanChunk>(nch);
      uint64_t* first_d = tmp.get<uint64_t>(ntiles + 1);
      uint64_t* csum = tmp.get<uint64_t>(nch);
      uint32_t* cnzc = tmp.get<uint32_t>(nch);
      uint32_t* cmx = tmp.get<uint32_t>(nch);
      uint64_t* coff = tmp.get<uint64_t>(nch);
      uint32_t* cnz = tmp.get<uint32_t>(nch);
      uint64_t* tnnz = tmp.get<uint64_t>(T);
      uint32_t* tnzr = tmp.get<uint32_t>(T);
      uint32_t* tmax = tmp.get<uint32_t>(T);
      dg->rowptr = dev_alloc<uint32_t>(ctx, L.rp_words, &dg->bytes);
      dg->rowptr_words = L.rp_words;
      dg->tiles = dev_alloc<TileDesc>(ctx, T, &dg->bytes);
      copy_to_device(ctx, s, off, g.off.data(), (uint64_t(g.n) + 1) * 8);
      copy_to_device(ctx, s, col, g.col.data(), g.m * 4);
      HIP_OK(hipMemcpyAsync(cuts_d, cuts.data(), (P + 1) * 4, hipMemcpyHostToDevice, s));
      HIP_OK(hipMemcpyAsync(ch, chunks.data(), nch * sizeof(ScanChunk), hipMemcpyHostToDevice, s));
      HIP_OK(hipMemcpyAsync(first_d, first.data(), (ntiles + 1) * 8, hipMemcpyHostToDevice, s));
      // row offset bases (the rest of the table is placed after pass 2)
      std::vector<TileDesc> t0(T, TileDesc{});
      for (uint32_t i = 0; i < P; ++i)
        for (uint32_t k = i; k < P; ++k) t0[L.tile(i, k)].rp = L.rp_base[L.tile(i, k)];
      HIP_OK(hipMemcpyAsync(dg->tiles, t0.data(), T * sizeof(TileDesc), hipMemcpyHostToDevice, s));
      hipLaunchKernelGGL(fill_u32, dim3(grid_for(L.rp_words)), dim3(kBlock), 0, s, dg->rowptr, L.rp_words);
      hipLaunchKernelGGL(row_counts, dim3(grid_for(g.n)), dim3(kBlock), 0, s, off, col, g.n, cuts_d, P, dg->tiles,
                         dg->rowptr);
      hipLaunchKernelGGL(chunk_sums, dim3(unsigned(std::min<uint64_t>(nch, 1u << 20))), dim3(kBlock), 0, s,
                         dg->rowptr, ch, nch, csum, cnzc, cmx);
      hipLaunchKernelGGL(tile_scan, dim3(grid_for(ntiles)), dim3(kBlock), 0, s, ch, first_d, ntiles, csum, cnzc, cmx,
                         coff, cnz, tnnz, tnzr, tmax);
      HIP_OK(hipGetLastError());
      std::vector<uint64_t> hn(T, 0);
      std::vector<uint32_t> hz(T, 0), hm(T, 0);
      HIP_OK(hipMemcpyAsync(hn.data(), tnnz, T * 8, hipMemcpyDeviceToHost, s));
      HIP_OK(hipMemcpyAsync(hz.data(), tnzr, T * 4, hipMemcpyDeviceToHost, s));
      HIP_OK(hipMemcpyAsync(hm.data(), tmax, T * 4, hipMemcpyDeviceToHost, s));
      HIP_OK(hipStreamSynchronize(s));
      for (uint32_t i = 0; i < P; ++i)
        for (uint32_t k = i; k < P; ++k) {
          const size_t t = L.tile(i, k);
          if (hn[t] > 0xFFFFFFFFull) fail(BBTC_ERR_INVALID, "a tile holds 2^32 or more entries; use more cuts");
          L.nnz[t] = hn[t];
          L.nz_rows[t] = hz[t];
          L.max_row[t] = hm[t];
        }
      layout_place(L);
      dg->htiles = tile_table(L);
      HIP_OK(hipMemcpyAsync(dg->tiles, dg->htiles.data(), T * sizeof(TileDesc), hipMemcpyHostToDevice, s));
      dg->nzrows = dev_alloc<uint32_t>(ctx, L.nz_words, &dg->bytes);
      dg->cols = dev_alloc<uint32_t>(ctx, L.col_words, &dg->bytes);
      // the first word of the row lists is zero where the layout has no entries
      hipLaunchKernelGGL(fill_u32, dim3(1), dim3(kBlock), 0, s, dg->nzrows, uint64_t(1));
      hipLaunchKernelGGL(chunk_write, dim3(unsigned(std::min<uint64_t>(nch, 1u << 20))), dim3(kBlock), 0, s,
                         dg->rowptr, ch, nch, coff, cnz, dg->tiles, dg->nzrows);
      hipLaunchKernelGGL(fill_u32, dim3(grid_for(L.col_words)), dim3(kBlock), 0, s, dg->cols, L.col_words);
      hipLaunchKernelGGL(scatter_cols, dim3(grid_for(g.n)), dim3(kBlock), 0, s, off, col, g.n, cuts_d, P, dg->tiles,
                         dg->rowptr, dg->cols);
      HIP_OK(hipGetLastError());
      HIP_OK(hipStreamSynchronize(s));
    }
    // bitmap rows of the dense tiles
    dg->dense = dev_alloc<uint32_t>(ctx, L.dense_words, &dg->bytes);
    {
      uint64_t rows = 0;
      const std::vector<DenseJob> jobs = dense_jobs(L, 0, 0, 0, 0, &rows);
      Temps tmp(ctx);
      DenseJob* jd = jobs.empty() ? nullptr : tmp.get<DenseJob>(jobs.size());
      if (jd)
        HIP_OK(hipMemcpyAsync(jd, jobs.data(), jobs.size() * sizeof(DenseJob), hipMemcpyHostToDevice, ctx->stream));
      enqueue_dense_build(dg->rowptr, dg->cols, dg->dense, L.dense_words, jd, jobs.size(), rows, ctx->stream);
      HIP_OK(hipStreamSynchronize(ctx->stream));
    }
    dg->deg = L.deg;
    dg->dwpr = L.dwpr;
    if (P <= kRuleMaxCuts) {
      // every task's work statistics, for the AUTO form rule of later plans
      dg->lex_ij.assign(T, 0);
      std::vector<bbtc_task> all;
      for (uint32_t i = 0; i < P; ++i)
        for (uint32_t j = i; j < P; ++j) {
          dg->lex_ij[size_t(i) * P + j] = all.size();
          for (uint32_t k = j; k < P; ++k) all.push_back(bbtc_task{i, j, k});
        }
      std::vector<uint64_t> st(all.size() * kTaskStats);
      device_task_stats(dg, L.deg, all.data(), all.size(), kRuleSampleUnits, st.data());
      dg->lex_work = work_from_stats(st.data(), all.size());
    }
    L.host_arrays = host_arrays;
    if (host_arrays) {
      L.rowptr.resize(L.rp_words);
      L.cols.resize(L.col_words);
      L.nzrows.resize(L.nz_words);
      copy_to_host(ctx->stream, L.rowptr.data(), dg->rowptr, L.rp_words * 4);
      copy_to_host(ctx->stream, L.cols.data(), dg->cols, L.col_words * 4);
      copy_to_host(ctx->stream, L.nzrows.data(), dg->nzrows, L.nz_words * 4);
    }
  } catch (...) {
    (void)hipStreamSynchronize(ctx->stream);
    bbtc_graph_release(dg);
    throw;
  }
  dg_out = dg;
}

}  // namespace bbtc

#define LB_TRY try {
#define LB_CATCH                                                                 \
  }                                                                              \
  catch (const bbtc::Error& e) {                                                 \
    bbtc::set_last_error(e.what());                                              \
    return e.code;                                                               \
  }                                                                              \
  catch (const std::bad_alloc&) {                                                \
    bbtc::set_last_error("host allocation failed");                              \
    return BBTC_ERR_NOMEM;                                                       \
  }                                                                              \
  catch (const std::exception& e) {                                              \
    bbtc::set_last_error(e.what());                                              \
    return BBTC_ERR_INTERNAL;                                                    \
  }                                                                              \
  return BBTC_OK;

extern "C" {

bbtc_status bbtc_graph_build(bbtc_ctx* ctx, const bbtc_hgraph* g, const uint32_t* cuts, uint32_t p, uint32_t flags,
                             bbtc_layout** layout_out, bbtc_dgraph** graph_out) {
  if (!ctx || !g || !cuts || !layout_out || !graph_out) {
    bbtc::set_last_error("null argument");
    return BBTC_ERR_INVALID;
  }
  if (p < 1) {
    bbtc::set_last_error("p must be >= 1");
    return BBTC_ERR_INVALID;
  }
  if (flags & ~uint32_t(BBTC_BUILD_HOST_ARRAYS)) {
    bbtc::set_last_error("unknown build flags");
    return BBTC_ERR_INVALID;
  }
  *layout_out = nullptr;
  *graph_out = nullptr;
  LB_TRY
  HIP_OK(hipSetDevice(ctx->device));
  auto* L = new bbtc_layout{};
  try {
    bbtc::build_on_device(ctx, g->g, std::vector<uint32_t>(cuts, cuts + p + 1), (flags & BBTC_BUILD_HOST_ARRAYS) != 0,
                          L->L, *graph_out);
  } catch (...) {
    delete L;
    throw;
  }
  *layout_out = L;
  LB_CATCH
}

bbtc_status bbtc_graph_tasks_stats(bbtc_dgraph* g, const bbtc_task* tasks, uint64_t n, uint32_t sample_units,
                                   uint64_t* stats) {
  if (!g || (!tasks && n) || (!stats && n)) {
    bbtc::set_last_error("null argument");
    return BBTC_ERR_INVALID;
  }
  for (uint64_t q = 0; q < n; ++q)
    if (!(tasks[q].i <= tasks[q].j && tasks[q].j <= tasks[q].k && tasks[q].k < g->P)) {
      bbtc::set_last_error("task out of range or not i <= j <= k");
      return BBTC_ERR_INVALID;
    }
  LB_TRY
  if (n) {
    HIP_OK(hipSetDevice(g->ctx->device));
    bbtc::device_task_stats(g, g->deg, tasks, n, sample_units, stats);
  }
  LB_CATCH
}

}  // extern "C"
