// Seeded synthetic graphs built on the GPU.
//
// Same graphs as host/generate.cpp, bit for bit: the same counter-based
// SplitMix64 streams pick the samples, and the canonical input depends only
// on the edge set and on the (degree, id) ranking, not on how they are
// computed. The host pipeline buckets and sorts on the CPU; this one sorts
// 64-bit (low, high) keys with a device radix sort:
//   sample keys -> sort -> unique -> degrees -> sort (degree, id) -> ranks
//   -> relabel to (lower rank, higher rank) -> sort -> CSR
// and copies the CSR to the host. Sample counts must stay below 2^31 (the
// radix sort's item count is an int).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <limits>

#include "objects.hpp"

namespace bbtc {
void set_last_error(const std::string& s);

namespace {

constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ull;
constexpr uint64_t kNoEdge = ~0ull;

__host__ __device__ inline uint64_t splitmix_out(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

inline uint64_t stream_key(uint64_t seed, uint64_t tag) {
  return splitmix_out(seed * kGolden + splitmix_out(tag + kGolden));
}

__device__ inline uint64_t stream_at(uint64_t key, uint64_t i) { return splitmix_out(key + (i + 1) * kGolden); }

__global__ void rmat_keys(uint64_t key, uint32_t scale, uint64_t tA, uint64_t tAB, uint64_t tABC,
                          uint64_t ns, uint64_t* __restrict__ out) {
  const uint64_t e = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= ns) return;
  const uint32_t draws = (scale + 1) / 2;
  uint32_t u = 0, v = 0;
  for (uint32_t d = 0; d < draws; ++d) {
    const uint64_t r = stream_at(key, e * draws + d);
    for (uint32_t h = 0; h < 2; ++h) {
      const uint32_t level = 2 * d + h;
      if (level >= scale) break;
      const uint64_t x = h == 0 ? (r >> 32) : (r & 0xFFFFFFFFull);
      const uint32_t bu = x >= tAB ? 1u : 0u;
      const uint32_t bv = ((x >= tA && x < tAB) || x >= tABC) ? 1u : 0u;
      const uint32_t sh = scale - 1 - level;
      u |= bu << sh;
      v |= bv << sh;
    }
  }
  out[e] = u == v ? kNoEdge : (uint64_t(min(u, v)) << 32) | max(u, v);
}

__global__ void gnm_keys(uint64_t key, uint32_t n, uint64_t nc, uint64_t* __restrict__ out,
                         uint32_t* __restrict__ idx) {
  const uint64_t c = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  const uint64_t r = stream_at(key, c);
  const uint32_t u = uint32_t(((r >> 32) * uint64_t(n)) >> 32);
  const uint32_t v = uint32_t(((r & 0xFFFFFFFFull) * uint64_t(n)) >> 32);
  out[c] = u == v ? kNoEdge : (uint64_t(min(u, v)) << 32) | max(u, v);
  idx[c] = uint32_t(c);
}

// first occurrence of each pair in key-sorted (key, index) arrays
__global__ void first_flags(const uint64_t* __restrict__ k, uint64_t n, uint8_t* __restrict__ f) {
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  f[i] = k[i] != kNoEdge && (i == 0 || k[i] != k[i - 1]);
}

// keys whose first index is at most `last`
__global__ void keep_upto(const uint64_t* __restrict__ k, const uint32_t* __restrict__ idx,
                          const uint8_t* __restrict__ first, uint64_t n, uint32_t last,
                          uint8_t* __restrict__ f) {
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  f[i] = first[i] && idx[i] <= last;
}

__global__ void degree_count(const uint64_t* __restrict__ k, uint64_t m, uint32_t* __restrict__ deg) {
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= m) return;
  atomicAdd(deg + uint32_t(k[i] >> 32), 1u);
  atomicAdd(deg + uint32_t(k[i]), 1u);
}

__global__ void degree_keys(const uint32_t* __restrict__ deg, uint32_t n, uint64_t* __restrict__ out) {
  const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v < n) out[v] = (uint64_t(deg[v]) << 32) | v;
}

__global__ void ranks_from_sorted(const uint64_t* __restrict__ s, uint32_t n, uint32_t* __restrict__ rank) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) rank[uint32_t(s[p])] = p;
}

__global__ void relabel(const uint64_t* __restrict__ k, uint64_t m, const uint32_t* __restrict__ rank,
                        uint64_t* __restrict__ out) {
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const uint32_t a = rank[uint32_t(k[i] >> 32)], b = rank[uint32_t(k[i])];
  out[i] = (uint64_t(min(a, b)) << 32) | max(a, b);
}

__global__ void row_counts(const uint64_t* __restrict__ k, uint64_t m, uint64_t* __restrict__ cnt) {
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < m) atomicAdd(reinterpret_cast<unsigned long long*>(cnt + uint32_t(k[i] >> 32)), 1ull);
}

__global__ void low_words(const uint64_t* __restrict__ k, uint64_t m, uint32_t* __restrict__ col) {
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < m) col[i] = uint32_t(k[i]);
}

inline dim3 grid_for(uint64_t n) { return dim3(unsigned((n + 255) / 256)); }

// Device scratch charged to the context's budget for its lifetime.
struct Scratch {
  bbtc_ctx* ctx;
  std::vector<std::pair<void*, uint64_t>> bufs;
  explicit Scratch(bbtc_ctx* c) : ctx(c) {}
  template <class T>
  T* get(uint64_t count) {
    const uint64_t bytes = std::max<uint64_t>(count, 1) * sizeof(T);
    if (ctx->budget && ctx->in_use + bytes > ctx->budget)
      fail(BBTC_ERR_BUDGET, "device budget exceeded while generating a graph");
    void* p = nullptr;
    HIP_OK(hipMalloc(&p, bytes));
    ctx->in_use += bytes;
    bufs.emplace_back(p, bytes);
    return static_cast<T*>(p);
  }
  ~Scratch() {
    for (auto& b : bufs) {
      (void)hipFree(b.first);
      ctx->in_use -= std::min(ctx->in_use, b.second);
    }
  }
};

// Radix sort of 64-bit keys in place (through a second buffer).
void sort_keys(Scratch& S, hipStream_t s, uint64_t*& keys, uint64_t*& alt, uint64_t n, int end_bit = 64) {
  if (n > uint64_t(std::numeric_limits<int>::max())) fail(BBTC_ERR_INVALID, "too many samples for one sort");
  hipcub::DoubleBuffer<uint64_t> db(keys, alt);
  size_t tmp = 0;
  HIP_OK(hipcub::DeviceRadixSort::SortKeys(nullptr, tmp, db, int(n), 0, end_bit, s));
  void* t = S.get<uint8_t>(tmp);
  HIP_OK(hipcub::DeviceRadixSort::SortKeys(t, tmp, db, int(n), 0, end_bit, s));
  if (db.Current() != keys) std::swap(keys, alt);
}

uint64_t select_flagged(Scratch& S, hipStream_t s, const uint64_t* in, const uint8_t* flags, uint64_t n,
                        uint64_t* out) {
  int* d_num = S.get<int>(1);
  size_t tmp = 0;
  HIP_OK(hipcub::DeviceSelect::Flagged(nullptr, tmp, in, flags, out, d_num, int(n), s));
  void* t = S.get<uint8_t>(tmp);
  HIP_OK(hipcub::DeviceSelect::Flagged(t, tmp, in, flags, out, d_num, int(n), s));
  int h = 0;
  HIP_OK(hipMemcpyAsync(&h, d_num, sizeof(int), hipMemcpyDeviceToHost, s));
  HIP_OK(hipStreamSynchronize(s));
  return uint64_t(h);
}

// Canonical CSR from m unique sorted (low, high) keys over n vertices.
Csr canonical_from_keys(Scratch& S, hipStream_t s, uint32_t n, uint64_t* keys, uint64_t* alt, uint64_t m) {
  uint32_t* deg = S.get<uint32_t>(n);
  HIP_OK(hipMemsetAsync(deg, 0, 4ull * n, s));
  if (m) hipLaunchKernelGGL(degree_count, grid_for(m), dim3(256), 0, s, keys, m, deg);
  uint64_t* dk = S.get<uint64_t>(n);
  uint64_t* dk2 = S.get<uint64_t>(n);
  hipLaunchKernelGGL(degree_keys, grid_for(n), dim3(256), 0, s, deg, n, dk);
  sort_keys(S, s, dk, dk2, n);
  uint32_t* rank = deg;  // degrees are no longer needed
  hipLaunchKernelGGL(ranks_from_sorted, grid_for(n), dim3(256), 0, s, dk, n, rank);
  if (m) {
    hipLaunchKernelGGL(relabel, grid_for(m), dim3(256), 0, s, keys, m, rank, alt);
    std::swap(keys, alt);
    sort_keys(S, s, keys, alt, m);
  }
  uint64_t* cnt = S.get<uint64_t>(uint64_t(n) + 1);
  HIP_OK(hipMemsetAsync(cnt, 0, 8ull * (n + 1ull), s));
  if (m) hipLaunchKernelGGL(row_counts, grid_for(m), dim3(256), 0, s, keys, m, cnt);
  uint64_t* off = S.get<uint64_t>(uint64_t(n) + 1);
  size_t tmp = 0;
  HIP_OK(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, cnt, off, int(n + 1ull), s));
  void* t = S.get<uint8_t>(tmp);
  HIP_OK(hipcub::DeviceScan::ExclusiveSum(t, tmp, cnt, off, int(n + 1ull), s));
  uint32_t* col = reinterpret_cast<uint32_t*>(alt);
  if (m) hipLaunchKernelGGL(low_words, grid_for(m), dim3(256), 0, s, keys, m, col);
  HIP_OK(hipGetLastError());
  Csr g;
  g.n = g.ncols = n;
  g.m = m;
  g.off.resize(size_t(n) + 1);
  g.col.resize(m);
  HIP_OK(hipMemcpyAsync(g.off.data(), off, 8ull * (n + 1ull), hipMemcpyDeviceToHost, s));
  if (m) HIP_OK(hipMemcpyAsync(g.col.data(), col, 4ull * m, hipMemcpyDeviceToHost, s));
  HIP_OK(hipStreamSynchronize(s));
  return g;
}

Csr device_rmat(bbtc_ctx* ctx, uint32_t scale, uint32_t ef, double a, double b, double c, uint64_t seed) {
  if (scale < 1 || scale > 31) fail(BBTC_ERR_INVALID, "RMAT scale must be in [1,31]");
  if (!(a >= 0 && b >= 0 && c >= 0 && a + b + c <= 1.0 + 1e-12))
    fail(BBTC_ERR_INVALID, "RMAT probabilities must be >= 0 with a+b+c <= 1");
  HIP_OK(hipSetDevice(ctx->device));
  const uint32_t n = 1u << scale;
  const uint64_t ns = uint64_t(ef) << scale;
  const double two32 = 4294967296.0;
  const uint64_t tA = uint64_t(std::floor(a * two32));
  const uint64_t tAB = uint64_t(std::floor((a + b) * two32));
  const uint64_t tABC = std::min<uint64_t>(uint64_t(std::floor((a + b + c) * two32)), 1ull << 32);
  Scratch S(ctx);
  hipStream_t s = ctx->stream;
  uint64_t* keys = S.get<uint64_t>(ns);
  uint64_t* alt = S.get<uint64_t>(ns);
  if (ns) hipLaunchKernelGGL(rmat_keys, grid_for(ns), dim3(256), 0, s, stream_key(seed, 0x524D4154ull),
                             scale, tA, tAB, tABC, ns, keys);
  HIP_OK(hipGetLastError());
  sort_keys(S, s, keys, alt, ns);
  uint8_t* f = S.get<uint8_t>(ns);
  if (ns) hipLaunchKernelGGL(first_flags, grid_for(ns), dim3(256), 0, s, keys, ns, f);
  const uint64_t m = ns ? select_flagged(S, s, keys, f, ns, alt) : 0;
  std::swap(keys, alt);
  return canonical_from_keys(S, s, n, keys, alt, m);
}

Csr device_gnm(bbtc_ctx* ctx, uint32_t n, uint64_t m, uint64_t seed) {
  if (n < 2) fail(BBTC_ERR_INVALID, "G(n,m) needs n >= 2");
  const uint64_t maxm = uint64_t(n) * (n - 1) / 2;
  if (m > maxm / 2) fail(BBTC_ERR_INVALID, "G(n,m): m above half of all pairs is not supported by the pair stream");
  HIP_OK(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  const uint64_t key = stream_key(seed, 0x474E4Dull);
  for (uint64_t C = m + m / 8 + 1024;; C *= 2) {
    if (C >= (1ull << 31)) fail(BBTC_ERR_INVALID, "G(n,m): too many candidates for one sort");
    Scratch S(ctx);
    uint64_t* keys = S.get<uint64_t>(C);
    uint64_t* keys2 = S.get<uint64_t>(C);
    uint32_t* idx = S.get<uint32_t>(C);
    uint32_t* idx2 = S.get<uint32_t>(C);
    hipLaunchKernelGGL(gnm_keys, grid_for(C), dim3(256), 0, s, key, n, C, keys, idx);
    HIP_OK(hipGetLastError());
    // stable sort by pair: equal pairs keep ascending candidate index
    hipcub::DoubleBuffer<uint64_t> dk(keys, keys2);
    hipcub::DoubleBuffer<uint32_t> dv(idx, idx2);
    size_t tmp = 0;
    HIP_OK(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, dk, dv, int(C), 0, 64, s));
    void* t = S.get<uint8_t>(tmp);
    HIP_OK(hipcub::DeviceRadixSort::SortPairs(t, tmp, dk, dv, int(C), 0, 64, s));
    uint64_t* sk = dk.Current();
    uint32_t* sv = dv.Current();
    uint8_t* first = S.get<uint8_t>(C);
    hipLaunchKernelGGL(first_flags, grid_for(C), dim3(256), 0, s, sk, C, first);
    // first-occurrence indices, sorted: the m-th smallest closes the prefix
    uint32_t* fidx = S.get<uint32_t>(C);
    int* d_num = S.get<int>(1);
    tmp = 0;
    HIP_OK(hipcub::DeviceSelect::Flagged(nullptr, tmp, sv, first, fidx, d_num, int(C), s));
    void* t2 = S.get<uint8_t>(tmp);
    HIP_OK(hipcub::DeviceSelect::Flagged(t2, tmp, sv, first, fidx, d_num, int(C), s));
    int distinct = 0;
    HIP_OK(hipMemcpyAsync(&distinct, d_num, sizeof(int), hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    if (uint64_t(distinct) < m) continue;
    uint32_t last = 0;
    if (m) {
      uint32_t* fidx2 = S.get<uint32_t>(uint64_t(distinct));
      hipcub::DoubleBuffer<uint32_t> df(fidx, fidx2);
      tmp = 0;
      HIP_OK(hipcub::DeviceRadixSort::SortKeys(nullptr, tmp, df, distinct, 0, 32, s));
      void* t3 = S.get<uint8_t>(tmp);
      HIP_OK(hipcub::DeviceRadixSort::SortKeys(t3, tmp, df, distinct, 0, 32, s));
      HIP_OK(hipMemcpyAsync(&last, df.Current() + (m - 1), 4, hipMemcpyDeviceToHost, s));
      HIP_OK(hipStreamSynchronize(s));
    }
    uint8_t* keep = S.get<uint8_t>(C);
    hipLaunchKernelGGL(keep_upto, grid_for(C), dim3(256), 0, s, sk, sv, first, C, last, keep);
    uint64_t* out = sk == keys ? keys2 : keys;
    const uint64_t got = m ? select_flagged(S, s, sk, keep, C, out) : 0;
    if (got != m) fail(BBTC_ERR_INTERNAL, "G(n,m) selection produced " + std::to_string(got) + " edges");
    uint64_t* other = out == keys ? keys2 : keys;
    return canonical_from_keys(S, s, n, out, other, m);
  }
}

}  // namespace
}  // namespace bbtc

extern "C" {

bbtc_status bbtc_hgraph_gen_rmat_device(bbtc_ctx* ctx, uint32_t scale, uint32_t edge_factor, double a,
                                        double b, double c, uint64_t seed, bbtc_hgraph** out) {
  if (!ctx || !out) {
    bbtc::set_last_error("null argument");
    return BBTC_ERR_INVALID;
  }
  try {
    *out = new bbtc_hgraph{bbtc::device_rmat(ctx, scale, edge_factor, a, b, c, seed)};
  } catch (const bbtc::Error& e) {
    bbtc::set_last_error(e.what());
    return e.code;
  } catch (const std::bad_alloc&) {
    bbtc::set_last_error("host allocation failed");
    return BBTC_ERR_NOMEM;
  }
  return BBTC_OK;
}

bbtc_status bbtc_hgraph_gen_gnm_device(bbtc_ctx* ctx, uint32_t n, uint64_t m, uint64_t seed,
                                       bbtc_hgraph** out) {
  if (!ctx || !out) {
    bbtc::set_last_error("null argument");
    return BBTC_ERR_INVALID;
  }
  try {
    *out = new bbtc_hgraph{bbtc::device_gnm(ctx, n, m, seed)};
  } catch (const bbtc::Error& e) {
    bbtc::set_last_error(e.what());
    return e.code;
  } catch (const std::bad_alloc&) {
    bbtc::set_last_error("host allocation failed");
    return BBTC_ERR_NOMEM;
  }
  return BBTC_OK;
}

}  // extern "C"
