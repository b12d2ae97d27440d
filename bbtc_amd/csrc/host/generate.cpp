// Seeded synthetic graphs and the canonical-input pipeline.
//
// Canonical input (what every .bbtc here holds): undirected simple graph,
// vertices ranked by (degree, id) ascending, each edge stored once in the
// row of its lower-ranked endpoint, rows sorted. This matches the reference
// ingest (utils.hpp:171-268 + sarma::Matrix::order(ASC, upper), csr_matrix.hpp
// 503-584) except that duplicates are removed and ties in degree are broken
// by vertex id, so the result does not depend on a sort's tie order.
//
// Random numbers come from counter-based SplitMix64 streams: draw i of a
// stream is a pure function of (seed, stream tag, i), so any thread count
// (and a GPU) produces the same graph.
#include <omp.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <vector>

#include "graph.hpp"

namespace bbtc {

namespace {

constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ull;

inline uint64_t splitmix_out(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Draw i of the stream keyed by (seed, tag): SplitMix64 output for state
// key + (i + 1) * golden.
struct Stream {
  uint64_t key;
  Stream(uint64_t seed, uint64_t tag) : key(splitmix_out(seed * kGolden + splitmix_out(tag + kGolden))) {}
  uint64_t at(uint64_t i) const { return splitmix_out(key + (i + 1) * kGolden); }
};

template <class T>
inline T fetch_add(T* p, T v) { return __atomic_fetch_add(p, v, __ATOMIC_RELAXED); }

// Exclusive prefix sum of cnt[0..n) into out[0..n] (out[n] = total), over
// T parts; each thread of the team (which may be smaller than T) takes parts
// tid, tid + team, ...
template <class C>
void exclusive_scan(const C* cnt, uint64_t* out, size_t n) {
  const int T = std::max(1, omp_get_max_threads());
  std::vector<uint64_t> part(size_t(T) + 1, 0);
#pragma omp parallel num_threads(T)
  {
    const int team = omp_get_num_threads();
    for (int t = omp_get_thread_num(); t < T; t += team) {
      const size_t b = n * t / T, e = n * (t + 1) / T;
      uint64_t s = 0;
      for (size_t i = b; i < e; ++i) s += cnt[i];
      part[t + 1] = s;
    }
#pragma omp barrier
#pragma omp single
    for (int i = 0; i < T; ++i) part[i + 1] += part[i];
    for (int t = omp_get_thread_num(); t < T; t += team) {
      const size_t b = n * t / T, e = n * (t + 1) / T;
      uint64_t acc = part[t];
      for (size_t i = b; i < e; ++i) { out[i] = acc; acc += cnt[i]; }
    }
  }
  out[n] = part[T];
}

// Bucket the pairs produced by gen(e, &lo, &hi) (lo < hi; gen returns false
// to drop e) by lo, sort each bucket and remove duplicates. Result: CSR over
// original ids holding each undirected edge once, in the row of its smaller id.
template <class Gen>
Csr dedup_by_low(uint32_t n, uint64_t nsamples, const Gen& gen) {
  Buf<uint32_t> cnt(n);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < int64_t(n); ++i) cnt[i] = 0;
#pragma omp parallel for schedule(static, 1 << 16)
  for (int64_t e = 0; e < int64_t(nsamples); ++e) {
    uint32_t lo, hi;
    if (gen(uint64_t(e), &lo, &hi)) fetch_add(&cnt[lo], 1u);
  }
  Buf<uint64_t> start(size_t(n) + 1);
  exclusive_scan(cnt.data(), start.data(), n);
  const uint64_t total = start[n];
  Buf<uint64_t> cur(n);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < int64_t(n); ++i) cur[i] = start[i];
  Buf<uint32_t> adj(total);
#pragma omp parallel for schedule(static, 1 << 16)
  for (int64_t e = 0; e < int64_t(nsamples); ++e) {
    uint32_t lo, hi;
    if (gen(uint64_t(e), &lo, &hi)) adj[fetch_add(&cur[lo], uint64_t(1))] = hi;
  }
  cur = Buf<uint64_t>();
  // sort + unique each bucket; cnt becomes the deduplicated length
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t u = 0; u < int64_t(n); ++u) {
    uint32_t* b = adj.data() + start[u];
    uint32_t* e = adj.data() + start[u + 1];
    std::sort(b, e);
    cnt[u] = uint32_t(std::unique(b, e) - b);
  }
  Csr d;
  d.n = d.ncols = n;
  d.off.resize(size_t(n) + 1);
  exclusive_scan(cnt.data(), d.off.data(), n);
  d.m = d.off[n];
  d.col.resize(d.m);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t u = 0; u < int64_t(n); ++u)
    std::memcpy(d.col.data() + d.off[u], adj.data() + start[u], 4ull * cnt[u]);
  return d;
}

// From a deduplicated low-row CSR over original ids to the canonical input.
Csr rank_and_orient(const Csr& d) {
  const uint32_t n = d.n;
  Buf<uint32_t> deg(n);
#pragma omp parallel for schedule(static)
  for (int64_t u = 0; u < int64_t(n); ++u) deg[u] = uint32_t(d.off[u + 1] - d.off[u]);
#pragma omp parallel for schedule(static, 1 << 16)
  for (int64_t x = 0; x < int64_t(d.m); ++x) fetch_add(&deg[d.col[x]], 1u);

  // rank by (degree, id): counting sort on degree, ids ascending within a degree
  uint32_t maxdeg = 0;
#pragma omp parallel for reduction(max : maxdeg)
  for (int64_t u = 0; u < int64_t(n); ++u) maxdeg = std::max(maxdeg, deg[u]);
  std::vector<uint64_t> pos(size_t(maxdeg) + 2, 0);
  for (uint32_t u = 0; u < n; ++u) pos[deg[u] + 1]++;
  for (size_t x = 1; x < pos.size(); ++x) pos[x] += pos[x - 1];
  Buf<uint32_t> rank(n);
  for (uint32_t u = 0; u < n; ++u) rank[u] = uint32_t(pos[deg[u]]++);

  // new row of each edge = lower rank; count, scan, scatter, sort rows
  Buf<uint32_t>& cnt = deg;
#pragma omp parallel for schedule(static)
  for (int64_t u = 0; u < int64_t(n); ++u) cnt[u] = 0;
#pragma omp parallel for schedule(dynamic, 4096)
  for (int64_t u = 0; u < int64_t(n); ++u) {
    const uint32_t ru = rank[u];
    for (uint64_t x = d.off[u]; x < d.off[u + 1]; ++x) {
      const uint32_t rv = rank[d.col[x]];
      fetch_add(&cnt[std::min(ru, rv)], 1u);
    }
  }
  Csr g;
  g.n = g.ncols = n;
  g.m = d.m;
  g.off.resize(size_t(n) + 1);
  exclusive_scan(cnt.data(), g.off.data(), n);
  Buf<uint64_t> cur(n);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < int64_t(n); ++i) cur[i] = g.off[i];
  g.col.resize(g.m);
#pragma omp parallel for schedule(dynamic, 4096)
  for (int64_t u = 0; u < int64_t(n); ++u) {
    const uint32_t ru = rank[u];
    for (uint64_t x = d.off[u]; x < d.off[u + 1]; ++x) {
      const uint32_t rv = rank[d.col[x]];
      const uint32_t lo = std::min(ru, rv), hi = std::max(ru, rv);
      g.col[fetch_add(&cur[lo], uint64_t(1))] = hi;
    }
  }
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t u = 0; u < int64_t(n); ++u)
    std::sort(g.col.data() + g.off[u], g.col.data() + g.off[u + 1]);
  return g;
}

}  // namespace

Csr canonical_from_edges(uint32_t n, uint64_t nedges, const uint32_t* src, const uint32_t* dst) {
  if (nedges && (!src || !dst)) fail(BBTC_ERR_INVALID, "null edge arrays");
  if (uint64_t(n) + 1 > 0xFFFFFFFFull) fail(BBTC_ERR_INVALID, "n+1 must fit in 32 bits");
  for (uint64_t e = 0; e < nedges; ++e)
    if (src[e] >= n || dst[e] >= n) fail(BBTC_ERR_INVALID, "edge endpoint >= n");
  auto gen = [&](uint64_t e, uint32_t* lo, uint32_t* hi) {
    const uint32_t u = src[e], v = dst[e];
    if (u == v) return false;
    *lo = std::min(u, v);
    *hi = std::max(u, v);
    return true;
  };
  Csr d = dedup_by_low(n, nedges, gen);
  return rank_and_orient(d);
}

Csr gen_rmat(uint32_t scale, uint32_t edge_factor, double a, double b, double c, uint64_t seed) {
  if (scale < 1 || scale > 31) fail(BBTC_ERR_INVALID, "RMAT scale must be in [1,31]");
  if (!(a >= 0 && b >= 0 && c >= 0 && a + b + c <= 1.0 + 1e-12))
    fail(BBTC_ERR_INVALID, "RMAT probabilities must be >= 0 with a+b+c <= 1");
  const uint32_t n = 1u << scale;
  const uint64_t ns = uint64_t(edge_factor) << scale;
  // 32-bit thresholds: a draw r in [0, 2^32) picks quadrant a if r < tA, etc.
  const double two32 = 4294967296.0;
  const uint64_t tA = uint64_t(std::floor(a * two32));
  const uint64_t tAB = uint64_t(std::floor((a + b) * two32));
  const uint64_t tABC = std::min<uint64_t>(uint64_t(std::floor((a + b + c) * two32)), 1ull << 32);
  const Stream st(seed, 0x524D4154ull);  // "RMAT"
  const uint32_t draws = (scale + 1) / 2;  // two 32-bit levels per 64-bit draw
  auto gen = [&](uint64_t e, uint32_t* lo, uint32_t* hi) {
    uint32_t u = 0, v = 0;
    for (uint32_t d = 0; d < draws; ++d) {
      const uint64_t r = st.at(e * draws + d);
      for (uint32_t h = 0; h < 2; ++h) {
        const uint32_t level = 2 * d + h;
        if (level >= scale) break;
        const uint64_t x = h == 0 ? (r >> 32) : (r & 0xFFFFFFFFull);
        const uint32_t bu = x >= tAB ? 1u : 0u;  // quadrants c, d set the row bit
        const uint32_t bv = (x >= tA && x < tAB) || x >= tABC ? 1u : 0u;  // b, d set the column bit
        const uint32_t sh = scale - 1 - level;
        u |= bu << sh;
        v |= bv << sh;
      }
    }
    if (u == v) return false;
    *lo = std::min(u, v);
    *hi = std::max(u, v);
    return true;
  };
  Csr d = dedup_by_low(n, ns, gen);
  return rank_and_orient(d);
}

Csr gen_gnm(uint32_t n, uint64_t m, uint64_t seed) {
  if (n < 2) fail(BBTC_ERR_INVALID, "G(n,m) needs n >= 2");
  const uint64_t maxm = uint64_t(n) * (n - 1) / 2;
  if (m > maxm) fail(BBTC_ERR_INVALID, "G(n,m): m exceeds n(n-1)/2");
  if (m > maxm / 2)
    fail(BBTC_ERR_INVALID, "G(n,m): m above half of all pairs is not supported by the pair stream");
  const Stream st(seed, 0x474E4Dull);  // "GNM"
  auto cand = [&](uint64_t c, uint32_t* lo, uint32_t* hi) {
    const uint64_t r = st.at(c);
    const uint32_t u = uint32_t(((r >> 32) * uint64_t(n)) >> 32);
    const uint32_t v = uint32_t(((r & 0xFFFFFFFFull) * uint64_t(n)) >> 32);
    if (u == v) return false;
    *lo = std::min(u, v);
    *hi = std::max(u, v);
    return true;
  };
  // Find the shortest prefix [0, C) of the candidate stream holding m distinct
  // pairs: collect (pair, first index) over a prefix long enough, then take the
  // m-th smallest first index.
  uint64_t C = m + m / 8 + 1024;
  uint64_t cstar = 0;
  for (;;) {
    Buf<uint32_t> cnt(n);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < int64_t(n); ++i) cnt[i] = 0;
#pragma omp parallel for schedule(static, 1 << 16)
    for (int64_t c = 0; c < int64_t(C); ++c) {
      uint32_t lo, hi;
      if (cand(uint64_t(c), &lo, &hi)) fetch_add(&cnt[lo], 1u);
    }
    Buf<uint64_t> start(size_t(n) + 1);
    exclusive_scan(cnt.data(), start.data(), n);
    Buf<uint64_t> cur(n);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < int64_t(n); ++i) cur[i] = start[i];
    Buf<uint64_t> item(start[n]);  // candidate index
    Buf<uint32_t> hiv(start[n]);   // its larger endpoint
#pragma omp parallel for schedule(static, 1 << 16)
    for (int64_t c = 0; c < int64_t(C); ++c) {
      uint32_t lo, hi;
      if (cand(uint64_t(c), &lo, &hi)) {
        const uint64_t at = fetch_add(&cur[lo], uint64_t(1));
        item[at] = uint64_t(c);
        hiv[at] = hi;
      }
    }
    // per bucket: sort by (hi, index), keep the first index of each hi
    Buf<uint64_t> firsts(start[n]);
#pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t u = 0; u < int64_t(n); ++u) {
      const uint64_t b = start[u], e = start[u + 1];
      if (b == e) continue;
      std::vector<std::pair<uint32_t, uint64_t>> v;
      v.reserve(e - b);
      for (uint64_t x = b; x < e; ++x) v.emplace_back(hiv[x], item[x]);
      std::sort(v.begin(), v.end());
      uint64_t w = b;
      for (size_t x = 0; x < v.size(); ++x)
        if (x == 0 || v[x].first != v[x - 1].first) firsts[w++] = v[x].second;
      for (; w < e; ++w) firsts[w] = ~0ull;
    }
    // gather the valid first indices
    std::vector<uint64_t> fi;
    fi.reserve(start[n]);
    for (uint64_t x = 0; x < start[n]; ++x)
      if (firsts[x] != ~0ull) fi.push_back(firsts[x]);
    if (fi.size() >= m) {
      if (m == 0) { cstar = 0; break; }
      std::nth_element(fi.begin(), fi.begin() + (m - 1), fi.end());
      cstar = fi[m - 1] + 1;
      break;
    }
    C = C * 2;
  }
  Csr d = dedup_by_low(n, cstar, cand);
  if (d.m != m) fail(BBTC_ERR_INTERNAL, "G(n,m) produced " + std::to_string(d.m) + " edges");
  return rank_and_orient(d);
}

}  // namespace bbtc
