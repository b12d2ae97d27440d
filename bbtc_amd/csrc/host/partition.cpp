// 2D partitioners behind the reference's `--partitioner uni|n1d|pal`.
//
// Each one reproduces the reference's cut vector bit for bit, including its
// integer widths (everything is 32-bit unsigned, as Ordinal/Value are in
// definitions.h:9-10) and its random sample (std::mt19937_64 seeded with
// 1091467301 feeding std::uniform_real_distribution<double>).
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <limits>
#include <random>
#include <vector>

#include "graph.hpp"

namespace bbtc {

// ---- uni: sarma::uniform::partition(N, P) (uniform.hpp:20-27) ---------------
std::vector<uint32_t> partition_uni(const Csr& g, uint32_t p) {
  std::vector<uint32_t> c(size_t(p) + 1);
  for (uint32_t i = 0; i <= p; ++i) c[i] = uint32_t((long long)i * (long long)g.n / p);
  return c;
}

// ---- n1d: sarma::nicol1d::partition(indptr, P) (nicol1d.hpp:30-216) ---------
namespace {

using u32 = uint32_t;

// Probe of nicol1d.hpp:30-65 for one prefix array (Q = 1): greedy cuts under
// load bound L, each cut as far right as possible but not beyond UB.
std::vector<u32> n1d_probe(const std::vector<u32>& prefix, u32 L, const std::vector<u32>& UB) {
  std::vector<u32> p;
  p.reserve(UB.size());
  p.push_back(0);
  while (p.size() < UB.size()) {
    u32 r = UB[p.size()];
    u32 l = p.back();
    const u32 base = prefix[p.back()];
    while (l < r) {
      const u32 m = (l + 1) + (r - (l + 1)) / 2;
      const u32 cl = prefix[m] - base;
      if (cl <= L) l = m;
      else r = m - 1;
    }
    p.push_back(r);
    if (p.back() == p[p.size() - 2]) {
      while (p.size() < UB.size()) p.push_back(p.back());
      break;
    }
  }
  return p;
}

// Optimal 1D cuts of a prefix array (nicol1d.hpp:106-195, one array): prefix
// has N + 1 entries (prefix[0] = 0) accumulated in 32 bits; returns P + 1
// cuts over [0, N].
std::vector<u32> nicol_prefix(std::vector<u32> prefix, u32 P) {
  const u32 N = u32(prefix.size() - 1);
  // One trailing zero: partition_indices reads prefix[l + 1] with l == N at
  // the last cut (nicol1d.hpp:144), one element past the array; the zero
  // filled pages a large allocation sits in supply 0 there.
  prefix.push_back(0);
  std::vector<u32> UB(size_t(P) + 1, N);
  u32 lower = prefix[N] / P, upper = prefix[N];
  auto cub = [&](u32 l, u32 r) { return u32(prefix[r] - prefix[l]); };
  u32 pback = 0;
  for (u32 i = 1; i <= P; ++i) {
    u32 l = pback, r = UB[i];
    while (l < r) {
      const u32 m = (l + 1) + (r - (l + 1)) / 2;
      const u32 b = cub(pback, m);
      if (b > upper) r = m - 1;
      else if (b < lower) l = m;
      else {
        auto pr = n1d_probe(prefix, b, UB);
        if (pr.back() < N) { l = m; lower = b; }
        else { r = m - 1; UB = pr; upper = b; }
      }
    }
    const u32 b = cub(pback, l + 1);
    if (lower < b && b < upper) {
      auto pr = n1d_probe(prefix, b, UB);
      if (pr.back() >= N) { UB = pr; upper = b; }
      else lower = b;
    }
    pback = l;
  }
  return n1d_probe(prefix, upper, UB);
}

}  // namespace

std::vector<uint32_t> partition_n1d(const Csr& g, uint32_t P) {
  // The weight of element i is indptr[i] itself (the reference passes the
  // offset array, utils.hpp:281) and the prefix is accumulated in 32 bits,
  // wrapping (nicol1d.hpp:211-214).
  const size_t asz = size_t(g.n) + 1;
  std::vector<u32> prefix(asz + 1, 0);
  for (size_t i = 0; i < asz; ++i) prefix[i + 1] = prefix[i] + u32(g.off[i]);
  return nicol_prefix(std::move(prefix), P);
}

// ---- rac: sarma::refine_a_cut::partition(A, P) (refine_a_cut.hpp:43-223) -------
namespace {

// Row ranges of g with about m / T entries each: T + 1 row bounds.
std::vector<u32> entry_split(const Csr& g, int T) {
  std::vector<u32> at(size_t(T) + 1, g.n);
  at[0] = 0;
  for (int t = 1; t < T; ++t) {
    const uint64_t want = g.m * uint64_t(t) / uint64_t(T);
    at[t] = u32(std::lower_bound(g.off.begin(), g.off.begin() + g.n, want) - g.off.begin());
  }
  for (int t = 1; t <= T; ++t) at[t] = std::max(at[t], at[t - 1]);
  return at;
}

// Rows of the transpose of an n x n CSR, each sorted. Each thread takes a
// range of rows, counts its entries per column in a histogram of its own and
// writes them after those of the threads before it, so every transposed row
// comes out ascending without a sort. When threads x columns would outgrow
// the entries, one thread does it.
void transpose_rows(const Csr& g, std::vector<u32>& tp, std::vector<u32>& ti) {
  const u32 n = g.n;
  tp.assign(size_t(n) + 1, 0);
  ti.resize(g.m);
  const int T = std::max(1, omp_get_max_threads());
  if (T == 1 || uint64_t(T) * n > 2 * g.m + (uint64_t(1) << 24)) {
    for (uint64_t e = 0; e < g.m; ++e) tp[size_t(g.col[e]) + 1]++;
    for (size_t i = 1; i < tp.size(); ++i) tp[i] += tp[i - 1];
    std::vector<u32> pos(tp.begin(), tp.end() - 1);
    for (u32 r = 0; r < n; ++r)
      for (uint64_t e = g.off[r]; e < g.off[r + 1]; ++e) ti[pos[g.col[e]]++] = r;
    return;
  }
  const std::vector<u32> at = entry_split(g, T);
  std::vector<u32> hist(size_t(T) * n, 0);
#pragma omp parallel num_threads(T)
  {
    // T parts; a team smaller than T takes parts tid, tid + team, ...
    const int team = omp_get_num_threads();
    for (int t = omp_get_thread_num(); t < T; t += team) {
      u32* h = hist.data() + size_t(t) * n;
      for (uint64_t e = g.off[at[t]]; e < g.off[at[t + 1]]; ++e) ++h[g.col[e]];
    }
#pragma omp barrier
#pragma omp for schedule(static)
    for (int64_t c = 0; c < int64_t(n); ++c) {
      u32 k = 0;
      for (int v = 0; v < T; ++v) k += hist[size_t(v) * n + size_t(c)];
      tp[size_t(c) + 1] = k;
    }
#pragma omp single
    for (size_t i = 1; i < tp.size(); ++i) tp[i] += tp[i - 1];
#pragma omp for schedule(static)
    for (int64_t c = 0; c < int64_t(n); ++c) {
      u32 slot = tp[size_t(c)];
      for (int v = 0; v < T; ++v) {
        u32& x = hist[size_t(v) * n + size_t(c)];
        const u32 k = x;
        x = slot;
        slot += k;
      }
    }
    for (int t = omp_get_thread_num(); t < T; t += team) {
      u32* h = hist.data() + size_t(t) * n;
      for (u32 r = at[t]; r < at[t + 1]; ++r)
        for (uint64_t e = g.off[r]; e < g.off[r + 1]; ++e) ti[h[g.col[e]]++] = r;
    }
  }
}

// refine_a_cut.hpp:43-73: per row, the largest count of its entries inside
// one interval of prev; prefix over rows in 32 bits; optimal 1D cuts.
template <class Off>
std::vector<u32> rac_refine(u32 n, const Off* off, const u32* col, const std::vector<u32>& prev, u32 P) {
  std::vector<u32> prefix(size_t(n) + 1, 0);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t i = 0; i < int64_t(n); ++i) {
    u32 max_i = 0, max_k = 0, k = 0;
    for (uint64_t j = off[i]; j < off[i + 1]; ++j) {
      while (col[j] >= prev[k]) {
        max_i = std::max(max_i, max_k);
        max_k = 0;
        ++k;
      }
      max_k += 1;
    }
    prefix[size_t(i) + 1] = std::max(max_i, max_k);
  }
  for (size_t i = 1; i < prefix.size(); ++i) prefix[i] += prefix[i - 1];
  return nicol_prefix(std::move(prefix), P);
}

// Matrix::compute_maxload(c, c) (csr_matrix.hpp:281-322): the largest tile
// entry count with the same cuts on rows and columns. Threads count ranges of
// rows into tile tables of their own, summed at the end (one table when
// threads x tiles would be large).
u32 maxload(const Csr& g, const std::vector<u32>& c) {
  const size_t B = c.size() - 1;
  const int T = B * B * size_t(omp_get_max_threads()) <= (size_t(1) << 24) ? std::max(1, omp_get_max_threads()) : 1;
  const std::vector<u32> at = entry_split(g, T);
  std::vector<u32> load(size_t(T) * B * B, 0);
#pragma omp parallel num_threads(T)
  for (int t = omp_get_thread_num(); t < T; t += omp_get_num_threads()) {
    u32* mine = load.data() + size_t(t) * B * B;
    // u: the number of cuts c[0..B) at or below row i (the row band is u - 1)
    size_t u = size_t(std::upper_bound(c.begin(), c.begin() + B, at[t]) - c.begin());
    for (u32 i = at[t]; i < at[t + 1]; ++i) {
      while (u < B && i >= c[u]) ++u;
      for (uint64_t e = g.off[i]; e < g.off[i + 1]; ++e) {
        const size_t q = size_t(std::lower_bound(c.begin(), c.end(), g.col[e] + 1) - c.begin()) - 1;
        mine[(u - 1) * B + q] += 1;
      }
    }
  }
  for (int t = 1; t < T; ++t)
    for (size_t x = 0; x < B * B; ++x) load[x] += load[size_t(t) * B * B + x];
  return *std::max_element(load.begin(), load.begin() + B * B);
}

}  // namespace

std::vector<uint32_t> partition_rac(const Csr& g, uint32_t P) {
  const u32 n = g.n;
  std::vector<u32> tp, ti;
  transpose_rows(g, tp, ti);
  std::vector<u32> cx(size_t(P) + 1, n), cy(size_t(P) + 1, n);
  cx[0] = cy[0] = 0;
  auto refine_x = [&]() { return rac_refine(n, g.off.data(), g.col.data(), cx, P); };
  auto refine_y = [&]() { return rac_refine(n, tp.data(), ti.data(), cy, P); };
  // direction with the smaller max tile load after one refinement
  // (refine_a_cut.hpp:153-166), then max(10, P) refinements along it
  const u32 max_x = maxload(g, refine_x());
  const u32 max_y = maxload(g, refine_y());
  const bool along_x = max_x <= max_y;
  const int iters = std::max(10, int(P));
  for (int it = 0; it < iters; ++it) {
    if (along_x) cx = refine_x();
    else cy = refine_y();
  }
  return along_x ? cx : cy;
}

// ---- pal: sparsify + sarma::probe_a_load::partition ----------------------------
namespace {

// Bernoulli sample of the nonzeros, as sarma::utils::sparse_mask
// (tools/utils.hpp:81-106) + Matrix::sparsify (csr_matrix.hpp:435-454).
// Returns the sampled entries as (row offsets, cols).
void sparsify(const Csr& g, double keep, int seed, std::vector<u32>& sp_off, std::vector<u32>& sp_col) {
  const size_t N = size_t(u32(g.m));  // the reference's nnz is a 32-bit Ordinal
  // positions of the kept entries, increasing: the draws of the mask, one per
  // entry or one per gap between kept entries
  std::vector<uint64_t> kept;
  std::mt19937_64 gen(seed);
  std::uniform_real_distribution<double> rnd;
  if (keep >= 1.0) {
    kept.resize(N);
    for (size_t i = 0; i < N; ++i) kept[i] = i;
  } else if (keep > 1.0 / 3) {
    for (size_t i = 0; i < N; ++i)
      if (rnd(gen) <= keep) kept.push_back(i);
  } else {
    const double Q = std::log(1 - keep);
    for (size_t i = std::numeric_limits<size_t>::max();;) {
      const size_t fc = size_t(std::log(1 - rnd(gen)) / Q);
      i += std::min(fc, N);
      if (++i < N) kept.push_back(i);
      else break;
    }
  }
  // the row of each kept entry; rows keep their entries in column order
  sp_off.assign(size_t(g.n) + 1, 0);
  sp_col.resize(kept.size());
  const uint64_t* off = g.off.data();
  for (size_t s = 0; s < kept.size(); ++s) {
    const uint64_t x = kept[s];
    const u32 r = u32(std::upper_bound(off, off + g.n + 1, x) - off - 1);
    sp_off[size_t(r) + 1]++;
    sp_col[s] = g.col[x];
  }
  for (size_t r = 0; r < g.n; ++r) sp_off[r + 1] += sp_off[r];
}

// Exact dominance counts over the sample: query(i, j) = #samples with
// row < i and col < j, the value sarma::sparse_prefix_sum::query returns
// (sparse_prefix_sum.hpp:318-322). A Fenwick tree over sample positions
// (row-major order) with a sorted column list per node.
class PrefixCounter {
 public:
  PrefixCounter(const std::vector<u32>& off, const std::vector<u32>& col) : off_(off) {
    S_ = col.size();
    node_start_.assign(S_ + 2, 0);
    for (size_t k = 1; k <= S_; ++k) node_start_[k + 1] = node_start_[k] + (k & (~k + 1));
    data_.resize(node_start_[S_ + 1]);
    for (size_t k = 1; k <= S_; ++k) {
      const size_t len = k & (~k + 1);
      u32* dst = data_.data() + node_start_[k];
      for (size_t q = 0; q < len; ++q) dst[q] = col[k - len + q];
      std::sort(dst, dst + len);
    }
  }
  u32 total() const { return u32(S_); }
  // Number of samples with row < i and col < j.
  u32 query(u32 i, u32 j) const {
    const size_t rows = off_.size() - 1;
    size_t r = off_[std::min<size_t>(i, rows)];
    u32 s = 0;
    for (; r > 0; r &= r - 1) {
      const u32* b = data_.data() + node_start_[r];
      const u32* e = data_.data() + node_start_[r + 1];
      s += u32(std::lower_bound(b, e, j) - b);
    }
    return s;
  }

 private:
  const std::vector<u32>& off_;
  size_t S_;
  std::vector<size_t> node_start_;
  std::vector<u32> data_;
};

// Max tile load over one row band (col == false) or one column band
// (col == true) of the cut vector p (sparse_prefix_sum.hpp:351-387).
u32 band_maxload(const PrefixCounter& pc, const std::vector<u32>& p, size_t idx, bool col) {
  u32 best = 0;
  if (col) {
    const u32 q0 = p[idx - 1], q1 = p[idx];
    u32 prev0 = pc.query(p[0], q0), prev1 = pc.query(p[0], q1);
    for (size_t a = 1; a < p.size(); ++a) {
      const u32 c0 = pc.query(p[a], q0), c1 = pc.query(p[a], q1);
      const u32 load = c1 - prev1 + prev0 - c0;
      best = std::max(best, load);
      prev0 = c0;
      prev1 = c1;
    }
  } else {
    const u32 r0 = p[idx - 1], r1 = p[idx];
    u32 prev0 = pc.query(r0, p[0]), prev1 = pc.query(r1, p[0]);
    for (size_t b = 1; b < p.size(); ++b) {
      const u32 c0 = pc.query(r0, p[b]), c1 = pc.query(r1, p[b]);
      const u32 load = c1 - c0 + prev0 - prev1;
      best = std::max(best, load);
      prev0 = c0;
      prev1 = c1;
    }
  }
  return best;
}

// probe_a_load::probe (probe_a_load.hpp:29-57) with an upper-bound vector.
std::vector<u32> pal_probe(const PrefixCounter& pc, u32 L, const std::vector<u32>& UB) {
  std::vector<u32> p;
  p.reserve(UB.size() + 1);
  p.push_back(0);
  while (p.size() < UB.size()) {
    u32 l = p.back(), r = UB[p.size()];
    while (l < r) {
      const u32 m = (l + 1) + (r - (l + 1)) / 2;
      p.push_back(m);
      const size_t idx = p.size() - 1;
      const u32 cl = std::max(band_maxload(pc, p, idx, true), band_maxload(pc, p, idx, false));
      if (cl <= L) l = m;
      else r = m - 1;
      p.pop_back();
    }
    p.push_back(l);
    if (p.back() == p[p.size() - 2]) {
      while (p.size() < UB.size()) p.push_back(p.back());
      break;
    }
  }
  return p;
}

}  // namespace

std::vector<uint32_t> partition_pal(const Csr& g, uint32_t P) {
  // keep probability get_prob(NNZ, p, p, 10.) (tools/utils.hpp:20-25), seed
  // 1091467301 (utils.hpp:271-273). The reference's generator is a
  // thread_local seeded on first use; its driver samples once per process, so
  // every call here starts a fresh stream.
  const u32 M32 = u32(g.m);
  const double keep = std::min(1.0, (P * 1.0 * 10.0) * (P * 1.0 * 10.0) / M32);
  std::vector<u32> sp_off, sp_col;
  sparsify(g, keep, 1091467301, sp_off, sp_col);
  PrefixCounter pc(sp_off, sp_col);
  // probe_a_load::partition (probe_a_load.hpp:70-89); N = sps.size() =
  // max(rows, ncols) (sparse_prefix_sum.hpp:257-259).
  const u32 N = std::max(g.n, g.ncols);
  const u32 M = pc.total();
  u32 l = (M - 1 + P * P) / P / P;
  u32 r = std::min(M, (M + P - 1) / P + N);
  std::vector<u32> UB(size_t(P) + 1, N);
  while (l + .5 < r) {
    const u32 m = l + (r - l) / 2;
    auto pr = pal_probe(pc, m, UB);
    if (pr.back() < N) l = m + 1;
    else { r = m; UB = pr; }
  }
  return pal_probe(pc, r, UB);
}

std::vector<uint32_t> partition(const Csr& g, bbtc_part kind, uint32_t p) {
  if (p < 1) fail(BBTC_ERR_INVALID, "number of cuts must be >= 1");
  std::vector<uint32_t> c;
  switch (kind) {
    case BBTC_PART_UNI: c = partition_uni(g, p); break;
    case BBTC_PART_N1D: c = partition_n1d(g, p); break;
    case BBTC_PART_PAL: c = partition_pal(g, p); break;
    case BBTC_PART_RAC: c = partition_rac(g, p); break;
    default: fail(BBTC_ERR_INVALID, "unknown partitioner");
  }
  if (c.size() != size_t(p) + 1) fail(BBTC_ERR_INTERNAL, "partitioner returned a wrong cut count");
  c[p] = g.n + 1;  // utils.hpp:288
  for (uint32_t b = 0; b < p; ++b)
    if (c[b] > c[b + 1]) fail(BBTC_ERR_INTERNAL, "partitioner returned decreasing cuts");
  return c;
}

}  // namespace bbtc
