// Block CSR layout, task list and algorithmic-byte accounting.
//
// The reference layout (block_csr.cpp:286-376) keeps (N+P)*P global 32-bit
// row pointers, lower tiles included, and every task rescans all h_i rows of
// its band. This layout stores only tiles (i,k) with i <= k, 32-bit
// tile-local row offsets, 128-byte aligned column segments and a compact list
// of non-empty rows per tile, so a task visits only rows that have entries.
#include <omp.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <vector>

#include "graph.hpp"

namespace bbtc {

namespace {
constexpr uint64_t kAlignEntries = 32;  // 128 bytes of uint32
inline uint64_t align_up(uint64_t x) { return (x + kAlignEntries - 1) & ~(kAlignEntries - 1); }
}  // namespace

// Bitmap form of the tiles whose column band is at most kDenseMaxBits wide,
// smallest tiles first until the total cap: only the placement; the rows are
// filled on the device from the uploaded tiles (bbtc_graph_upload), so they
// never cross PCIe.
static void build_dense(Layout& L) {
  const uint32_t P = L.P;
  L.dense_base.assign(size_t(P) * P, ~0ull);
  L.dwpr.assign(P, 0);
  for (uint32_t k = 0; k < P; ++k) L.dwpr[k] = (((L.height(k) + 31) / 32) + 3) & ~3u;
  std::vector<std::pair<uint64_t, size_t>> cand;
  for (uint32_t i = 0; i < P; ++i)
    for (uint32_t k = i; k < P; ++k) {
      const uint64_t words = uint64_t(L.height(i)) * L.dwpr[k];
      if (L.height(k) == 0 || L.height(k) > kDenseMaxBits || words > kDenseTileWords) continue;
      if (L.nnz[L.tile(i, k)] == 0 || words > kDenseWordsPerNnz * L.nnz[L.tile(i, k)]) continue;
      cand.emplace_back(words, L.tile(i, k));
    }
  std::sort(cand.begin(), cand.end());
  uint64_t total = 0;
  for (auto& c : cand) {
    if (total + c.first > kDenseTotalWords) break;
    L.dense_base[c.second] = total;
    total += c.first;
  }
  L.dense_words = std::max<uint64_t>(total, 4);
}

Layout layout_frame(uint32_t n, uint64_t m, const std::vector<uint32_t>& cuts) {
  if (cuts.size() < 2) fail(BBTC_ERR_INVALID, "need at least one band");
  const uint32_t P = uint32_t(cuts.size() - 1);
  if (cuts[0] != 0) fail(BBTC_ERR_INVALID, "cuts[0] must be 0");
  for (uint32_t b = 0; b < P; ++b)
    if (cuts[b] > cuts[b + 1]) fail(BBTC_ERR_INVALID, "cuts must not decrease");
  if (cuts[P] < n) fail(BBTC_ERR_INVALID, "cuts[P] must cover every vertex");

  Layout L;
  L.P = P;
  L.n = n;
  L.m = m;
  L.cuts = cuts;
  const size_t T = size_t(P) * P;
  L.rp_base.assign(T, ~0ull);
  L.col_base.assign(T, ~0ull);
  L.nz_base.assign(T, ~0ull);
  L.nnz.assign(T, 0);
  L.nz_rows.assign(T, 0);
  L.max_row.assign(T, 0);
  L.deg.assign(T, 0.f);
  uint64_t rp_total = 0;
  for (uint32_t i = 0; i < P; ++i)
    for (uint32_t k = i; k < P; ++k) {
      L.rp_base[L.tile(i, k)] = rp_total;
      rp_total += uint64_t(L.height(i)) + 1;
    }
  L.rp_words = rp_total;
  return L;
}

void layout_place(Layout& L) {
  const uint32_t P = L.P;
  uint64_t col_total = 0, nz_total = 0;
  for (uint32_t i = 0; i < P; ++i)
    for (uint32_t k = i; k < P; ++k) {
      const size_t t = L.tile(i, k);
      L.col_base[t] = col_total;
      col_total = align_up(col_total + L.nnz[t]);
      L.nz_base[t] = nz_total;
      nz_total += L.nz_rows[t];
    }
  L.col_words = std::max<uint64_t>(col_total, 1);
  L.nz_words = std::max<uint64_t>(nz_total, 1);
  for (uint32_t i = 0; i < P; ++i)
    for (uint32_t k = 0; k < P; ++k) {
      const size_t t = L.tile(i, k);
      // block_csr.cpp:272-278: 32-bit sum over the rows, then float division
      const uint32_t nnz32 = uint32_t(L.nnz[t]);
      L.deg[t] = static_cast<float>(nnz32) / static_cast<float>(L.height(i));
    }
  build_dense(L);
}

void require_host_arrays(const Layout& L, const char* what) {
  if (!L.host_arrays)
    fail(BBTC_ERR_INVALID, std::string(what) +
                               " reads the layout's arrays, which a layout built on the device without "
                               "BBTC_BUILD_HOST_ARRAYS does not hold on the host");
}

void parallel_copy(void* dst, const void* src, size_t bytes) {
  constexpr size_t kPiece = 1 << 20;
  const int64_t pieces = int64_t((bytes + kPiece - 1) / kPiece);
  // pieces taken dynamically: a copy made from a thread other than the
  // caller's main one (the server's upload thread) shares the cores with
  // other busy threads, and a statically split copy waits for its slowest
  // thread
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t q = 0; q < pieces; ++q) {
    const size_t o = size_t(q) * kPiece;
    std::memcpy(static_cast<char*>(dst) + o, static_cast<const char*>(src) + o, std::min(kPiece, bytes - o));
  }
}

uint64_t column_end(const Csr& g) {
  uint64_t mx = 0;
#pragma omp parallel for reduction(max : mx) schedule(static, 1 << 16)
  for (int64_t u = 0; u < int64_t(g.n); ++u)
    if (g.off[u + 1] > g.off[u]) mx = std::max<uint64_t>(mx, uint64_t(g.col[g.off[u + 1] - 1]) + 1);
  return mx;
}

uint32_t longest_row(const Csr& g) {
  uint64_t best = 0;
#pragma omp parallel for schedule(static, 1 << 16) reduction(max : best)
  for (int64_t u = 0; u < int64_t(g.n); ++u) best = std::max<uint64_t>(best, g.off[u + 1] - g.off[u]);
  return uint32_t(std::min<uint64_t>(best, 0xFFFFFFFFull));
}

Layout build_layout(const Csr& g, const std::vector<uint32_t>& cuts) {
  Layout L = layout_frame(g.n, g.m, cuts);
  L.longest_row = longest_row(g);
  L.longest_known = true;
  const uint32_t P = L.P;
  if (column_end(g) > cuts[P]) fail(BBTC_ERR_INVALID, "a column lies at or beyond cuts[P]");
  const uint64_t rp_total = L.rp_words;
  L.rowptr.resize(rp_total);
  uint32_t* rp = L.rowptr.data();
#pragma omp parallel for schedule(static, 1 << 16)
  for (int64_t x = 0; x < int64_t(rp_total); ++x) rp[x] = 0;

  // band of every row (the non-empty band that holds it)
  auto band_of = [&](uint32_t v) {
    return uint32_t(std::upper_bound(cuts.begin(), cuts.end(), v) - cuts.begin() - 1);
  };

  // pass 1: per (tile, local row) entry counts into rowptr[r + 1]
#pragma omp parallel for schedule(dynamic, 4096)
  for (int64_t u = 0; u < int64_t(g.n); ++u) {
    const uint64_t b = g.off[u], e = g.off[u + 1];
    if (b == e) continue;
    const uint32_t i = band_of(uint32_t(u));
    const uint32_t r = uint32_t(u) - cuts[i];
    uint32_t k = band_of(g.col[b]);
    uint64_t x = b;
    while (x < e) {
      const uint32_t hi = cuts[k + 1];
      uint64_t y = x;
      while (y < e && g.col[y] < hi) ++y;
      if (y > x) rp[L.rp_base[L.tile(i, k)] + r + 1] = uint32_t(y - x);
      x = y;
      ++k;
    }
  }

  // pass 2: per-tile scan, sizes and non-empty row counts
  std::vector<std::pair<uint32_t, uint32_t>> upper;
  for (uint32_t i = 0; i < P; ++i)
    for (uint32_t k = i; k < P; ++k) upper.emplace_back(i, k);
  int overflow = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(| : overflow)
  for (int64_t q = 0; q < int64_t(upper.size()); ++q) {
    const uint32_t i = upper[q].first, k = upper[q].second;
    const size_t t = L.tile(i, k);
    uint32_t* row = rp + L.rp_base[t];
    const uint32_t h = L.height(i);
    uint64_t acc = 0;
    uint32_t nzr = 0, longest = 0;
    for (uint32_t r = 0; r <= h; ++r) {
      const uint64_t c = row[r];
      nzr += (r > 0 && c > 0) ? 1u : 0u;
      longest = std::max(longest, uint32_t(c));
      acc += c;
      if (acc > 0xFFFFFFFFull) { overflow = 1; break; }
      row[r] = uint32_t(acc);
    }
    L.nnz[t] = acc;
    L.nz_rows[t] = nzr;
    L.max_row[t] = longest;
  }
  if (overflow) fail(BBTC_ERR_INVALID, "a tile holds 2^32 or more entries; use more cuts");

  layout_place(L);
  L.cols.resize(L.col_words);
  L.nzrows.resize(L.nz_words);
  // a layout with no entries keeps one zero word in each array
  L.cols[0] = 0;
  L.nzrows[0] = 0;
  uint32_t* cl = L.cols.data();
  // zero the alignment padding so the array content is deterministic
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t q = 0; q < int64_t(upper.size()); ++q) {
    const size_t t = L.tile(upper[q].first, upper[q].second);
    const uint64_t end = L.col_base[t] + L.nnz[t];
    const uint64_t pad = align_up(end);
    for (uint64_t x = end; x < pad && x < L.cols.size(); ++x) cl[x] = 0;
  }

  // pass 3: scatter tile-local columns
#pragma omp parallel for schedule(dynamic, 4096)
  for (int64_t u = 0; u < int64_t(g.n); ++u) {
    const uint64_t b = g.off[u], e = g.off[u + 1];
    if (b == e) continue;
    const uint32_t i = band_of(uint32_t(u));
    const uint32_t r = uint32_t(u) - cuts[i];
    uint32_t k = band_of(g.col[b]);
    uint64_t x = b;
    while (x < e) {
      const uint32_t lo = cuts[k], hi = cuts[k + 1];
      const size_t t = L.tile(i, k);
      uint32_t* dst = cl + L.col_base[t] + rp[L.rp_base[t] + r];
      while (x < e && g.col[x] < hi) *dst++ = g.col[x++] - lo;
      ++k;
    }
  }

  // pass 4: non-empty row lists and average degrees
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t q = 0; q < int64_t(upper.size()); ++q) {
    const uint32_t i = upper[q].first, k = upper[q].second;
    const size_t t = L.tile(i, k);
    const uint32_t* row = rp + L.rp_base[t];
    uint32_t* nz = L.nzrows.data() + L.nz_base[t];
    const uint32_t h = L.height(i);
    uint32_t w = 0;
    for (uint32_t r = 0; r < h; ++r)
      if (row[r + 1] > row[r]) nz[w++] = r;
  }
  return L;
}

std::vector<bbtc_task> task_list(uint32_t p) {
  std::vector<bbtc_task> t;
  t.reserve(size_t(p) * (p + 1) * (p + 2) / 6);
  for (uint32_t i = 0; i < p; ++i)
    for (uint32_t j = i; j < p; ++j)
      for (uint32_t k = j; k < p; ++k) t.push_back(bbtc_task{i, j, k});
  return t;
}

void sort_tasks(std::vector<bbtc_task>& t, const Layout& L, bbtc_order order) {
  if (order == BBTC_ORDER_LEX) return;
  // tasks.hpp:21-23: nnz(ab) * max(deg(bc), deg(ac)), float arithmetic.
  auto weight = [&](const bbtc_task& x) {
    const uint32_t nnz_ab = uint32_t(L.nnz[L.tile(x.i, x.j)]);
    return nnz_ab * std::max(L.deg[L.tile(x.j, x.k)], L.deg[L.tile(x.i, x.k)]);
  };
  if (order == BBTC_ORDER_DESC)
    std::sort(t.begin(), t.end(), [&](const bbtc_task& l, const bbtc_task& r) { return weight(l) > weight(r); });
  else
    std::sort(t.begin(), t.end(), [&](const bbtc_task& l, const bbtc_task& r) { return weight(l) < weight(r); });
}

bool task_skipped(const Layout& L, const bbtc_task& x) {
  // bb_tc.cpp:86-91 (and 124-129): float compare, so a NaN degree does not skip
  return L.deg[L.tile(x.j, x.k)] == 0 || L.deg[L.tile(x.i, x.k)] == 0 ||
         L.deg[L.tile(x.i, x.j)] == 0;
}

void task_bytes(const Layout& L, const bbtc_task* tasks, uint64_t n, uint32_t o, uint64_t* per_task,
                uint64_t* total, uint64_t* wedges) {
  require_host_arrays(L, "bbtc_tasks_bytes");
  uint64_t tot = 0, wtot = 0;
  const uint32_t* rp = L.rowptr.data();
  const uint32_t* cl = L.cols.data();
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : tot, wtot)
  for (int64_t q = 0; q < int64_t(n); ++q) {
    const bbtc_task x = tasks[q];
    uint64_t B = 0, W = 0;
    if (!task_skipped(L, x)) {
      const size_t ab = L.tile(x.i, x.j), bc = L.tile(x.j, x.k), ac = L.tile(x.i, x.k);
      const uint32_t* ab_rp = rp + L.rp_base[ab];
      const uint32_t* ac_rp = rp + L.rp_base[ac];
      const uint32_t* bc_rp = rp + L.rp_base[bc];
      const uint32_t* ab_c = cl + L.col_base[ab];
      const uint32_t* nz = L.nzrows.data() + L.nz_base[ab];
      for (uint32_t z = 0; z < L.nz_rows[ab]; ++z) {
        const uint32_t r = nz[z];
        const uint64_t dac = ac_rp[r + 1] - ac_rp[r];
        if (!dac) continue;
        const uint32_t a0 = ab_rp[r], a1 = ab_rp[r + 1];
        B += (2ull * o + 4ull * (a1 - a0)) + (2ull * o + 4ull * dac);
        W += a1 - a0;
        for (uint32_t y = a0; y < a1; ++y) {
          const uint32_t v = ab_c[y];
          B += 2ull * o + 4ull * (bc_rp[v + 1] - bc_rp[v]);
        }
      }
    }
    if (per_task) per_task[q] = B;
    tot += B;
    wtot += W;
  }
  if (total) *total = tot;
  if (wedges) *wedges = wtot;
}

void task_stats(const Layout& L, const bbtc_task* tasks, uint64_t n, uint64_t* out, uint32_t sample_units) {
  require_host_arrays(L, "bbtc_tasks_stats");
  const uint32_t* rp = L.rowptr.data();
  const uint32_t* cl = L.cols.data();
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t q = 0; q < int64_t(n); ++q) {
    const bbtc_task x = tasks[q];
    double R = 0, W = 0, A = 0, Pbc = 0, Pac = 0, Lac = 0;
    if (!task_skipped(L, x)) {
      // 128-byte lines of a bitmap row of band k
      const uint64_t lines = (uint64_t(L.dwpr[x.k]) + 31) / 32;
      const size_t ab = L.tile(x.i, x.j), bc = L.tile(x.j, x.k), ac = L.tile(x.i, x.k);
      const uint32_t* ab_rp = rp + L.rp_base[ab];
      const uint32_t* ac_rp = rp + L.rp_base[ac];
      const uint32_t* bc_rp = rp + L.rp_base[bc];
      const uint32_t* ab_c = cl + L.col_base[ab];
      const uint32_t* nz = L.nzrows.data() + L.nz_base[ab];
      // rows in units of kStatUnitRows consecutive non-empty rows; with
      // sample_units > 0, at most that many units spread evenly over the list
      const uint32_t nzr = L.nz_rows[ab];
      const uint64_t units = (uint64_t(nzr) + kStatUnitRows - 1) / kStatUnitRows;
      const uint64_t take = sample_units && units > sample_units ? sample_units : units;
      uint64_t seen = 0;
      double r_acc = 0, w_acc = 0, a_acc = 0, pbc_acc = 0, pac_acc = 0, lac_acc = 0;
      for (uint64_t m = 0; m < take; ++m) {
        const uint64_t unit = take == units ? m : (2 * m + 1) * units / (2 * take);
        const uint32_t z0 = uint32_t(unit * kStatUnitRows);
        const uint32_t z1 = std::min<uint32_t>(nzr, z0 + kStatUnitRows);
        seen += z1 - z0;
        for (uint32_t z = z0; z < z1; ++z) {
          const uint32_t r = nz[z];
          const uint64_t dac = ac_rp[r + 1] - ac_rp[r];
          if (!dac) continue;
          const uint32_t a0 = ab_rp[r], a1 = ab_rp[r + 1];
          uint64_t pbc = 0;
          for (uint32_t y = a0; y < a1; ++y) {
            const uint32_t v = ab_c[y];
            pbc += bc_rp[v + 1] - bc_rp[v];
          }
          r_acc += 1;
          w_acc += a1 - a0;
          a_acc += double(dac);
          pbc_acc += double(pbc);
          pac_acc += double(uint64_t(a1 - a0) * dac);
          lac_acc += double(uint64_t(a1 - a0) * std::min(dac, lines));
        }
      }
      // scaled from the rows seen to all rows of the list
      const double f = seen ? double(nzr) / double(seen) : 0.0;
      R = r_acc * f;
      W = w_acc * f;
      A = a_acc * f;
      Pbc = pbc_acc * f;
      Pac = pac_acc * f;
      Lac = lac_acc * f;
    }
    uint64_t* o = out + kTaskStats * q;
    o[0] = uint64_t(std::llround(R));
    o[1] = uint64_t(std::llround(W));
    o[2] = uint64_t(std::llround(A));
    o[3] = uint64_t(std::llround(Pbc));
    o[4] = uint64_t(std::llround(Pac));
    o[5] = uint64_t(std::llround(Lac));
  }
}

}  // namespace bbtc
