// Text ingest: edge lists and Matrix Market files to the canonical
// degree-ordered upper-triangular CSR, written as the reference's driver
// writes it (utils.hpp:171-268).
//
// Steps, as the reference performs them:
//   1. parse (PIGO COO reader without flags: coo.impl.hpp:318-466 with the
//      FileReader grammar of pigo.impl.hpp:184-345): one entry per line,
//      (x, y) are the first two digit runs; '%'/'#' start comments where an
//      integer is searched for; the entry whose second integer ends exactly
//      at end of file is dropped; Matrix Market ids stay 1-based;
//   2. CSR rows sorted (csr.impl.hpp:102-243 then Matrix::sort);
//   3. symmetrise by merging row i of A and of A^T, dropping the diagonal
//      (utils.hpp:175-252); a value present a times in A's row and b times
//      in A^T's row appears max(a, b) times;
//   4. order by degree ascending with the reference's std::sort over
//      (id, degree) pairs compared by degree (csr_matrix.hpp:503-584), keep
//      the upper triangle, sort rows;
//   5. optionally drop duplicate entries (the reference keeps them).
#include <omp.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <utility>
#include <vector>

#include "handles.hpp"

namespace bbtc {
void set_last_error(const std::string& s);

namespace {

// A cursor over the file bytes with the reference reader's moves.
struct Reader {
  const char* d;
  const char* end;
  bool good() const { return d < end; }
  bool comment_at() const { return d < end && (*d == '%' || *d == '#'); }
  void skip_comments() {
    while (d < end && (*d == '%' || *d == '#'))
      while (d < end && *d++ != '\n') {
      }
  }
  // unsigned 32-bit, wrapping as the reference's accumulator does
  uint32_t read_int() {
    uint32_t r = 0;
    while (d < end && (*d < '0' || *d > '9')) ++d;
    while (d < end && *d >= '0' && *d <= '9') r = r * 10u + uint32_t(*d++ - '0');
    return r;
  }
  void move_to_first_int() {
    if (comment_at()) skip_comments();
    while (d < end && (*d < '0' || *d > '9')) {
      ++d;
      if (comment_at()) skip_comments();
    }
  }
  void move_to_non_int() {
    while (d < end && *d >= '0' && *d <= '9') ++d;
  }
  void move_to_next_int() {
    move_to_non_int();
    move_to_first_int();
  }
  void move_to_eol() {
    while (d < end && *d != '\n') ++d;
  }
  bool at_str(const char* s) const {
    const size_t n = std::strlen(s);
    if (d + n >= end) return false;
    return std::memcmp(d, s, n) == 0;
  }
};

struct Coo {
  std::vector<uint32_t> x, y;
  uint32_t nrows = 0, ncols = 0;
};

// Entries of [b, e): split into T chunks on line boundaries as the
// reference's reader does, each chunk parsed by one thread (a team smaller
// than T takes chunks tid, tid + team, ...).
Coo parse_entries(const char* b, const char* e) {
  const int T = std::max(1, omp_get_max_threads());
  std::vector<std::vector<uint32_t>> xs(T), ys(T);
  std::vector<uint32_t> mr(T, 0), mc(T, 0);
  const size_t size = size_t(e - b);
#pragma omp parallel num_threads(T)
  for (int tid = omp_get_thread_num(); tid < T; tid += omp_get_num_threads()) {
    Reader rs{b + (size_t(tid) * size) / size_t(T), e};
    Reader re{b + (size_t(tid + 1) * size) / size_t(T), e};
    re.move_to_eol();
    re.move_to_next_int();
    if (tid != 0) {
      rs.move_to_eol();
      rs.move_to_next_int();
    } else {
      rs.move_to_first_int();
    }
    if (re.d < rs.end) rs.end = re.d;
    uint32_t max_r = 0, max_c = 0;
    while (rs.good()) {
      const uint32_t x = rs.read_int();
      rs.move_to_next_int();
      const uint32_t y = rs.read_int();
      if (!rs.good()) break;
      rs.move_to_eol();
      rs.move_to_next_int();
      xs[tid].push_back(x);
      ys[tid].push_back(y);
      if (x > max_r) max_r = x;
      if (y > max_c) max_c = y;
    }
    mr[tid] = max_r;
    mc[tid] = max_c;
  }
  Coo c;
  size_t m = 0;
  for (int t = 0; t < T; ++t) m += xs[t].size();
  c.x.reserve(m);
  c.y.reserve(m);
  uint32_t max_r = 0, max_c = 0;
  for (int t = 0; t < T; ++t) {
    c.x.insert(c.x.end(), xs[t].begin(), xs[t].end());
    c.y.insert(c.y.end(), ys[t].begin(), ys[t].end());
    max_r = std::max(max_r, mr[t]);
    max_c = std::max(max_c, mc[t]);
  }
  c.nrows = max_r + 1;
  c.ncols = max_c + 1;
  return c;
}

Coo parse_matrix_market(const char* b, const char* e) {
  Reader r{b, e};
  if (!r.at_str("%%MatrixMarket matrix coordinate"))
    fail(BBTC_ERR_FORMAT, "Unable to handle different MatrixMarket formats");
  r.move_to_next_int();
  const uint32_t nrows = r.read_int() + 1;  // ids start at 1 and are kept
  r.move_to_next_int();
  const uint32_t ncols = r.read_int() + 1;
  r.move_to_next_int();
  const uint32_t nnz = r.read_int();
  r.move_to_eol();
  r.move_to_next_int();
  Coo c = parse_entries(r.d, e);
  if (nrows >= c.nrows) c.nrows = nrows;
  else fail(BBTC_ERR_FORMAT, "Too many row labels in file contradicting header");
  if (ncols >= c.ncols) c.ncols = ncols;
  else fail(BBTC_ERR_FORMAT, "Too many col labels in file contradicting header");
  if (uint64_t(nnz) != c.x.size()) fail(BBTC_ERR_FORMAT, "Header contradicts number of read non-zeros");
  return c;
}

bool ends_with(const std::string& s, const char* suf) {
  const size_t n = std::strlen(suf);
  return s.size() >= n && s.compare(s.size() - n, n, suf) == 0;
}

// Rows of a matrix with n rows from (row, col) pairs, each row sorted: the
// slots are claimed with atomic counters (for id spaces too sparse for
// per-thread histograms), so each row is sorted afterwards.
void rows_of(uint32_t n, const std::vector<uint32_t>& row, const std::vector<uint32_t>& col,
             std::vector<uint32_t>& ptr, std::vector<uint32_t>& idx) {
  const size_t m = row.size();
  ptr.assign(size_t(n) + 1, 0);
  idx.resize(m);
  uint32_t* cnt = ptr.data() + 1;
#pragma omp parallel for schedule(static, 1 << 16)
  for (int64_t q = 0; q < int64_t(m); ++q) __atomic_fetch_add(&cnt[row[q]], 1u, __ATOMIC_RELAXED);
  for (size_t i = 1; i < ptr.size(); ++i) ptr[i] += ptr[i - 1];
  std::vector<uint32_t> pos(ptr.begin(), ptr.end() - 1);
  uint32_t* at = pos.data();
#pragma omp parallel for schedule(static, 1 << 16)
  for (int64_t q = 0; q < int64_t(m); ++q) idx[__atomic_fetch_add(&at[row[q]], 1u, __ATOMIC_RELAXED)] = col[q];
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t i = 0; i < int64_t(n); ++i) std::sort(idx.begin() + ptr[i], idx.begin() + ptr[i + 1]);
}

// Whether per-thread histograms over n keys stay within the memory of the
// m pairs they place (a much sparser id space takes rows_of's atomics).
bool histograms_fit(uint32_t n, size_t m, int T) {
  return size_t(T) * n <= 2 * m + (size_t(1) << 24);
}

// Rows over n keys from m (key, value) pairs in a fixed order, each row
// keeping that order (a stable counting scatter). The pairs come in T
// consecutive parts; part(t, f) calls f(key, value) for the pairs of part t
// in order. Part t is counted into a histogram of its own and placed after
// the parts before it; thread tid of the team (which may be smaller than T)
// takes parts tid, tid + team, ...
template <class Part>
void stable_rows(uint32_t n, size_t m, int T, Part part, std::vector<uint32_t>& ptr,
                 std::vector<uint32_t>& idx) {
  ptr.assign(size_t(n) + 1, 0);
  idx.resize(m);
  std::vector<uint32_t> hist(size_t(T) * n, 0);
#pragma omp parallel num_threads(T)
  {
    const int team = omp_get_num_threads();
    for (int t = omp_get_thread_num(); t < T; t += team) {
      uint32_t* h = hist.data() + size_t(t) * n;
      part(t, [&](uint32_t key, uint32_t) { ++h[key]; });
    }
#pragma omp barrier
#pragma omp for schedule(static)
    for (int64_t i = 0; i < int64_t(n); ++i) {
      uint32_t k = 0;
      for (int u = 0; u < T; ++u) k += hist[size_t(u) * n + size_t(i)];
      ptr[size_t(i) + 1] = k;
    }
#pragma omp single
    for (size_t i = 1; i < ptr.size(); ++i) ptr[i] += ptr[i - 1];
#pragma omp for schedule(static)
    for (int64_t i = 0; i < int64_t(n); ++i) {
      uint32_t at = ptr[size_t(i)];
      for (int u = 0; u < T; ++u) {
        uint32_t& c = hist[size_t(u) * n + size_t(i)];
        const uint32_t k = c;
        c = at;
        at += k;
      }
    }
    for (int t = omp_get_thread_num(); t < T; t += team) {
      uint32_t* h = hist.data() + size_t(t) * n;
      part(t, [&](uint32_t key, uint32_t val) { idx[h[key]++] = val; });
    }
  }
}

// The pairs of a CSR in row order (key: the entry, value: its row), cut into
// T parts of about equal entries.
struct CsrPairs {
  const std::vector<uint32_t>& ptr;
  const std::vector<uint32_t>& idx;
  std::vector<uint32_t> at;  // T + 1 row bounds
  CsrPairs(const std::vector<uint32_t>& p, const std::vector<uint32_t>& i, int T) : ptr(p), idx(i), at(size_t(T) + 1) {
    const uint32_t n = uint32_t(ptr.size() - 1);
    for (int t = 0; t <= T; ++t) {
      const uint64_t want = uint64_t(idx.size()) * uint64_t(t) / uint64_t(T);
      at[size_t(t)] = t == T ? n : uint32_t(std::lower_bound(ptr.begin(), ptr.end() - 1, want) - ptr.begin());
    }
  }
  template <class F>
  void operator()(int t, F f) const {
    for (uint32_t r = at[size_t(t)]; r < at[size_t(t) + 1]; ++r)
      for (uint32_t q = ptr[r]; q < ptr[size_t(r) + 1]; ++q) f(idx[q], r);
  }
};

}  // namespace

// Canonical upper-triangular CSR of a text graph file.
Csr csr_from_text(const std::string& path, bool dedup) {
  if (ends_with(path, ".graph"))
    fail(BBTC_ERR_FORMAT, "METIS .graph input is not supported; convert it to an edge list");
  std::ifstream f(path, std::ios::binary | std::ios::ate);
  if (!f) fail(BBTC_ERR_IO, "cannot open " + path);
  const std::streamoff size = f.tellg();
  if (size < 0) fail(BBTC_ERR_IO, "cannot read " + path);
  std::vector<char> bytes(static_cast<size_t>(size));
  f.seekg(0);
  if (size && !f.read(bytes.data(), size)) fail(BBTC_ERR_IO, "cannot read " + path);
  const char* b = bytes.data();
  const char* e = b + bytes.size();
  if (bytes.size() >= 4 && std::memcmp(b, "PIGO", 4) == 0)
    fail(BBTC_ERR_FORMAT, "PIGO binary input other than .bbtc is not supported");
  Coo c = ends_with(path, ".mtx") ? parse_matrix_market(b, e) : parse_entries(b, e);
  std::vector<char>().swap(bytes);
  const uint32_t n = std::max(c.nrows, c.ncols);
  if (c.x.size() >= (1ull << 32)) fail(BBTC_ERR_FORMAT, "more than 2^32-1 entries");

  // A and A^T as sorted rows over n x n. With histograms: three stable
  // scatters, by y, then by x, then by y, sort the rows of both (a two-key
  // radix sort); otherwise each row is sorted after an atomic scatter.
  std::vector<uint32_t> ap, ai, tp, ti;
  const int T = std::max(1, omp_get_max_threads());
  const size_t m = c.x.size();
  if (histograms_fit(n, m, T)) {
    auto coo = [&](int t, auto f) {
      const size_t q0 = m * size_t(t) / size_t(T), q1 = m * size_t(t + 1) / size_t(T);
      for (size_t q = q0; q < q1; ++q) f(c.y[q], c.x[q]);
    };
    stable_rows(n, m, T, coo, tp, ti);  // rows of y in input order
    c = Coo{};
    stable_rows(n, m, T, CsrPairs(tp, ti, T), ap, ai);  // A: rows of x, y ascending
    stable_rows(n, m, T, CsrPairs(ap, ai, T), tp, ti);  // A^T: rows of y, x ascending
  } else {
    rows_of(n, c.x, c.y, ap, ai);
    rows_of(n, c.y, c.x, tp, ti);
    c = Coo{};
  }

  // symmetric merge without the diagonal
  std::vector<uint32_t> sp(size_t(n) + 1, 0);
  auto merge_row = [&](uint32_t i, uint32_t* out) -> uint32_t {
    uint32_t j = ap[i], k = tp[i], cnt = 0;
    const uint32_t je = ap[i + 1], ke = tp[i + 1];
    auto put = [&](uint32_t v) {
      if (v != i) {
        if (out) out[cnt] = v;
        ++cnt;
      }
    };
    while (j < je && k < ke) {
      if (ai[j] == ti[k]) {
        put(ai[j++]);
        ++k;
      } else if (ai[j] < ti[k]) {
        put(ai[j++]);
      } else {
        put(ti[k++]);
      }
    }
    while (j < je) put(ai[j++]);
    while (k < ke) put(ti[k++]);
    return cnt;
  };
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t i = 0; i < int64_t(n); ++i) sp[size_t(i) + 1] = merge_row(uint32_t(i), nullptr);
  for (size_t i = 1; i < sp.size(); ++i) {
    if (uint64_t(sp[i]) + sp[i - 1] >= (1ull << 32)) fail(BBTC_ERR_FORMAT, "more than 2^32-1 entries");
    sp[i] += sp[i - 1];
  }
  std::vector<uint32_t> si(sp[n]);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t i = 0; i < int64_t(n); ++i) merge_row(uint32_t(i), si.data() + sp[i]);
  std::vector<uint32_t>().swap(ai);
  std::vector<uint32_t>().swap(ti);

  // degree order: the same std::sort over the same (id, degree) pairs
  std::vector<std::pair<uint32_t, uint32_t>> deg(n);
  for (uint32_t i = 0; i < n; ++i) deg[i] = {i, sp[i + 1] - sp[i]};
  std::sort(deg.begin(), deg.end(), [](const auto& u, const auto& v) { return u.second < v.second; });
  std::vector<uint32_t> map(n);
  for (uint32_t i = 0; i < n; ++i) map[deg[i].first] = i;
  std::vector<std::pair<uint32_t, uint32_t>>().swap(deg);

  // upper triangle in the new ids
  std::vector<uint64_t> cnt(size_t(n) + 1, 0);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t i = 0; i < int64_t(n); ++i) {
    uint64_t k = 0;
    for (uint32_t q = sp[i]; q < sp[i + 1]; ++q) k += map[si[q]] > map[i];
    cnt[size_t(map[i]) + 1] = k;
  }
  for (size_t i = 1; i < cnt.size(); ++i) cnt[i] += cnt[i - 1];
  std::vector<uint32_t> ui(cnt[n]);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t i = 0; i < int64_t(n); ++i) {
    uint64_t st = cnt[map[i]];
    const uint64_t b0 = st;
    for (uint32_t q = sp[i]; q < sp[i + 1]; ++q)
      if (map[si[q]] > map[i]) ui[st++] = map[si[q]];
    std::sort(ui.begin() + b0, ui.begin() + st);
  }

  Csr g;
  g.n = n;
  g.ncols = n;
  if (!dedup) {
    g.m = ui.size();
    g.off.resize(size_t(n) + 1);
    g.col.resize(std::max<size_t>(ui.size(), 1));
    for (size_t i = 0; i <= n; ++i) g.off[i] = cnt[i];
    std::copy(ui.begin(), ui.end(), g.col.data());
    return g;
  }
  std::vector<uint64_t> kept(size_t(n) + 1, 0);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t r = 0; r < int64_t(n); ++r) {
    uint64_t k = 0;
    for (uint64_t q = cnt[r]; q < cnt[r + 1]; ++q) k += (q == cnt[r] || ui[q] != ui[q - 1]);
    kept[size_t(r) + 1] = k;
  }
  for (size_t i = 1; i < kept.size(); ++i) kept[i] += kept[i - 1];
  g.m = kept[n];
  g.off.resize(size_t(n) + 1);
  g.col.resize(std::max<uint64_t>(g.m, 1));
  for (size_t i = 0; i <= n; ++i) g.off[i] = kept[i];
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t r = 0; r < int64_t(n); ++r) {
    uint64_t o = kept[r];
    for (uint64_t q = cnt[r]; q < cnt[r + 1]; ++q)
      if (q == cnt[r] || ui[q] != ui[q - 1]) g.col[o++] = ui[q];
  }
  return g;
}

}  // namespace bbtc
