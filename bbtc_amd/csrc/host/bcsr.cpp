// Block CSR files (.bcsr), the reference's serialised BlockCSR
// (block_csr.cpp:215-255): uint32 N, E, P; cuts[P+1]; sparse_tiles[P*P+1];
// rows[N*P]; cols[E]. Tiles are in row-major (i, k) order over the full
// P x P grid; tile t = (i, k) owns h_i = cuts[i+1] - cuts[i] consecutive
// slots of rows starting at sparse_tiles[t], each the global offset in cols
// of that local row's first entry; cols hold tile-local column ids.
//
// The tiles' row slots are N*P words (cuts[P] = N), and the end of the last
// tile's last row is one slot further. The reference's reader allocates and
// reads (header N) * P words, and its counters read that end slot, so the
// file announces N + 1 in the header: N*P slots, the end slot (E), and P - 1
// more words of E, which the reader takes and nothing uses (its writer side
// does not exist). Files whose header holds N itself, with the end slot
// implied as E, are read too. Reading rebuilds the CSR and this library's
// tile-major layout from the stored cuts.
#include <cstdio>
#include <cstring>
#include <vector>

#include "handles.hpp"

namespace bbtc {
void set_last_error(const std::string& s);

namespace {

struct FileCloser {
  std::FILE* f = nullptr;
  ~FileCloser() {
    if (f) std::fclose(f);
  }
};

void put(std::FILE* f, const void* p, size_t bytes, const std::string& path) {
  if (bytes && std::fwrite(p, 1, bytes, f) != bytes) fail(BBTC_ERR_IO, "write failed for " + path);
}

void get(std::FILE* f, void* p, size_t bytes, const std::string& path) {
  if (bytes && std::fread(p, 1, bytes, f) != bytes) fail(BBTC_ERR_FORMAT, "truncated .bcsr file " + path);
}

}  // namespace

void save_bcsr(const Layout& L, const std::string& path) {
  require_host_arrays(L, "bbtc_layout_save_bcsr");
  const uint32_t P = L.P, N = L.n;
  if (L.m >= (1ull << 32)) fail(BBTC_ERR_INVALID, ".bcsr holds at most 2^32-1 entries");
  if ((uint64_t(N) + 1) * P >= (1ull << 32)) fail(BBTC_ERR_INVALID, ".bcsr holds at most 2^32-1 row slots");
  std::vector<uint32_t> cuts(L.cuts.begin(), L.cuts.end());
  cuts[P] = N;
  std::vector<uint32_t> tiles(size_t(P) * P + 1, 0);
  for (uint32_t i = 0; i < P; ++i)
    for (uint32_t k = 0; k < P; ++k) tiles[size_t(i) * P + k + 1] = tiles[size_t(i) * P + k] + (cuts[i + 1] - cuts[i]);
  std::vector<uint32_t> rows((size_t(N) + 1) * P, uint32_t(L.m));  // the slots past N*P: E
  std::vector<uint32_t> cols(L.m);
  uint64_t at = 0;  // entries written before the current tile
  for (uint32_t i = 0; i < P; ++i)
    for (uint32_t k = 0; k < P; ++k) {
      const size_t t = L.tile(i, k);
      const uint32_t h = cuts[i + 1] - cuts[i];
      const bool stored = i <= k && L.nnz[t] > 0;
      for (uint32_t r = 0; r < h; ++r)
        rows[tiles[t] + r] = uint32_t(at + (stored ? L.rowptr[L.rp_base[t] + r] : 0));
      if (stored) {
        std::memcpy(cols.data() + at, L.cols.data() + L.col_base[t], L.nnz[t] * 4);
        at += L.nnz[t];
      }
    }
  FileCloser fc;
  fc.f = std::fopen(path.c_str(), "wb");
  if (!fc.f) fail(BBTC_ERR_IO, "cannot create " + path);
  const uint32_t E = uint32_t(L.m);
  const uint32_t Nfile = N + 1;
  put(fc.f, &Nfile, 4, path);
  put(fc.f, &E, 4, path);
  put(fc.f, &P, 4, path);
  put(fc.f, cuts.data(), cuts.size() * 4, path);
  put(fc.f, tiles.data(), tiles.size() * 4, path);
  put(fc.f, rows.data(), rows.size() * 4, path);
  put(fc.f, cols.data(), cols.size() * 4, path);
  if (std::fflush(fc.f) != 0) fail(BBTC_ERR_IO, "flush failed for " + path);
}

void load_bcsr(const std::string& path, Csr& g, std::vector<uint32_t>& cuts_out) {
  FileCloser fc;
  fc.f = std::fopen(path.c_str(), "rb");
  if (!fc.f) fail(BBTC_ERR_IO, "cannot open " + path);
  uint32_t N = 0, E = 0, P = 0;
  get(fc.f, &N, 4, path);
  get(fc.f, &E, 4, path);
  get(fc.f, &P, 4, path);
  if (P < 1 || P > 65535) fail(BBTC_ERR_FORMAT, "bad number of cuts in " + path);
  if (uint64_t(N) * P >= (1ull << 32)) fail(BBTC_ERR_FORMAT, "too many row slots in " + path);
  // N here is the header's; the graph has cuts[P] vertices (N, or N - 1 when
  // the file carries the end slot)
  // the arrays the header announces must be in the file before they are
  // allocated, so a damaged header fails at once instead of after a large
  // allocation
  const uint64_t need = 4 * ((uint64_t(P) + 1) + (uint64_t(P) * P + 1) + uint64_t(N) * P + E);
  if (std::fseek(fc.f, 0, SEEK_END) != 0) fail(BBTC_ERR_IO, "cannot read " + path);
  const long size = std::ftell(fc.f);
  if (size < 0 || std::fseek(fc.f, 12, SEEK_SET) != 0) fail(BBTC_ERR_IO, "cannot read " + path);
  if (uint64_t(size) - 12 < need) fail(BBTC_ERR_FORMAT, "truncated .bcsr file " + path);
  std::vector<uint32_t> cuts(size_t(P) + 1), tiles(size_t(P) * P + 1), rows(size_t(N) * P), cols(E);
  get(fc.f, cuts.data(), cuts.size() * 4, path);
  get(fc.f, tiles.data(), tiles.size() * 4, path);
  get(fc.f, rows.data(), rows.size() * 4, path);
  get(fc.f, cols.data(), cols.size() * 4, path);
  if (cuts[0] != 0 || (cuts[P] != N && uint64_t(cuts[P]) + 1 != N))
    fail(BBTC_ERR_FORMAT, "cuts must run from 0 to N or N - 1 in " + path);
  const bool end_slot = cuts[P] != N;  // N*P slots of rows, then the end slot and padding
  N = cuts[P];
  const uint64_t used = uint64_t(N) * P;
  if (end_slot)
    for (uint64_t s = used; s < rows.size(); ++s)
      if (rows[s] != E) fail(BBTC_ERR_FORMAT, "end slot is not E in " + path);
  for (uint32_t b = 0; b < P; ++b)
    if (cuts[b] > cuts[b + 1]) fail(BBTC_ERR_FORMAT, "decreasing cuts in " + path);
  for (uint32_t i = 0; i < P; ++i)
    for (uint32_t k = 0; k < P; ++k) {
      const size_t t = size_t(i) * P + k;
      if (tiles[t + 1] - tiles[t] != cuts[i + 1] - cuts[i] || (t == 0 && tiles[0] != 0))
        fail(BBTC_ERR_FORMAT, "tile slots do not match the cuts in " + path);
    }
  // slot s ends where slot s + 1 starts; the last slot ends at E
  auto slot_end = [&](uint64_t s) { return s + 1 < used ? rows[s + 1] : E; };
  for (uint64_t s = 0; s < used; ++s)
    if (rows[s] > slot_end(s) || slot_end(s) > E) fail(BBTC_ERR_FORMAT, "row offsets out of order in " + path);
  std::vector<uint32_t> band_of(N);
  for (uint32_t b = 0; b < P; ++b)
    for (uint32_t u = cuts[b]; u < cuts[b + 1]; ++u) band_of[u] = b;
  g = Csr{};
  g.n = N;
  g.ncols = N;
  g.m = E;
  g.off.resize(size_t(N) + 1);
  g.col.resize(std::max<size_t>(E, 1));
  uint64_t o = 0;
  for (uint32_t u = 0; u < N; ++u) {
    g.off[u] = o;
    const uint32_t y = band_of[u], r = u - cuts[y];
    for (uint32_t x = 0; x < P; ++x) {
      const uint64_t s = tiles[size_t(y) * P + x] + r;
      for (uint64_t q = rows[s]; q < slot_end(s); ++q) {
        const uint64_t c = uint64_t(cols[q]) + cuts[x];
        if (c >= N || o >= E) fail(BBTC_ERR_FORMAT, "column out of range in " + path);
        g.col[o++] = uint32_t(c);
      }
    }
  }
  if (o != E) fail(BBTC_ERR_FORMAT, "entry count does not match E in " + path);
  g.off[N] = o;
  validate_csr(g);
  cuts_out = cuts;
  cuts_out[P] = N + 1;  // the driver's convention (utils.hpp:288)
}

}  // namespace bbtc

extern "C" {

bbtc_status bbtc_layout_save_bcsr(const bbtc_layout* L, const char* path) {
  if (!L || !path) {
    bbtc::set_last_error("null argument");
    return BBTC_ERR_INVALID;
  }
  try {
    bbtc::save_bcsr(L->L, path);
  } catch (const bbtc::Error& e) {
    bbtc::set_last_error(e.what());
    return e.code;
  } catch (const std::exception& e) {
    bbtc::set_last_error(e.what());
    return BBTC_ERR_INTERNAL;
  }
  return BBTC_OK;
}

bbtc_status bbtc_layout_load_bcsr(const char* path, bbtc_hgraph** graph_out, bbtc_layout** layout_out) {
  if (!path || !graph_out || !layout_out) {
    bbtc::set_last_error("null argument");
    return BBTC_ERR_INVALID;
  }
  *graph_out = nullptr;
  *layout_out = nullptr;
  try {
    auto* h = new bbtc_hgraph{};
    std::vector<uint32_t> cuts;
    try {
      bbtc::load_bcsr(path, h->g, cuts);
      *layout_out = new bbtc_layout{bbtc::build_layout(h->g, cuts)};
    } catch (...) {
      delete h;
      throw;
    }
    *graph_out = h;
  } catch (const bbtc::Error& e) {
    bbtc::set_last_error(e.what());
    return e.code;
  } catch (const std::bad_alloc&) {
    bbtc::set_last_error("host allocation failed");
    return BBTC_ERR_NOMEM;
  } catch (const std::exception& e) {
    bbtc::set_last_error(e.what());
    return BBTC_ERR_INTERNAL;
  }
  return BBTC_OK;
}

}  // extern "C"
