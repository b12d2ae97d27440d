// Internal C++ interfaces of the host library (graph I/O, generators,
// partitioners, block layout, tasks, CPU count).
#pragma once

#include <vector>

#include "common.hpp"

namespace bbtc {

// ---- CSR ------------------------------------------------------------------
// Throws Error(BBTC_ERR_FORMAT) unless offsets start at 0, never decrease,
// end at m, and every row is strictly increasing with all columns > row and
// < max(n, ncols).
void validate_csr(const Csr& g);
Csr load_bbtc(const std::string& path);
void save_bbtc(const Csr& g, const std::string& path);
Csr copy_csr(uint32_t n, uint32_t ncols, uint64_t m, const uint64_t* off, const uint32_t* col);
// Edge list or Matrix Market text to the canonical upper-triangular CSR, as
// the reference's driver converts it; dedup drops repeated entries (which
// the reference keeps).
Csr csr_from_text(const std::string& path, bool dedup);

// ---- generators -----------------------------------------------------------
// Canonical input from an undirected pair list: symmetrise, drop self loops
// and duplicates, rank by (degree, id) ascending, keep the upper triangle.
Csr canonical_from_edges(uint32_t n, uint64_t nedges, const uint32_t* src, const uint32_t* dst);
Csr gen_rmat(uint32_t scale, uint32_t edge_factor, double a, double b, double c, uint64_t seed);
Csr gen_gnm(uint32_t n, uint64_t m, uint64_t seed);

// ---- partitioners ---------------------------------------------------------
std::vector<uint32_t> partition_uni(const Csr& g, uint32_t p);
std::vector<uint32_t> partition_n1d(const Csr& g, uint32_t p);
std::vector<uint32_t> partition_pal(const Csr& g, uint32_t p);
std::vector<uint32_t> partition_rac(const Csr& g, uint32_t p);
// Dispatch + cuts[p] = n + 1, as the reference driver does.
std::vector<uint32_t> partition(const Csr& g, bbtc_part kind, uint32_t p);

// ---- block layout -----------------------------------------------------------
// Tile (i,k), i <= k, of a P x P grid over the degree-ordered upper triangle.
// Stored per upper tile t (dense P*P indexing, lower tiles unused):
//   rowptr[rp_base[t] + r]   tile-local offset of local row r, r in [0, h_i]
//   cols[col_base[t] + e]    tile-local column id (global col - cuts[k])
//   nzrows[nz_base[t] + q]   q-th non-empty local row
// col_base is a multiple of 32 entries (128 bytes) so every tile segment
// starts on a cache line.
//
// Tiles whose column band is narrow also keep a bitmap form: row r of tile
// t is dense[dense_base[t] + r * dwpr[k] .. + dwpr[k]), one bit per column of
// band k, dwpr[k] = ceil(width/32) rounded up to 4 words (16-byte rows). The
// host layout holds only the placement (dense_words in all); the rows are
// built on the device from the uploaded tiles.
constexpr uint32_t kDenseMaxBits = 16384;            // widest band kept as bitmaps
constexpr uint64_t kDenseTileWords = 64ull << 20;    // largest bitmap tile (256 MiB)
constexpr uint64_t kDenseTotalWords = 4096ull << 20; // all bitmap tiles (16 GiB)
// a tile is kept as bitmap rows only if they take at most this many words per
// non-zero (a density of at least 1/512): sparser tiles gain nothing from the
// bitmap forms on RMAT scale 26 and would only hold memory
constexpr uint64_t kDenseWordsPerNnz = 16;

struct Layout {
  uint32_t P = 0;
  uint32_t n = 0;
  uint64_t m = 0;
  std::vector<uint32_t> cuts;      // P+1, cuts[P] == n + 1
  std::vector<uint64_t> rp_base;   // P*P
  std::vector<uint64_t> col_base;  // P*P
  std::vector<uint64_t> nz_base;   // P*P
  std::vector<uint64_t> nnz;       // P*P
  std::vector<uint32_t> nz_rows;   // P*P
  std::vector<uint32_t> max_row;   // P*P, entries in the tile's longest row
  std::vector<float> deg;          // P*P, (float)nnz / (float)h
  std::vector<uint64_t> dense_base; // P*P, ~0 when the tile has no bitmap form
  std::vector<uint32_t> dwpr;       // P, bitmap words per row of band k
  uint64_t dense_words = 4;         // words of all bitmap rows (device only)
  uint64_t rp_words = 0;            // entries of rowptr, cols and nzrows
  uint64_t col_words = 1;
  uint64_t nz_words = 1;
  // false for a layout built on the device without a host copy of its arrays:
  // only the per-tile fields above are valid then
  bool host_arrays = true;
  // entries in the graph's longest row (over all bands), when known (the
  // layout was built from a CSR)
  uint32_t longest_row = 0;
  bool longest_known = false;
  Buf<uint32_t> rowptr;
  Buf<uint32_t> cols;
  Buf<uint32_t> nzrows;

  uint32_t height(uint32_t b) const { return cuts[b + 1] - cuts[b]; }
  size_t tile(uint32_t i, uint32_t k) const { return size_t(i) * P + k; }
  uint64_t device_bytes() const {
    return rp_words * 4 + col_words * 4 + nz_words * 4 + dense_words * 4 +
           cuts.size() * 4 + 4 * rp_base.size() * 8 + nnz.size() * 8;
  }
};
Layout build_layout(const Csr& g, const std::vector<uint32_t>& cuts);
// The parts of build_layout shared with the device build: the checked cuts,
// per-tile vectors and row-offset bases of an n-row graph (no arrays)...
Layout layout_frame(uint32_t n, uint64_t m, const std::vector<uint32_t>& cuts);
// ...and, once nnz, nz_rows and max_row of every tile are known, the column
// and row-list bases, the tile degrees and the bitmap placement.
void layout_place(Layout& L);
// memcpy by all OpenMP threads in 1 MiB pieces taken dynamically (the
// destination's pages are touched in parallel, which a fresh host array
// needs to fill at memory speed).
void parallel_copy(void* dst, const void* src, size_t bytes);
// One past the largest column of g (0 when it has no entries); a layout
// needs it at most cuts[P].
uint64_t column_end(const Csr& g);
// Entries in g's longest row.
uint32_t longest_row(const Csr& g);
// Throws BBTC_ERR_INVALID naming `what` when L holds no host arrays.
void require_host_arrays(const Layout& L, const char* what);

// ---- tasks ----------------------------------------------------------------
std::vector<bbtc_task> task_list(uint32_t p);
void sort_tasks(std::vector<bbtc_task>& t, const Layout& L, bbtc_order order);
// Algorithmic bytes (see bbtc_tasks_bytes).
void task_bytes(const Layout& L, const bbtc_task* t, uint64_t n, uint32_t o, uint64_t* per_task,
                uint64_t* total, uint64_t* wedges);
// Work statistics per task (see bbtc_tasks_stats): out[6 q .. 6 q + 6).
// sample_units 0: exact; else at most that many units of kStatUnitRows
// consecutive non-empty ab rows per task, spread evenly over its rows, and
// the sums scaled to all its rows.
constexpr int kTaskStats = 6;
constexpr uint32_t kStatUnitRows = 64;
void task_stats(const Layout& L, const bbtc_task* t, uint64_t n, uint64_t* out, uint32_t sample_units = 0);
// True when the reference would skip the task: a tile with average degree 0.
bool task_skipped(const Layout& L, const bbtc_task& t);

// ---- CPU count ------------------------------------------------------------
uint64_t count_cpu(const Layout& L, const bbtc_task* t, uint64_t n, int nthreads,
                   uint64_t* per_task);
// A CPU worker's tables, kept from task to task. Dense tasks: stamp[c] == r
// marks column c of ac row r (the table is cleared for each task). Sparse
// tasks: mark[c] == epoch marks column c of the current long ac row; every
// such row takes a new epoch, so the table is cleared only when the epoch
// counter wraps or the table grows.
struct CpuWorker {
  std::vector<uint32_t> stamp;
  std::vector<uint32_t> mark;
  uint32_t epoch = 0;
};
// One task with the worker's tables.
uint64_t count_task_cpu(const Layout& L, const bbtc_task& t, CpuWorker& w);

}  // namespace bbtc
