// C-ABI of the host library: handles, status codes, thread-local messages.
#include <new>
#include <string>

#include "handles.hpp"

namespace {
thread_local std::string g_last_error;
}

namespace bbtc {
void set_last_error(const std::string& s) { g_last_error = s; }
}  // namespace bbtc

#define BBTC_TRY try {
#define BBTC_CATCH                                                              \
  }                                                                             \
  catch (const bbtc::Error& e) {                                                \
    g_last_error = e.what();                                                    \
    return e.code;                                                              \
  }                                                                             \
  catch (const std::bad_alloc&) {                                               \
    g_last_error = "host allocation failed";                                    \
    return BBTC_ERR_NOMEM;                                                      \
  }                                                                             \
  catch (const std::exception& e) {                                             \
    g_last_error = e.what();                                                    \
    return BBTC_ERR_INTERNAL;                                                   \
  }                                                                             \
  return BBTC_OK;

#define REQUIRE(cond, msg)                                                      \
  do {                                                                          \
    if (!(cond)) {                                                              \
      g_last_error = msg;                                                       \
      return BBTC_ERR_INVALID;                                                  \
    }                                                                           \
  } while (0)

extern "C" {

const char* bbtc_strerror(bbtc_status s) {
  switch (s) {
    case BBTC_OK: return "ok";
    case BBTC_ERR_INVALID: return "invalid argument";
    case BBTC_ERR_IO: return "I/O error";
    case BBTC_ERR_FORMAT: return "malformed graph";
    case BBTC_ERR_NOMEM: return "out of host memory";
    case BBTC_ERR_HIP: return "HIP runtime error";
    case BBTC_ERR_NODEVICE: return "no usable GPU";
    case BBTC_ERR_BUDGET: return "device memory budget exceeded";
    case BBTC_ERR_INTERNAL: return "internal error";
  }
  return "unknown status";
}

const char* bbtc_last_error(void) { return g_last_error.c_str(); }

bbtc_status bbtc_hgraph_load(const char* path, bbtc_hgraph** out) {
  REQUIRE(path && out, "null argument");
  BBTC_TRY
  auto* h = new bbtc_hgraph{bbtc::load_bbtc(path)};
  *out = h;
  BBTC_CATCH
}

bbtc_status bbtc_convert_text(const char* in_path, const char* out_path, int dedup) {
  REQUIRE(in_path && out_path, "null argument");
  BBTC_TRY
  bbtc::save_bbtc(bbtc::csr_from_text(in_path, dedup != 0), out_path);
  BBTC_CATCH
}

bbtc_status bbtc_hgraph_load_text(const char* path, bbtc_hgraph** out) {
  REQUIRE(path && out, "null argument");
  BBTC_TRY
  bbtc::Csr g = bbtc::csr_from_text(path, true);
  bbtc::validate_csr(g);
  *out = new bbtc_hgraph{std::move(g)};
  BBTC_CATCH
}

bbtc_status bbtc_hgraph_save(const bbtc_hgraph* g, const char* path) {
  REQUIRE(g && path, "null argument");
  BBTC_TRY
  bbtc::save_bbtc(g->g, path);
  BBTC_CATCH
}

bbtc_status bbtc_hgraph_from_csr(uint32_t n, uint32_t ncols, uint64_t m, const uint64_t* offsets,
                                 const uint32_t* cols, bbtc_hgraph** out) {
  REQUIRE(out && offsets, "null argument");
  BBTC_TRY
  *out = new bbtc_hgraph{bbtc::copy_csr(n, ncols, m, offsets, cols)};
  BBTC_CATCH
}

bbtc_status bbtc_hgraph_from_edges(uint32_t n, uint64_t nedges, const uint32_t* src,
                                   const uint32_t* dst, bbtc_hgraph** out) {
  REQUIRE(out, "null argument");
  BBTC_TRY
  *out = new bbtc_hgraph{bbtc::canonical_from_edges(n, nedges, src, dst)};
  BBTC_CATCH
}

bbtc_status bbtc_hgraph_gen_rmat(uint32_t scale, uint32_t edge_factor, double a, double b, double c,
                                 uint64_t seed, bbtc_hgraph** out) {
  REQUIRE(out, "null argument");
  BBTC_TRY
  *out = new bbtc_hgraph{bbtc::gen_rmat(scale, edge_factor, a, b, c, seed)};
  BBTC_CATCH
}

bbtc_status bbtc_hgraph_gen_gnm(uint32_t n, uint64_t m, uint64_t seed, bbtc_hgraph** out) {
  REQUIRE(out, "null argument");
  BBTC_TRY
  *out = new bbtc_hgraph{bbtc::gen_gnm(n, m, seed)};
  BBTC_CATCH
}

void bbtc_hgraph_info(const bbtc_hgraph* g, uint32_t* n, uint32_t* ncols, uint64_t* m) {
  if (!g) return;
  if (n) *n = g->g.n;
  if (ncols) *ncols = g->g.ncols;
  if (m) *m = g->g.m;
}
const uint64_t* bbtc_hgraph_offsets(const bbtc_hgraph* g) { return g ? g->g.off.data() : nullptr; }
const uint32_t* bbtc_hgraph_cols(const bbtc_hgraph* g) { return g ? g->g.col.data() : nullptr; }
void bbtc_hgraph_free(bbtc_hgraph* g) { delete g; }

bbtc_status bbtc_partition(const bbtc_hgraph* g, bbtc_part kind, uint32_t p, uint32_t* cuts_out) {
  REQUIRE(g && cuts_out, "null argument");
  REQUIRE(p >= 1, "p must be >= 1");
  BBTC_TRY
  auto c = bbtc::partition(g->g, kind, p);
  for (uint32_t b = 0; b <= p; ++b) cuts_out[b] = c[b];
  BBTC_CATCH
}

bbtc_status bbtc_layout_build(const bbtc_hgraph* g, const uint32_t* cuts, uint32_t p,
                              bbtc_layout** out) {
  REQUIRE(g && cuts && out, "null argument");
  REQUIRE(p >= 1, "p must be >= 1");
  BBTC_TRY
  std::vector<uint32_t> c(cuts, cuts + p + 1);
  *out = new bbtc_layout{bbtc::build_layout(g->g, c)};
  BBTC_CATCH
}

uint32_t bbtc_layout_ncut(const bbtc_layout* L) { return L ? L->L.P : 0; }
const uint32_t* bbtc_layout_cuts(const bbtc_layout* L) { return L ? L->L.cuts.data() : nullptr; }

bbtc_status bbtc_layout_tile(const bbtc_layout* L, uint32_t i, uint32_t k, bbtc_tile_info* out) {
  REQUIRE(L && out, "null argument");
  REQUIRE(i < L->L.P && k < L->L.P, "tile index out of range");
  const size_t t = L->L.tile(i, k);
  out->nnz = L->L.nnz[t];
  out->height = L->L.height(i);
  out->width = L->L.height(k);
  out->nz_rows = L->L.nz_rows[t];
  out->avg_degree = L->L.deg[t];
  out->max_row = L->L.max_row[t];
  return BBTC_OK;
}

uint64_t bbtc_layout_device_bytes(const bbtc_layout* L) { return L ? L->L.device_bytes() : 0; }

bbtc_status bbtc_layout_arrays(const bbtc_layout* L, const uint32_t** rowptr, uint64_t* rowptr_n,
                               const uint32_t** cols, uint64_t* cols_n, const uint32_t** nzrows,
                               uint64_t* nzrows_n) {
  REQUIRE(L && rowptr && rowptr_n && cols && cols_n && nzrows && nzrows_n, "null argument");
  const bbtc::Layout& x = L->L;
  const bool h = x.host_arrays;
  *rowptr = h ? x.rowptr.data() : nullptr;
  *rowptr_n = h ? x.rowptr.size() : 0;
  *cols = h ? x.cols.data() : nullptr;
  *cols_n = h ? x.cols.size() : 0;
  *nzrows = h ? x.nzrows.data() : nullptr;
  *nzrows_n = h ? x.nzrows.size() : 0;
  return BBTC_OK;
}
void bbtc_layout_free(bbtc_layout* L) { delete L; }

uint64_t bbtc_task_count(uint32_t p) { return uint64_t(p) * (p + 1) * (p + 2) / 6; }

bbtc_status bbtc_tasks_build(const bbtc_layout* L, bbtc_order order, bbtc_task* out, uint64_t cap,
                             uint64_t* n_out) {
  REQUIRE(L && n_out, "null argument");
  REQUIRE(order == BBTC_ORDER_LEX || order == BBTC_ORDER_DESC || order == BBTC_ORDER_ASC,
          "unknown task order");
  BBTC_TRY
  auto t = bbtc::task_list(L->L.P);
  bbtc::sort_tasks(t, L->L, order);
  *n_out = t.size();
  if (out) {
    if (cap < t.size()) bbtc::fail(BBTC_ERR_INVALID, "task buffer too small");
    std::copy(t.begin(), t.end(), out);
  }
  BBTC_CATCH
}

bbtc_status bbtc_tasks_bytes(const bbtc_layout* L, const bbtc_task* tasks, uint64_t n,
                             uint32_t offset_bytes, uint64_t* per_task_bytes, uint64_t* total_bytes,
                             uint64_t* total_wedges) {
  REQUIRE(L && (tasks || n == 0), "null argument");
  for (uint64_t q = 0; q < n; ++q)
    REQUIRE(tasks[q].i <= tasks[q].j && tasks[q].j <= tasks[q].k && tasks[q].k < L->L.P,
            "task out of range or not i <= j <= k");
  BBTC_TRY
  bbtc::task_bytes(L->L, tasks, n, offset_bytes, per_task_bytes, total_bytes, total_wedges);
  BBTC_CATCH
}

bbtc_status bbtc_tasks_stats(const bbtc_layout* L, const bbtc_task* tasks, uint64_t n, uint32_t sample_units,
                             uint64_t* stats) {
  REQUIRE(L && (tasks || n == 0) && (stats || n == 0), "null argument");
  for (uint64_t q = 0; q < n; ++q)
    REQUIRE(tasks[q].i <= tasks[q].j && tasks[q].j <= tasks[q].k && tasks[q].k < L->L.P,
            "task out of range or not i <= j <= k");
  BBTC_TRY
  bbtc::task_stats(L->L, tasks, n, stats, sample_units);
  BBTC_CATCH
}

bbtc_status bbtc_count_cpu(const bbtc_layout* L, const bbtc_task* tasks, uint64_t n, int nthreads,
                           uint64_t* per_task, uint64_t* total) {
  REQUIRE(L && (tasks || n == 0), "null argument");
  for (uint64_t q = 0; q < n; ++q)
    REQUIRE(tasks[q].i <= tasks[q].j && tasks[q].j <= tasks[q].k && tasks[q].k < L->L.P,
            "task out of range or not i <= j <= k");
  BBTC_TRY
  const uint64_t t = bbtc::count_cpu(L->L, tasks, n, nthreads, per_task);
  if (total) *total = t;
  BBTC_CATCH
}

}  // extern "C"
