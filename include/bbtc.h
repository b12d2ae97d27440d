/*
 * bbtc.h - C-ABI of the MI355X block-based triangle counter.
 *
 * Plain C: opaque handles, fixed-width integers, pointers and sizes. No HIP,
 * torch or C++ types cross this boundary (streams are passed as void*, which
 * is a hipStream_t; NULL means the library's own stream).
 *
 * Every entry point returns a bbtc_status; nothing calls exit(). On failure
 * bbtc_last_error() returns a message for the calling thread.
 *
 * Each declaration cites the reference interface it replaces (paths are
 * relative to the GT-TDAlab/bbTC source tree).
 */
#ifndef BBTC_H_
#define BBTC_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum bbtc_status {
  BBTC_OK = 0,
  BBTC_ERR_INVALID = 1,   /* bad argument or unsupported parameter */
  BBTC_ERR_IO = 2,        /* file could not be opened, read or written */
  BBTC_ERR_FORMAT = 3,    /* malformed .bbtc file or CSR content */
  BBTC_ERR_NOMEM = 4,     /* host allocation failed */
  BBTC_ERR_HIP = 5,       /* a HIP runtime call failed */
  BBTC_ERR_NODEVICE = 6,  /* no usable GPU */
  BBTC_ERR_BUDGET = 7,    /* device memory budget of the context exceeded */
  BBTC_ERR_INTERNAL = 8
} bbtc_status;

/* Replaces nothing: the reference reports no errors (utils.hpp:79-85 exits
 * with status 0; CUDA calls are unchecked). */
const char* bbtc_strerror(bbtc_status s);
const char* bbtc_last_error(void);

/* ------------------------------------------------------------------------ */
/* Host graph: degree-ordered upper-triangular CSR (.bbtc content).          */
/* ------------------------------------------------------------------------ */
typedef struct bbtc_hgraph bbtc_hgraph;

/* Replaces sarma::Matrix(path) -> pigo::CSR::read_bin_
 * (SARMA csr_matrix.hpp:132-152; PIGO csr.impl.hpp:475-512). Accepts the
 * PIGO-CSR-v2 binary with L=4 and O in {4,8}; rejects rows that are not
 * strictly increasing or not above the diagonal. */
bbtc_status bbtc_hgraph_load(const char* path, bbtc_hgraph** out);

/* Replaces pigo::CSR::save (PIGO csr.impl.hpp:428-473) as used by
 * utils.hpp:260-268. Writes L=4 and O=4 when m < 2^32, else O=8. */
bbtc_status bbtc_hgraph_save(const bbtc_hgraph* g, const char* path);

/* Replaces the text branch of utils::ReadGraph (utils.hpp:171-268) with
 * PIGO's edge-list / Matrix Market reader (coo.impl.hpp:318-466,
 * pigo.impl.hpp:184-345) and SARMA's degree order (csr_matrix.hpp:503-584):
 * symmetrise (diagonal dropped), rank by degree ascending with the same
 * std::sort over (id, degree) pairs, keep the upper triangle. Writes the
 * .bbtc file the reference writes, byte for byte, when dedup is 0; dedup 1
 * also drops repeated entries (the reference keeps them, and the counters
 * here accept only simple graphs). METIS .graph input is refused. */
bbtc_status bbtc_convert_text(const char* in_path, const char* out_path, int dedup);
/* The same conversion (with dedup) straight into a host graph. */
bbtc_status bbtc_hgraph_load_text(const char* path, bbtc_hgraph** out);

/* Copies a caller CSR (validated like bbtc_hgraph_load). */
bbtc_status bbtc_hgraph_from_csr(uint32_t n, uint32_t ncols, uint64_t m,
                                 const uint64_t* offsets, const uint32_t* cols,
                                 bbtc_hgraph** out);

/* Builds the canonical input from an undirected edge list: symmetrise, drop
 * self loops and duplicates, rank vertices by (degree, id) ascending, keep
 * the upper triangle, sort rows. The reference's text ingest
 * (utils.hpp:171-268) does the same except deduplication. */
bbtc_status bbtc_hgraph_from_edges(uint32_t n, uint64_t nedges, const uint32_t* src,
                                   const uint32_t* dst, bbtc_hgraph** out);

/* Seeded synthetic inputs (counter-based SplitMix64 streams; identical for any
 * thread count). RMAT: edge_factor * 2^scale samples with quadrant
 * probabilities a, b, c (d = 1-a-b-c). G(n,m): the first m distinct
 * non-loop pairs of a uniform pair stream. */
bbtc_status bbtc_hgraph_gen_rmat(uint32_t scale, uint32_t edge_factor, double a, double b,
                                 double c, uint64_t seed, bbtc_hgraph** out);
bbtc_status bbtc_hgraph_gen_gnm(uint32_t n, uint64_t m, uint64_t seed, bbtc_hgraph** out);

void bbtc_hgraph_info(const bbtc_hgraph* g, uint32_t* n, uint32_t* ncols, uint64_t* m);
const uint64_t* bbtc_hgraph_offsets(const bbtc_hgraph* g);
const uint32_t* bbtc_hgraph_cols(const bbtc_hgraph* g);
void bbtc_hgraph_free(bbtc_hgraph* g);

/* ------------------------------------------------------------------------ */
/* Partitioners.                                                             */
/* ------------------------------------------------------------------------ */
typedef enum bbtc_part {
  BBTC_PART_UNI = 0,
  BBTC_PART_N1D = 1,
  BBTC_PART_PAL = 2,
  BBTC_PART_RAC = 3
} bbtc_part;

/* Replaces the partition dispatch of utils.hpp:271-288:
 *   uni -> sarma::uniform::partition (uniform.hpp:20-27, 44-49)
 *   n1d -> sarma::nicol1d::partition(indptr) (nicol1d.hpp:106-216)
 *   pal -> sparsify (csr_matrix.hpp:435-454, tools/utils.hpp:20-25, 81-106)
 *          then sarma::probe_a_load::partition (probe_a_load.hpp:29-89)
 *   rac -> sarma::refine_a_cut::partition(A, p) on the whole matrix
 *          (refine_a_cut.hpp:43-73, 129-223; max(10, p) refinements)
 * Writes p+1 cuts into cuts_out exactly as the driver uses them, i.e. with
 * cuts_out[p] overwritten by n+1 (utils.hpp:288). Requires p >= 1. */
bbtc_status bbtc_partition(const bbtc_hgraph* g, bbtc_part kind, uint32_t p, uint32_t* cuts_out);

/* ------------------------------------------------------------------------ */
/* Block CSR layout (host copy; uploaded with bbtc_graph_upload).            */
/* ------------------------------------------------------------------------ */
typedef struct bbtc_layout bbtc_layout;

typedef struct bbtc_tile_info {
  uint64_t nnz;         /* entries in the tile */
  uint32_t height;      /* rows of band i (cuts[i+1]-cuts[i]) */
  uint32_t width;       /* columns of band k */
  uint32_t nz_rows;     /* rows with at least one entry */
  float avg_degree;     /* (float)nnz / height, as block_csr.cpp:278 (NaN when height is 0) */
  uint32_t max_row;     /* entries in the tile's longest row (the lane form needs <= 32) */
} bbtc_tile_info;

/* Replaces BlockCSR(cuts, x_adj, adj) -> CSRtoBCSR + ComputeTileProps
 * (block_csr.cpp:34-42, 257-376). Only tiles (i,k) with i <= k are stored;
 * each keeps tile-local 32-bit row offsets, tile-local column ids in a
 * 128-byte aligned segment, and a compact list of its non-empty rows. */
bbtc_status bbtc_layout_build(const bbtc_hgraph* g, const uint32_t* cuts, uint32_t p,
                              bbtc_layout** out);
uint32_t bbtc_layout_ncut(const bbtc_layout* L);
const uint32_t* bbtc_layout_cuts(const bbtc_layout* L);
/* Replaces the BlockCSR getters GetNNZ/GetHeight/GetWidth/GetTileDeg
 * (block_csr.cpp:87-186). Lower tiles (i > k) report zero entries. */
bbtc_status bbtc_layout_tile(const bbtc_layout* L, uint32_t i, uint32_t k, bbtc_tile_info* out);
/* Block CSR files (.bcsr): replaces BlockCSR::DeSerialize
 * (block_csr.cpp:215-255) and supplies the writer the reference lacks. Same
 * fields and widths (uint32 N, E, P, cuts[P+1], sparse_tiles[P*P+1],
 * rows[N*P], cols[E], tiles row-major over the full grid). A graph of n
 * vertices is written with cuts[P] = n, so n*P row slots are the tiles'
 * rows, and with N = n + 1 in the header: the slot after them holds the end
 * of the last row (E), which the reference's counters read, and the P - 1
 * words after it repeat E. Files whose header holds n itself (the end slot
 * implied) are read too. Loading rebuilds the graph and the layout with the
 * stored cuts (cuts[P] set to n+1 as the driver does), skipping
 * partitioning and the CSR-to-tile pass. */
bbtc_status bbtc_layout_save_bcsr(const bbtc_layout* L, const char* path);
bbtc_status bbtc_layout_load_bcsr(const char* path, bbtc_hgraph** graph_out, bbtc_layout** layout_out);

/* Bytes the layout occupies once uploaded. */
uint64_t bbtc_layout_device_bytes(const bbtc_layout* L);
/* The layout's three arrays as the device holds them (read-only, owned by
 * L): per upper tile its h + 1 tile-local row offsets, its tile-local
 * columns (each tile from a 32-entry boundary, padding zero) and its
 * non-empty local rows, tiles in (i, k) order, i <= k. Sets every pointer to
 * NULL and every size to 0 for a layout built on the device without its host
 * arrays. No counterpart in the reference (its BlockCSR keeps (N+P)*P
 * global row offsets and no row lists). */
bbtc_status bbtc_layout_arrays(const bbtc_layout* L, const uint32_t** rowptr, uint64_t* rowptr_n,
                               const uint32_t** cols, uint64_t* cols_n, const uint32_t** nzrows,
                               uint64_t* nzrows_n);
void bbtc_layout_free(bbtc_layout* L);

/* ------------------------------------------------------------------------ */
/* Tasks.                                                                    */
/* ------------------------------------------------------------------------ */
/* A task (i,j,k), i <= j <= k, counts triangles u<v<w with u in band i,
 * v in band j, w in band k, over tiles ab=(i,j), bc=(j,k), ac=(i,k). */
typedef struct bbtc_task { uint32_t i, j, k; } bbtc_task;

typedef enum bbtc_order {
  BBTC_ORDER_LEX = 0,   /* i, j, k lexicographic (tasks.hpp:7-17) */
  BBTC_ORDER_DESC = 1,  /* by weight, descending: the GPU build (tasks.hpp:25-28) */
  BBTC_ORDER_ASC = 2    /* by weight, ascending: the CPU build (tasks.hpp:29-32) */
} bbtc_order;

/* C(p+2,3). */
uint64_t bbtc_task_count(uint32_t p);

/* Replaces tasks::TaskList + tasks::SortTasks (tasks.hpp:7-34): weight is
 * nnz(ab) * max(deg(bc), deg(ac)) in float, sorted with std::sort. */
bbtc_status bbtc_tasks_build(const bbtc_layout* L, bbtc_order order, bbtc_task* out,
                             uint64_t cap, uint64_t* n_out);

/* Exact algorithmic bytes per task for the roofline:
 *   B = sum over rows u of band i with nonempty ab(u) and ac(u) of
 *       (2o + 4 d_ab(u)) + (2o + 4 d_ac(u)) + sum_{v in ab(u)} (2o + 4 d_bc(v))
 * with o = offset_bytes. Also returns the wedge count sum d_ab(u). Tasks with
 * an empty tile contribute nothing (bb_tc.cpp:86-91). Either output array may
 * be NULL. */
bbtc_status bbtc_tasks_bytes(const bbtc_layout* L, const bbtc_task* tasks, uint64_t n,
                             uint32_t offset_bytes, uint64_t* per_task_bytes,
                             uint64_t* total_bytes, uint64_t* total_wedges);

/* Work of each task, as the device kernels meet it (no counterpart in the
 * reference; the measure behind the kernel-form cost rule). stats[6q..6q+6)
 * receives, over the rows u of band i with non-empty ab(u) and ac(u):
 *   [0] rows, [1] wedges sum d_ab(u), [2] sum d_ac(u),
 *   [3] probes of bc lists sum_{v in ab(u)} d_bc(v),
 *   [4] probes of ac lists sum d_ab(u) d_ac(u),
 *   [5] sum d_ab(u) min(d_ac(u), lines), lines = 128-byte lines of a bitmap
 *       row of band k (the lines an ac list's probes of a bc bitmap row meet).
 * Tasks the reference skips (bb_tc.cpp:86-91) get zeros. sample_units 0
 * gives exact sums; otherwise each task's non-empty ab rows are read in
 * units of 64 consecutive rows, at most sample_units units spread evenly
 * over the rows, and the sums are scaled to all rows (estimates). */
bbtc_status bbtc_tasks_stats(const bbtc_layout* L, const bbtc_task* tasks, uint64_t n, uint32_t sample_units,
                             uint64_t* stats);
/* ------------------------------------------------------------------------ */
/* CPU count: the --ngpu 0 path and the host tail of --border.               */
/* ------------------------------------------------------------------------ */
/* Replaces BBTc::NoOfTriangles (CPU build, bb_tc.cpp:255-273) with
 * RunCpuKernels / NoOfTrianglesInCpus / TrianglesIn{Sparse,Dense}Task
 * (bb_tc.cpp:118-139, 184-197; cpu_sparse_kernel.cpp; cpu_dense_kernel.cpp):
 * workers claim tasks from the back of the list, in units of 1,024
 * non-empty rows so that all of them share a heavy task (pass the tasks in
 * ascending weight, BBTC_ORDER_ASC, to start with the heaviest); the ac
 * rows of sparse tasks are marked once and probed rather than merged per
 * wedge, with the same counts. nthreads <= 0 means all OpenMP threads.
 * per_task (may be NULL) receives one count per task, the count of the whole
 * task. */
bbtc_status bbtc_count_cpu(const bbtc_layout* L, const bbtc_task* tasks, uint64_t n,
                           int nthreads, uint64_t* per_task, uint64_t* total);

/* ------------------------------------------------------------------------ */
/* Device side.                                                              */
/* ------------------------------------------------------------------------ */
typedef struct bbtc_ctx bbtc_ctx;
typedef struct bbtc_dgraph bbtc_dgraph;
typedef struct bbtc_plan bbtc_plan;

typedef enum bbtc_kernel {
  BBTC_KERNEL_AUTO = 0,    /* per task the form of least estimated cost from the tasks' work
                              statistics (a linear model of measured clocks); LANE for every task
                              when every task with work admits it, and then ROWS for a single
                              graph whose rows fit its row table (bbtc_autotune measures
                              instead) */
  BBTC_KERNEL_MERGE = 1,   /* bc(v) entries binary-searched in ac(u) (LDS copy) */
  BBTC_KERNEL_BITMAP = 2,  /* bc(v) entries tested against an LDS bitmap of ac(u), where the
                              band fits the slice */
  BBTC_KERNEL_DENSE = 3,   /* LDS bitmap of ac(u) ANDed with the bitmap row of bc(v) where the
                              bc tile has one, else as BITMAP */
  BBTC_KERNEL_RDENSE = 4,  /* ac(u) entries tested against the bitmap row of bc(v) where the
                              bc tile has one, else as MERGE */
  BBTC_KERNEL_LANE = 5,    /* one lane per row u merges bc(v) with ac(u) in HBM, the
                              per-thread row merge of SparseKernel (gpu_sparse_kernel.cu:6-49),
                              where no row of the three tiles holds more than 32 entries,
                              else as MERGE */
  BBTC_KERNEL_ROWS = 6     /* as AUTO, and a plan of one graph whose tasks all take the LANE
                              form counts every task in one launch over whole rows: a wedge
                              (u, v) merges N+(u) with N+(v) over all bands once (rows of at
                              most 64 entries, kept in a row table built on first use) and
                              each common w adds to task (band(u), band(v), band(w)) */
} bbtc_kernel;

typedef struct bbtc_opts {
  uint64_t hbm_budget_bytes;  /* 0: no limit beyond the device's own */
} bbtc_opts;

/* Number of visible HIP devices (0 when there is none); replaces the
 * cudaGetDeviceCount calls of utils.hpp:94 and bb_tc.cpp:19. */
bbtc_status bbtc_device_count(int* n);

/* Replaces BBTc::PrepareGpu (bb_tc_helper.cpp:53-73): one context per
 * device holds the stream, counters and scratch. */
bbtc_status bbtc_ctx_create(int device, const bbtc_opts* opts, bbtc_ctx** out);
void bbtc_ctx_destroy(bbtc_ctx* ctx);
/* Waits until all work queued on the context's device has finished. */
bbtc_status bbtc_ctx_synchronize(bbtc_ctx* ctx);
/* Device bytes currently held by graphs and plans of this context. */
uint64_t bbtc_ctx_bytes_in_use(const bbtc_ctx* ctx);
/* Achievable HBM read bandwidth of the context's device: a streaming read
 * kernel over `bytes` (at least 1 MiB, rounded down to 16-byte words) of
 * zeroed device memory, one untimed pass and then `iters` passes timed with
 * HIP events; gbps_out[i] = bytes read / time of pass i, in GB/s (1e9 B/s).
 * The buffer is allocated and freed inside the call. No counterpart in the
 * reference: it is the measured ceiling reported beside the spec peak. */
bbtc_status bbtc_ctx_read_bandwidth(bbtc_ctx* ctx, uint64_t bytes, int iters, double* gbps_out);

/* The seeded generators above run on the context's device (radix sorts in
 * HBM) and return the same host CSR, bit for bit, much faster for large
 * graphs. Sample or candidate counts must stay below 2^31. */
bbtc_status bbtc_hgraph_gen_rmat_device(bbtc_ctx* ctx, uint32_t scale, uint32_t edge_factor, double a,
                                        double b, double c, uint64_t seed, bbtc_hgraph** out);
bbtc_status bbtc_hgraph_gen_gnm_device(bbtc_ctx* ctx, uint32_t n, uint64_t m, uint64_t seed,
                                       bbtc_hgraph** out);

/* Replaces CopyTilesToGpu / CopyTileToGpu (bb_tc_helper.cpp:75-172, the
 * --copy-first path): the whole block CSR becomes device resident once. */
bbtc_status bbtc_graph_upload(bbtc_ctx* ctx, const bbtc_layout* L, bbtc_dgraph** out);
/* Splits tasks[0..n) of a resident graph into `shards` shares, one per GPU
 * of a multi-GPU run, and writes each task's share to shard_out[q]: longest
 * estimated time first onto the least loaded share (LPT). The estimate
 * (cost_out, may be NULL) is the
 * AUTO rule's shader clocks for the task's form, or the reference's task
 * weight (tasks.hpp:21-23) where the graph has no statistics (more than 128
 * cuts); tasks with an empty tile cost 0. Replaces the reference's dealing
 * of task q to GPU q mod G (bb_tc.cpp:142-181, thread_id % n_gpu with a
 * stride of all workers). One share's tasks counted alone on one GPU are
 * that GPU's part of the run at its real shape. */
bbtc_status bbtc_tasks_shard(const bbtc_dgraph* g, const bbtc_task* tasks, uint64_t n, uint32_t shards,
                             uint32_t* shard_out, double* cost_out);
/* Moves L's block-CSR arrays (row offsets, columns, row lists) into
 * page-locked host memory, as the reference allocates its BlockCSR's tiles
 * with cudaMallocHost (block_csr.cpp:296-299): uploads, batches and served
 * requests then copy them straight over PCIe (the arrays of an unpinned
 * layout pass through the context's two pinned staging buffers, filled by the
 * host threads). The contents do not change; a pinned layout must not be
 * freed while a copy from it may be in flight (a served request until it is
 * polled). Costs one host copy of the arrays; BBTC_ERR_HIP when the memory
 * cannot be locked (the layout is then unchanged or partly pinned, and valid
 * either way). bbtc_layout_pinned: 1 when every array is pinned. */
bbtc_status bbtc_layout_pin(bbtc_ctx* ctx, bbtc_layout* L);
int bbtc_layout_pinned(const bbtc_layout* L);
uint64_t bbtc_graph_bytes(const bbtc_dgraph* g);
void bbtc_graph_release(bbtc_dgraph* g);

/* Replaces BlockCSR(cuts, x_adj, adj) (block_csr.cpp:286-376) followed by
 * CopyTilesToGpu (bb_tc_helper.cpp:75-172): the CSR is copied to the device
 * once and the block layout is built there by data-parallel kernels (row
 * counts, per-tile scans, column scatter), then stays resident. The device
 * arrays are those bbtc_layout_build + bbtc_graph_upload make, bit for bit,
 * and so are the work statistics the AUTO rule reads (computed on the
 * device). *layout_out receives the host side: the cuts and every per-tile
 * field (bbtc_layout_tile, bbtc_tasks_build). With BBTC_BUILD_HOST_ARRAYS
 * the arrays are also copied back, so every entry point that takes a
 * bbtc_layout works on it; without, those that read the arrays
 * (bbtc_tasks_bytes, bbtc_tasks_stats, bbtc_count_cpu, the host tail of
 * bbtc_count_hybrid, bbtc_graph_upload, bbtc_layout_save_bcsr, batches and
 * the server) return BBTC_ERR_INVALID. Requires every column of g below
 * cuts[p]. Free both outputs (bbtc_layout_free, bbtc_graph_release). */
#define BBTC_BUILD_HOST_ARRAYS 1u
bbtc_status bbtc_graph_build(bbtc_ctx* ctx, const bbtc_hgraph* g, const uint32_t* cuts, uint32_t p,
                             uint32_t flags, bbtc_layout** layout_out, bbtc_dgraph** graph_out);
/* bbtc_tasks_stats computed on the device from a resident graph: the same
 * six numbers per task, bit for bit. No counterpart in the reference. */
bbtc_status bbtc_graph_tasks_stats(bbtc_dgraph* g, const bbtc_task* tasks, uint64_t n, uint32_t sample_units,
                                   uint64_t* stats);

/* A plan binds a graph to a task list and a kernel choice: it uploads task
 * descriptors and the work-item table once, so repeated launches do no host
 * work. Replaces the per-task launch loop of BBTc::RunGpuKernels /
 * NoOfTrianglesInGPU (bb_tc.cpp:76-115, 142-181). With BBTC_KERNEL_AUTO and
 * no choice array the plan runs the tasks in an order of its own (each
 * form's tasks grouped by bc tile, the forms interleaved by estimated
 * work); every per-task output (results, kinds, clocks, times) is in the
 * order of `tasks`. */
bbtc_status bbtc_plan_create(bbtc_ctx* ctx, bbtc_dgraph* g, const bbtc_task* tasks, uint64_t n,
                             bbtc_kernel kernel, bbtc_plan** out);
/* As bbtc_plan_create, with a per-task override choice[q]: in bits 0-2, 0
 * keeps `kernel`'s rule, 1 forces MERGE, 2 BITMAP, 3 DENSE, 4 RDENSE, 5 LANE
 * (each where it applies); bits 4-5 halve the rows of a work item that many
 * times (0: 64 rows, or as many as the form's LDS bitmaps fit).
 * Used with bbtc_autotune's output; choice may be NULL. */
bbtc_status bbtc_plan_create_tuned(bbtc_ctx* ctx, bbtc_dgraph* g, const bbtc_task* tasks,
                                   uint64_t n, bbtc_kernel kernel, const uint8_t* choice,
                                   bbtc_plan** out);
/* Zeroes the plan's per-task counters and enqueues the count on stream
 * (NULL: the context stream). Asynchronous. */
bbtc_status bbtc_plan_launch(bbtc_plan* plan, void* stream);
/* Waits for the plan's last launch and copies per-task counts (may be NULL)
 * and their sum. */
bbtc_status bbtc_plan_results(bbtc_plan* plan, uint64_t* per_task, uint64_t* total);
/* Launches the plan iters times on the context stream, recording a HIP event
 * pair around each launch; ms_each[iters] receives the event times. */
bbtc_status bbtc_plan_time(bbtc_plan* plan, int iters, float* ms_each);
/* Counts every task of the plan alone, one launch per task in list order with
 * a HIP event pair around each; ms_per_task[n] receives the times. This is
 * the measurement the autotuner builds its per-task cost table from. */
bbtc_status bbtc_plan_time_tasks(bbtc_plan* plan, float* ms_per_task);
/* Runs the plan once with per-item clock counters: cycles_per_task[n]
 * receives the shader clocks the device spent on each task's work items
 * while all tasks share the device (the autotuner's cost measure). */
bbtc_status bbtc_plan_profile(bbtc_plan* plan, uint64_t* cycles_per_task);
/* Autotuner: measures every form on the whole task list (each forced where it
 * applies, per-item clock counters) and writes per task the cheapest of the
 * forms 1 merge, 2 bitmap, 3 dense, 4 rdense into bits 0-2 of choice_out[n],
 * ready for bbtc_plan_create_tuned. The lane form (5) runs in a kernel of its
 * own only when every task of a launch uses it, so where it applies to every
 * task with work, that per-task choice is timed against all tasks in the lane
 * form and the faster one is written. Then the chosen forms are measured with
 * 64 (or as many as fit), 32 and 16 rows per work item, and each task's
 * cheapest goes into bits 4-5 (the rows' halvings). cycles_out (may be NULL)
 * receives n x 5 clock counts (forms 1-5 in order, at 64 rows), UINT64_MAX
 * where a form does not apply. Replaces the fixed threshold of bb_tc.cpp:98
 * and the fixed 256-thread row blocks of bb_tc.cpp:96-105. */
bbtc_status bbtc_autotune(bbtc_ctx* ctx, bbtc_dgraph* g, const bbtc_task* tasks, uint64_t n,
                          uint8_t* choice_out, uint64_t* cycles_out);
/* Form the plan chose per task: 1 merge, 2 bitmap, 3 dense, 4 rdense, 5 lane, 6 rows (the
 * plan's tasks with work are counted by the row form's launch). */
bbtc_status bbtc_plan_task_kinds(const bbtc_plan* plan, uint8_t* kinds);
/* Number of tasks the plan routed to an LDS bitmap form (bitmap or dense). */
uint64_t bbtc_plan_bitmap_tasks(const bbtc_plan* plan);
void bbtc_plan_destroy(bbtc_plan* plan);

/* One-shot convenience: plan, launch, wait, copy, destroy. Replaces
 * BBTc::NoOfTriangles (GPU build, bb_tc.cpp:200-252) for a fully resident
 * graph; per_task may be NULL. */
bbtc_status bbtc_count(bbtc_ctx* ctx, bbtc_dgraph* g, const bbtc_task* tasks, uint64_t n,
                       bbtc_kernel kernel, uint64_t* per_task, uint64_t* total, void* stream);
/* No counterpart in the reference. For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): bbtc_count with no device allocation and no free. The
 * task descriptors, item prefix, per-task counters, launch control words
 * and spill records live in a fixed 12 MiB array of each device's global
 * memory, held until the process ends, so at most 65,536 tasks per call
 * (a larger n returns BBTC_ERR_INVALID; 2^18 spill records, past which heavy
 * items are not split, which costs time, not counts). Per-task counts
 * (per_task, may be NULL) and their sum (total, may be NULL; not both) are
 * the same as bbtc_count's. The call waits for its work before returning.
 * Every call on one device shares that device's array: calls that overlap on
 * one device (other threads, other contexts of the same device) race and
 * must be serialised by the caller. */
bbtc_status bbtc_count_noalloc(bbtc_ctx* ctx, bbtc_dgraph* g, const bbtc_task* tasks, uint64_t n,
                               bbtc_kernel kernel, uint64_t* per_task, uint64_t* total, void* stream);

/* ------------------------------------------------------------------------ */
/* Batches: many independent graphs, one persistent launch.                  */
/* ------------------------------------------------------------------------ */
/* No counterpart in the reference (one graph per run). Block layouts of n
 * graphs built in parallel, graph x with ncut[x] cuts of partitioner kind. */
bbtc_status bbtc_layouts_build(const bbtc_hgraph* const* graphs, uint64_t n, const uint32_t* ncut,
                               bbtc_part kind, bbtc_layout** out);

/* Cut-count autotuner: counts every sample graph with every candidate cut
 * count and measures the device work, with the forms autotuned per graph
 * (flag BBTC_TUNE_FORMS in tune_forms) or chosen by BBTC_KERNEL_AUTO. The
 * work is the summed per-item clocks (the share of the machine a graph takes
 * inside a batch), or with flag BBTC_TUNE_BY_TIME the wall time of its
 * upload plus the event time of a launch of that graph alone, in ns (what a
 * graph counted by itself waits for).
 * rule[64] receives, per size class floor(log2(m+1)), the candidate with the
 * least work; classes without samples copy the nearest sampled class (the
 * smaller one first), and no class takes fewer cuts than a smaller class.
 * work_out (may be NULL) receives the n x ncand measurements. */
#define BBTC_TUNE_FORMS 1
#define BBTC_TUNE_BY_TIME 2
bbtc_status bbtc_tune_ncut(bbtc_ctx* ctx, const bbtc_hgraph* const* sample, uint64_t n,
                           const uint32_t* cands, uint32_t ncand, bbtc_part kind, int tune_forms,
                           uint32_t* rule, uint64_t* work_out);
/* The rule's cut count for a graph with m edges (0 when rule is NULL). */
uint32_t bbtc_rule_ncut(const uint32_t* rule, uint64_t m);

typedef struct bbtc_batch bbtc_batch;
/* Uploads n layouts into one device arena and plans every task of every
 * graph (heavy tasks of all graphs first) for one persistent launch. */
typedef enum bbtc_batch_order {
  BBTC_BATCH_HEAVY_FIRST = 0,        /* tasks of all graphs by weight, heaviest first */
  BBTC_BATCH_SMALL_GRAPHS_FIRST = 1  /* graphs by edge count, smallest first; heavy tasks first within */
} bbtc_batch_order;
bbtc_status bbtc_batch_create(bbtc_ctx* ctx, const bbtc_layout* const* layouts, uint64_t n,
                              bbtc_kernel kernel, bbtc_batch_order order, bbtc_batch** out);
uint64_t bbtc_batch_ntasks(const bbtc_batch* b);
uint64_t bbtc_batch_bytes(const bbtc_batch* b);
/* The batch's task order: graph index and (i,j,k) of each task (either may be NULL). */
bbtc_status bbtc_batch_tasks(const bbtc_batch* b, uint32_t* graph, bbtc_task* tasks);
/* Form the batch's plan chose per task (batch task order): 1 merge, 2 bitmap,
 * 3 dense, 4 rdense, 5 lane. */
bbtc_status bbtc_batch_task_kinds(const bbtc_batch* b, uint8_t* kinds);
/* Measures every form on the batch (as bbtc_autotune) and replans with the
 * cheapest per task; choice_out (may be NULL) receives the forms. */
bbtc_status bbtc_batch_autotune(bbtc_batch* b, uint8_t* choice_out);
/* Replans with a kernel rule and optional per-task forms (batch task order). */
bbtc_status bbtc_batch_set_forms(bbtc_batch* b, bbtc_kernel kernel, const uint8_t* choice);
/* Records the launch sequence (clock stamp, counter resets, persistent count)
 * as a HIP graph; later bbtc_batch_launch calls replay it. */
bbtc_status bbtc_batch_capture(bbtc_batch* b);
/* Enqueues one count of the whole batch on the context stream. */
bbtc_status bbtc_batch_launch(bbtc_batch* b);
/* Waits, then writes per-graph counts (n), per-task counts (batch order) and
 * each graph's completion time after the launch began, in ms, from the
 * device wall clock. Any output may be NULL. */
bbtc_status bbtc_batch_results(bbtc_batch* b, uint64_t* per_graph, uint64_t* per_task,
                               double* latency_ms);
void bbtc_batch_destroy(bbtc_batch* b);

/* Replaces BBTc::NoOfTriangles of the GPU build with its --border semantics
 * (bbtc.cu:54-62; bb_tc.cpp:142-252): tasks [0, separator) run on the device
 * only; tasks [separator, n) are shared, the device taking them from the
 * front while nthreads-1 host workers take single tasks from the back, each
 * only while it would finish before the device reaches it at the measured
 * rates. Where the device's atomics reach host memory, both sides claim from
 * one word in host-coherent memory: one persistent launch counts the whole
 * list and its waves move the front with system-scope atomics as they reach
 * the shared tasks (those of the launch's last part). Otherwise a host thread
 * claims device batches from the front on the device's behalf (up to
 * nstreams launches in flight). ctx and g NULL: every task on the host
 * (--ngpu 0). choice (may be NULL) gives per-task device forms as in
 * bbtc_plan_create_tuned. gpu_tasks_out (may be NULL) receives how many
 * leading tasks the device counted. */
bbtc_status bbtc_count_hybrid(bbtc_ctx* ctx, bbtc_dgraph* g, const bbtc_layout* L,
                              const bbtc_task* tasks, uint64_t n, uint64_t separator, int nthreads,
                              int nstreams, bbtc_kernel kernel, const uint8_t* choice,
                              uint64_t* per_task, uint64_t* total, uint64_t* gpu_tasks_out);

/* ------------------------------------------------------------------------- */
/* Serving: count requests for many graphs on one device with continuous      */
/* batching. Replaces running the reference's main (bbtc.cu:22-89) once per   */
/* graph: requests are admitted in arrival order while their device footprint */
/* (block CSR, bitmap rows of dense tiles, plan tables) fits the server heap; */
/* an upload thread copies them in while earlier requests count; each launch  */
/* fills a budget of work items: half kept by the oldest resident request,    */
/* the rest least remaining work first, so arrivals join the next launch and  */
/* share the machine with requests already running.                           */
/* ------------------------------------------------------------------------- */
typedef struct bbtc_server bbtc_server;

typedef struct bbtc_server_opts {
  uint64_t heap_bytes;     /* device heap for resident requests (0: 200 GB, 2e11 bytes) */
  uint64_t quantum_items;  /* work items per launch (0: adaptive, see launch_us) */
  uint32_t max_inflight;   /* launches queued on the device at once (0: 2) */
  uint32_t max_segments;   /* requests per launch (0: 1024) */
  uint32_t sequential;     /* 1: one request at a time, whole plan per launch */
  uint32_t lane_form;      /* 1: BBTC_KERNEL_AUTO requests whose tasks all admit the lane form
                              take it and run in launches of their own (the lane kernel),
                              taking turns with the others; 0: they keep the per-task forms,
                              so any requests share a launch */
  uint32_t launch_us;      /* with quantum_items 0: launches are composed by estimated work
                              (the reference's task weight spread over a task's items), and
                              the work budget follows the measured rate of finished launches
                              so that a launch takes about this long (0: 40000 us) */
  uint32_t oldest_pct;     /* percent of each launch's budget the oldest request with work
                              keeps (0: 90); the rest goes shortest remaining work first */
  uint32_t list_chunks;    /* 0: a launch's items are split into one contiguous range per
                              claim part (each XCD on its own range, as for one graph);
                              1: 4,096-item chunks dealt to the parts in turn */
  uint32_t defer_steps;    /* wave steps after which a work item is split (0: 1,024 steps for
                              launches of at least 64 items per resident wave without
                              list_chunks, else 512) */
} bbtc_server_opts;

typedef struct bbtc_server_result {
  uint64_t id;
  uint64_t total;      /* triangles of the request's tasks */
  int32_t status;      /* BBTC_OK, or why the request was dropped */
  uint32_t launches;   /* launches its items were spread over */
  uint64_t bytes;      /* device footprint while resident */
  /* server clock (seconds since create / reset_clock) */
  double arrival_s, admit_s, ready_s, start_s, done_s;
} bbtc_server_result;

typedef struct bbtc_server_stats {
  uint64_t launches, segments, items;
  uint64_t peak_resident_bytes, heap_bytes;
  uint32_t peak_active, peak_segments;
  double busy_s;    /* sum of the counting launches' device times (HIP events) */
  double upload_s;  /* host time the upload thread spent copying requests in */
  uint64_t row_form_requests;  /* requests counted in the row form (a launch of their own) */
} bbtc_server_stats;

bbtc_status bbtc_server_create(bbtc_ctx* ctx, const bbtc_server_opts* opts, bbtc_server** out);
/* Restarts the server clock and statistics; no request may be outstanding. */
bbtc_status bbtc_server_reset_clock(bbtc_server* s);
double bbtc_server_now(const bbtc_server* s);
/* Queues a count of tasks[0..n) of L, arriving at arrival_s on the server
 * clock (not decreasing between calls; the request waits until then). L and
 * per_task_out (may be NULL) must stay valid until the result is polled.
 * A request that bbtc_count would count in the row form (BBTC_KERNEL_ROWS,
 * or AUTO where it pays; L built from a CSR with rows of at most 64 entries)
 * is counted in that form, in a launch of its own that takes turns with the
 * shared launches. */
bbtc_status bbtc_server_submit(bbtc_server* s, const bbtc_layout* L, const bbtc_task* tasks, uint64_t n,
                               bbtc_kernel kernel, double arrival_s, uint64_t* per_task_out,
                               uint64_t* id_out);
/* Up to max finished requests in completion order; waits up to wait_ms
 * (-1: until one finishes or nothing is outstanding). */
bbtc_status bbtc_server_poll(bbtc_server* s, bbtc_server_result* out, uint64_t max, int wait_ms,
                             uint64_t* got);
/* Waits until every submitted request has finished. */
bbtc_status bbtc_server_wait_idle(bbtc_server* s);
bbtc_status bbtc_server_get_stats(bbtc_server* s, bbtc_server_stats* out);
/* Finishes admitted requests, drops those not yet admitted, frees the heap. */
void bbtc_server_destroy(bbtc_server* s);

#ifdef __cplusplus
}
#endif

#endif /* BBTC_H_ */
