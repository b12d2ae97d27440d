/*
 * bbtc_oracle.c - CPU restatement of the reference block-based triangle count.
 *
 * TEST INFRASTRUCTURE ONLY. Results are checked against this file: only
 * tests/, the smoke entry point and the bench's CPU-baseline leg may load it.
 * The product path (libbbtc.so) never links, calls or falls back to it.
 *
 * It restates, in plain C with OpenMP, the reference's CPU build (paths
 * relative to the GT-TDAlab/bbTC tree):
 *   - .bbtc reader            PIGO csr.impl.hpp:475-512
 *   - uni cuts                SARMA algorithms/uniform.hpp:20-27
 *   - n1d cuts                SARMA algorithms/nicol1d.hpp:30-216
 *   - rac cuts                SARMA algorithms/refine_a_cut.hpp:43-73, 129-223
 *   - pal cuts                SARMA tools/utils.hpp:20-25, 81-106 (sample),
 *                             csr_matrix.hpp:435-454 (sparsify),
 *                             sparse_prefix_sum.hpp:318-387 (loads),
 *                             algorithms/probe_a_load.hpp:29-89 (search)
 *   - driver cuts[p] = N+1    include/utils/utils.hpp:288
 *   - BlockCSR layout         src/data_structures/block_csr.cpp:257-376
 *   - task list and weights   include/utils/tasks.hpp:7-34
 *   - per-task count          src/triangle_counting/bb_tc.cpp:118-139, 184-197,
 *                             255-273; cpu_sparse_kernel.cpp:3-59;
 *                             cpu_dense_kernel.cpp:3-59
 *
 * Pinning (each function is checked against something it does not restate):
 *   - counts: brute force over the adjacency matrix (orc_brute_force) and the
 *     per-task definition counter (orc_count_definition), closed forms in the
 *     tests, and totals invariant across cuts;
 *   - cuts: golden vectors written by the reference's own code built from its
 *     sources (oracle/ref_harness.cpp) in tests/golden;
 *   - layout: reconstruction of the input CSR (orc_bcsr_row) in the tests;
 *   - algorithmic bytes: an independent recomputation from the plain CSR in
 *     the tests.
 *
 * Deviations, all without effect on counts:
 *   - tasks with equal weight keep lexicographic order (a stable merge sort);
 *     std::sort's order for ties is unspecified;
 *   - the per-worker stamp table of a dense task is as wide as the ac band
 *     (the reference sizes it by the widest dense tile, cpu_dense_kernel.cpp
 *     53-59, and would write past it for a wider ac band);
 *   - rows_ has one more word than (N+P)*P and the prefix sum covers it: with
 *     P = 1 the reference reads that word past its array (block_csr.cpp:292,
 *     348-351); its CLI refuses --ncut 1 (utils.hpp:122-125).
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define ORC_API __attribute__((visibility("default")))

/* ------------------------------------------------------------------------ */
/* .bbtc reader                                                              */
/* ------------------------------------------------------------------------ */
typedef struct {
  uint32_t n, ncols;
  uint64_t m;
  uint64_t* off;
  uint32_t* col;
} orc_csr;

ORC_API int orc_load_bbtc(const char* path, orc_csr* g) {
  memset(g, 0, sizeof *g);
  FILE* f = fopen(path, "rb");
  if (!f) return -1;
  char hdr[11];
  uint8_t lo[2];
  uint32_t n, nrows, ncols;
  uint64_t m = 0;
  int rc = -2;
  if (fread(hdr, 1, 11, f) != 11 || memcmp(hdr, "PIGO-CSR-v2", 11) != 0) goto out;
  if (fread(lo, 1, 2, f) != 2 || lo[0] != 4 || (lo[1] != 4 && lo[1] != 8)) goto out;
  if (fread(&n, 4, 1, f) != 1) goto out;
  if (lo[1] == 4) {
    uint32_t m32;
    if (fread(&m32, 4, 1, f) != 1) goto out;
    m = m32;
  } else if (fread(&m, 8, 1, f) != 1) goto out;
  if (fread(&nrows, 4, 1, f) != 1 || fread(&ncols, 4, 1, f) != 1) goto out;
  {
    const uint32_t rows = nrows > n ? nrows : n;
    g->n = rows;
    g->ncols = ncols;
    g->m = m;
    g->off = (uint64_t*)malloc(8ull * (rows + 1ull));
    g->col = (uint32_t*)malloc(4ull * (m ? m : 1));
    if (!g->off || !g->col) goto out;
    for (uint64_t i = 0; i <= n; ++i) {
      if (lo[1] == 4) {
        uint32_t x;
        if (fread(&x, 4, 1, f) != 1) goto out;
        g->off[i] = x;
      } else if (fread(&g->off[i], 8, 1, f) != 1) goto out;
    }
    for (uint64_t i = (uint64_t)n + 1; i <= rows; ++i) g->off[i] = g->off[n];
    if (m && fread(g->col, 4, m, f) != m) goto out;
    rc = 0;
  }
out:
  fclose(f);
  return rc;
}

ORC_API void orc_free_csr(orc_csr* g) {
  free(g->off);
  free(g->col);
  memset(g, 0, sizeof *g);
}

/* ------------------------------------------------------------------------ */
/* uni                                                                       */
/* ------------------------------------------------------------------------ */
ORC_API void orc_cuts_uni(uint32_t n, uint32_t p, uint32_t* cuts) {
  for (uint32_t i = 0; i <= p; ++i) cuts[i] = (uint32_t)((long long)i * (long long)n / p);
}

/* ------------------------------------------------------------------------ */
/* n1d: Nicol's 1D partition of the offset array itself, 32-bit wrapping      */
/* ------------------------------------------------------------------------ */
static void n1d_probe(const uint32_t* prefix, uint32_t L, const uint32_t* UB, uint32_t np1,
                      uint32_t* out) {
  uint32_t sz = 1;
  out[0] = 0;
  while (sz < np1) {
    uint32_t r = UB[sz];
    uint32_t l = out[sz - 1];
    const uint32_t base = prefix[out[sz - 1]];
    while (l < r) {
      const uint32_t mid = (l + 1) + (r - (l + 1)) / 2;
      if ((uint32_t)(prefix[mid] - base) <= L) l = mid;
      else r = mid - 1;
    }
    out[sz++] = r;
    if (out[sz - 1] == out[sz - 2]) {
      while (sz < np1) { out[sz] = out[sz - 1]; ++sz; }
      break;
    }
  }
}

/* Nicol's optimal 1D cuts (nicol1d.hpp:106-195) of prefix[0..N]; prefix
 * holds one more zero at N+1 (the reference reads prefix[N+1] at its last
 * cut, nicol1d.hpp:144). Writes p+1 cuts. */
static void nicol_core(const uint32_t* prefix, uint32_t N, uint32_t p, uint32_t* cuts) {
  uint32_t* UB = (uint32_t*)malloc(4ull * (p + 1));
  uint32_t* pr = (uint32_t*)malloc(4ull * (p + 1));
  for (uint32_t i = 0; i <= p; ++i) UB[i] = N;
  uint32_t lower = prefix[N] / p, upper = prefix[N], pback = 0;
  for (uint32_t i = 1; i <= p; ++i) {
    uint32_t l = pback, r = UB[i];
    while (l < r) {
      const uint32_t mid = (l + 1) + (r - (l + 1)) / 2;
      const uint32_t b = prefix[mid] - prefix[pback];
      if (b > upper) r = mid - 1;
      else if (b < lower) l = mid;
      else {
        n1d_probe(prefix, b, UB, p + 1, pr);
        if (pr[p] < N) { l = mid; lower = b; }
        else { r = mid - 1; memcpy(UB, pr, 4ull * (p + 1)); upper = b; }
      }
    }
    const uint32_t b = prefix[l + 1] - prefix[pback];
    if (lower < b && b < upper) {
      n1d_probe(prefix, b, UB, p + 1, pr);
      if (pr[p] >= N) { memcpy(UB, pr, 4ull * (p + 1)); upper = b; }
      else lower = b;
    }
    pback = l;
  }
  n1d_probe(prefix, upper, UB, p + 1, cuts);
  free(UB);
  free(pr);
}

ORC_API void orc_cuts_n1d(const uint64_t* off, uint32_t n, uint32_t p, uint32_t* cuts) {
  const uint64_t asz = (uint64_t)n + 1;
  /* prefix of the 32-bit offsets (plus the zero past the end) */
  uint32_t* prefix = (uint32_t*)calloc(asz + 2, 4);
  for (uint64_t i = 0; i < asz; ++i) prefix[i + 1] = prefix[i] + (uint32_t)off[i];
  nicol_core(prefix, (uint32_t)asz, p, cuts);
  free(prefix);
}

/* ------------------------------------------------------------------------ */
/* rac: refine a cut (refine_a_cut.hpp:43-73, 129-223) on the whole matrix    */
/* ------------------------------------------------------------------------ */
/* One refinement: per row the largest number of its entries inside one
 * interval of prev, a 32-bit prefix over rows, then Nicol's cuts. */
static void rac_step(const uint64_t* off, const uint32_t* col, uint32_t n, const uint32_t* prev,
                     uint32_t p, uint32_t* out) {
  uint32_t* prefix = (uint32_t*)calloc((size_t)n + 2, 4);
  for (uint32_t i = 0; i < n; ++i) {
    uint32_t best = 0, run = 0, k = 0;
    for (uint64_t j = off[i]; j < off[i + 1]; ++j) {
      while (col[j] >= prev[k]) {
        if (run > best) best = run;
        run = 0;
        ++k;
      }
      ++run;
    }
    prefix[i + 1] = run > best ? run : best;
  }
  for (uint32_t i = 1; i <= n; ++i) prefix[i] += prefix[i - 1];
  nicol_core(prefix, n, p, out);
  free(prefix);
}

/* Largest tile entry count with the same cuts on rows and columns
 * (csr_matrix.hpp:281-322). */
static uint32_t rac_maxload(const uint64_t* off, const uint32_t* col, uint32_t n, const uint32_t* c,
                            uint32_t p) {
  uint32_t* load = (uint32_t*)calloc((size_t)p * p, 4);
  uint32_t rb = 0, mx = 0;
  for (uint32_t i = 0; i < n; ++i) {
    while (rb < p && i >= c[rb + 1]) ++rb;  /* band of row i: c[rb] <= i < c[rb+1] */
    for (uint64_t j = off[i]; j < off[i + 1]; ++j) {
      uint32_t cb = 0;
      while (cb + 1 < p && col[j] >= c[cb + 1]) ++cb;
      const uint32_t v = ++load[(size_t)rb * p + cb];
      if (v > mx) mx = v;
    }
  }
  free(load);
  return mx;
}

ORC_API void orc_cuts_rac(const uint64_t* off, const uint32_t* col, uint32_t n, uint32_t p, uint32_t* cuts) {
  const uint64_t m = off[n];
  /* the transpose, rows ascending */
  uint64_t* toff = (uint64_t*)calloc((size_t)n + 1, 8);
  uint32_t* tcol = (uint32_t*)malloc(4ull * (m ? m : 1));
  for (uint64_t e = 0; e < m; ++e) toff[col[e] + 1]++;
  for (uint32_t i = 1; i <= n; ++i) toff[i] += toff[i - 1];
  uint64_t* fill = (uint64_t*)malloc(8ull * (n ? n : 1));
  for (uint32_t i = 0; i < n; ++i) fill[i] = toff[i];
  for (uint32_t r = 0; r < n; ++r)
    for (uint64_t e = off[r]; e < off[r + 1]; ++e) tcol[fill[col[e]]++] = r;
  free(fill);
  uint32_t* cx = (uint32_t*)malloc(4ull * (p + 1));
  uint32_t* cy = (uint32_t*)malloc(4ull * (p + 1));
  uint32_t* t = (uint32_t*)malloc(4ull * (p + 1));
  for (uint32_t i = 0; i <= p; ++i) cx[i] = cy[i] = n;
  cx[0] = cy[0] = 0;
  rac_step(off, col, n, cx, p, t);
  const uint32_t mx = rac_maxload(off, col, n, t, p);
  rac_step(toff, tcol, n, cy, p, t);
  const uint32_t my = rac_maxload(off, col, n, t, p);
  const int iters = p > 10 ? (int)p : 10;
  uint32_t* c = mx <= my ? cx : cy;
  for (int it = 0; it < iters; ++it) {
    if (mx <= my) rac_step(off, col, n, c, p, t);
    else rac_step(toff, tcol, n, c, p, t);
    memcpy(c, t, 4ull * (p + 1));
  }
  memcpy(cuts, c, 4ull * (p + 1));
  cuts[p] = n + 1;  /* the driver, utils.hpp:288 */
  free(cx);
  free(cy);
  free(t);
  free(toff);
  free(tcol);
}

/* ------------------------------------------------------------------------ */
/* pal                                                                       */
/* ------------------------------------------------------------------------ */
/* MT19937-64 (Matsumoto & Nishimura 2000), the engine std::mt19937_64
 * names: w=64 n=312 m=156 r=31 a=0xB5026F5AA96619E9 u=29 d=0x5555555555555555
 * s=17 b=0x71D67FFFEDA60000 t=37 c=0xFFF7EEE000000000 l=43 f=6364136223846793005 */
typedef struct {
  uint64_t mt[312];
  int idx;
} mt64;

static void mt64_seed(mt64* s, uint64_t seed) {
  s->mt[0] = seed;
  for (int i = 1; i < 312; ++i)
    s->mt[i] = 6364136223846793005ull * (s->mt[i - 1] ^ (s->mt[i - 1] >> 62)) + (uint64_t)i;
  s->idx = 312;
}

static uint64_t mt64_next(mt64* s) {
  if (s->idx >= 312) {
    const uint64_t UM = 0xFFFFFFFF80000000ull, LM = 0x7FFFFFFFull;
    for (int i = 0; i < 312; ++i) {
      const uint64_t x = (s->mt[i] & UM) | (s->mt[(i + 1) % 312] & LM);
      uint64_t xa = x >> 1;
      if (x & 1) xa ^= 0xB5026F5AA96619E9ull;
      s->mt[i] = s->mt[(i + 156) % 312] ^ xa;
    }
    s->idx = 0;
  }
  uint64_t y = s->mt[s->idx++];
  y ^= (y >> 29) & 0x5555555555555555ull;
  y ^= (y << 17) & 0x71D67FFFEDA60000ull;
  y ^= (y << 37) & 0xFFF7EEE000000000ull;
  y ^= y >> 43;
  return y;
}

/* uniform double in [0,1) from one 64-bit draw: the draw rounded to double,
 * divided by 2^64, pulled below 1 when rounding reached it (the behaviour of
 * libstdc++'s generate_canonical<double,53> for a 64-bit engine). */
static double unit_double(mt64* s) {
  const double r = (double)mt64_next(s) / 18446744073709551616.0;
  return r >= 1.0 ? nextafter(1.0, 0.0) : r;
}

typedef struct {
  uint32_t S;
  uint32_t* byrow_off; /* samples with row < i: byrow_off[i] */
  /* column-major view: samples sorted by (col, row); Fenwick over that order
   * with sorted rows per node */
  uint32_t* colkeys;   /* sorted columns */
  uint64_t* node_start;
  uint32_t* node_rows;
} pal_counts;

static int cmp_u32(const void* a, const void* b) {
  const uint32_t x = *(const uint32_t*)a, y = *(const uint32_t*)b;
  return x < y ? -1 : x > y;
}
static int cmp_pair(const void* a, const void* b) {
  const uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b;
  return x < y ? -1 : x > y;
}

static void pal_counts_build(pal_counts* pc, const uint32_t* srow, const uint32_t* scol, uint32_t S) {
  pc->S = S;
  uint64_t* keys = (uint64_t*)malloc(8ull * (S ? S : 1));
  for (uint32_t x = 0; x < S; ++x) keys[x] = ((uint64_t)scol[x] << 32) | srow[x];
  qsort(keys, S, 8, cmp_pair);
  pc->colkeys = (uint32_t*)malloc(4ull * (S ? S : 1));
  uint32_t* rows = (uint32_t*)malloc(4ull * (S ? S : 1));
  for (uint32_t x = 0; x < S; ++x) { pc->colkeys[x] = (uint32_t)(keys[x] >> 32); rows[x] = (uint32_t)keys[x]; }
  free(keys);
  pc->node_start = (uint64_t*)calloc((size_t)S + 2, 8);
  for (uint64_t k = 1; k <= S; ++k) pc->node_start[k + 1] = pc->node_start[k] + (k & (~k + 1));
  pc->node_rows = (uint32_t*)malloc(4ull * (pc->node_start[S + 1] ? pc->node_start[S + 1] : 1));
  for (uint64_t k = 1; k <= S; ++k) {
    const uint64_t len = k & (~k + 1);
    uint32_t* dst = pc->node_rows + pc->node_start[k];
    memcpy(dst, rows + (k - len), 4 * len);
    qsort(dst, len, 4, cmp_u32);
  }
  free(rows);
}

static void pal_counts_free(pal_counts* pc) {
  free(pc->colkeys);
  free(pc->node_start);
  free(pc->node_rows);
}

/* #samples with row < i and col < j */
static uint32_t pal_query(const pal_counts* pc, uint32_t i, uint32_t j) {
  /* samples with col < j: a prefix of the column-major order */
  uint32_t lo = 0, hi = pc->S;
  while (lo < hi) {
    const uint32_t mid = lo + (hi - lo) / 2;
    if (pc->colkeys[mid] < j) lo = mid + 1;
    else hi = mid;
  }
  uint32_t s = 0;
  for (uint64_t r = lo; r > 0; r &= r - 1) {
    const uint32_t* b = pc->node_rows + pc->node_start[r];
    uint32_t l2 = 0, h2 = (uint32_t)(pc->node_start[r + 1] - pc->node_start[r]);
    while (l2 < h2) {
      const uint32_t mid = l2 + (h2 - l2) / 2;
      if (b[mid] < i) l2 = mid + 1;
      else h2 = mid;
    }
    s += l2;
  }
  return s;
}

/* load of tile (row band [r0,r1), col band [c0,c1)), by inclusion-exclusion
 * in 32-bit arithmetic as compute_loads does */
static uint32_t tile_load(const pal_counts* pc, uint32_t r0, uint32_t r1, uint32_t c0, uint32_t c1) {
  return pal_query(pc, r1, c1) - pal_query(pc, r0, c1) + pal_query(pc, r0, c0) - pal_query(pc, r1, c0);
}

/* max load over the last row band and last column band of p[0..sz) */
static uint32_t last_band_maxload(const pal_counts* pc, const uint32_t* p, uint32_t sz) {
  const uint32_t a = p[sz - 2], b = p[sz - 1];
  uint32_t colmax = 0, rowmax = 0;
  for (uint32_t x = 0; x + 1 < sz; ++x) {
    const uint32_t lc = tile_load(pc, p[x], p[x + 1], a, b);
    if (lc > colmax) colmax = lc;
    const uint32_t lr = tile_load(pc, a, b, p[x], p[x + 1]);
    if (lr > rowmax) rowmax = lr;
  }
  return colmax > rowmax ? colmax : rowmax;
}

static void pal_probe(const pal_counts* pc, uint32_t L, const uint32_t* UB, uint32_t np1, uint32_t* out) {
  uint32_t* p = (uint32_t*)malloc(4ull * (np1 + 1));
  uint32_t sz = 1;
  p[0] = 0;
  while (sz < np1) {
    uint32_t l = p[sz - 1], r = UB[sz];
    while (l < r) {
      const uint32_t mid = (l + 1) + (r - (l + 1)) / 2;
      p[sz] = mid;
      if (last_band_maxload(pc, p, sz + 1) <= L) l = mid;
      else r = mid - 1;
    }
    p[sz++] = l;
    if (p[sz - 1] == p[sz - 2]) {
      while (sz < np1) { p[sz] = p[sz - 1]; ++sz; }
      break;
    }
  }
  memcpy(out, p, 4ull * np1);
  free(p);
}

ORC_API void orc_cuts_pal(const uint64_t* off, const uint32_t* col, uint32_t n, uint32_t ncols,
                          uint64_t m, uint32_t p, uint32_t* cuts) {
  /* sample: keep probability min(1, (10p)^2 / nnz) with nnz as 32 bits */
  const uint32_t M32 = (uint32_t)m;
  double keep = ((double)p * 10.0) * ((double)p * 10.0) / (double)M32;
  if (!(keep < 1.0)) keep = 1.0;
  const uint64_t N = M32;
  unsigned char* mask = (unsigned char*)calloc(N ? N : 1, 1);
  mt64 gen;
  mt64_seed(&gen, 1091467301ull);
  if (keep >= 1.0) {
    memset(mask, 1, N);
  } else if (keep > 1.0 / 3) {
    for (uint64_t i = 0; i < N; ++i) mask[i] = unit_double(&gen) <= keep;
  } else {
    const double Q = log(1 - keep);
    for (uint64_t i = UINT64_MAX;;) {
      const uint64_t fc = (uint64_t)(log(1 - unit_double(&gen)) / Q);
      i += fc < N ? fc : N;
      if (++i < N) mask[i] = 1;
      else break;
    }
  }
  uint32_t S = 0;
  for (uint64_t x = 0; x < N; ++x) S += mask[x];
  uint32_t* srow = (uint32_t*)malloc(4ull * (S ? S : 1));
  uint32_t* scol = (uint32_t*)malloc(4ull * (S ? S : 1));
  uint32_t w = 0;
  for (uint32_t r = 0; r < n; ++r)
    for (uint64_t x = off[r]; x < off[r + 1]; ++x)
      if (x < N && mask[x]) { srow[w] = r; scol[w] = col[x]; ++w; }
  free(mask);
  pal_counts pc;
  pal_counts_build(&pc, srow, scol, S);
  free(srow);
  free(scol);

  const uint32_t NN = n > ncols ? n : ncols;
  const uint32_t M = S;
  uint32_t l = (M - 1 + p * p) / p / p;
  uint32_t r = (M + p - 1) / p + NN;
  if (M < r) r = M;
  uint32_t* UB = (uint32_t*)malloc(4ull * (p + 1));
  uint32_t* pr = (uint32_t*)malloc(4ull * (p + 1));
  for (uint32_t i = 0; i <= p; ++i) UB[i] = NN;
  while ((double)l + .5 < (double)r) {
    const uint32_t mid = l + (r - l) / 2;
    pal_probe(&pc, mid, UB, p + 1, pr);
    if (pr[p] < NN) l = mid + 1;
    else { r = mid; memcpy(UB, pr, 4ull * (p + 1)); }
  }
  pal_probe(&pc, r, UB, p + 1, cuts);
  free(UB);
  free(pr);
  pal_counts_free(&pc);
}

/* ------------------------------------------------------------------------ */
/* Reference block CSR (block_csr.cpp:286-376)                               */
/* ------------------------------------------------------------------------ */
typedef struct {
  uint32_t P, N;
  uint64_t E;
  uint32_t* cuts;     /* P+1 */
  uint64_t* tiles;    /* P*P+1: first slot of tile t in rows */
  uint32_t* rows;     /* (N+P)*P global offsets (32-bit, as Ordinal) */
  uint32_t* cols;     /* E tile-local columns */
  float* deg;         /* P*P */
} orc_bcsr;

ORC_API orc_bcsr* orc_bcsr_build(const uint64_t* off, const uint32_t* col, uint32_t n, uint64_t m,
                                 const uint32_t* cuts, uint32_t P) {
  orc_bcsr* B = (orc_bcsr*)calloc(1, sizeof *B);
  B->P = P;
  B->N = n;
  B->E = m;
  B->cuts = (uint32_t*)malloc(4ull * (P + 1));
  memcpy(B->cuts, cuts, 4ull * (P + 1));
  const uint64_t ntiles = (uint64_t)P * P + 1;
  const uint64_t nrpart = ((uint64_t)n + P) * P;
  B->tiles = (uint64_t*)malloc(8 * ntiles);
  B->rows = (uint32_t*)calloc(nrpart + 1, 4);
  B->cols = (uint32_t*)malloc(4ull * (m ? m : 1));
  B->deg = (float*)malloc(4ull * P * P);
  B->tiles[0] = 0;
  for (uint32_t i = 0; i < P; ++i)
    for (uint32_t j = 0; j < P; ++j)
      B->tiles[(uint64_t)i * P + j + 1] = B->tiles[(uint64_t)i * P + j] + (cuts[i + 1] - cuts[i]);

#pragma omp parallel for schedule(dynamic, 2048)
  for (int64_t u = 0; u < (int64_t)n; ++u) {
    uint32_t x = 0, y = 0;
    while ((uint64_t)u >= cuts[y + 1]) ++y;
    for (uint64_t v = off[u]; v < off[u + 1]; ++v) {
      while (col[v] >= cuts[x + 1]) ++x;
      __atomic_fetch_add(&B->rows[B->tiles[(uint64_t)y * P + x] + ((uint64_t)u - cuts[y]) + 1], 1u,
                         __ATOMIC_RELAXED);
    }
  }
  /* the prefix also covers the one extra word, which only P = 1 reaches */
  for (uint64_t i = 1; i <= nrpart; ++i) B->rows[i] += B->rows[i - 1];
#pragma omp parallel for schedule(dynamic, 2048)
  for (int64_t u = 0; u < (int64_t)n; ++u) {
    uint32_t x = 0, y = 0;
    while ((uint64_t)u >= cuts[y + 1]) ++y;
    uint64_t i = B->rows[B->tiles[(uint64_t)y * P] + ((uint64_t)u - cuts[y])];
    for (uint64_t v = off[u]; v < off[u + 1]; ++v) {
      while (col[v] >= cuts[x + 1]) {
        ++x;
        i = B->rows[B->tiles[(uint64_t)y * P + x] + ((uint64_t)u - cuts[y])];
      }
      B->cols[i++] = col[v] - cuts[x];
    }
  }
  /* ComputeTileProps (block_csr.cpp:257-284): 32-bit row-length sum over the
   * tile's rows, divided as float by the tile height */
  for (uint64_t t = 0; t < (uint64_t)P * P; ++t) {
    uint32_t s = 0;
    for (uint64_t j = B->tiles[t]; j < B->tiles[t + 1]; ++j) s += B->rows[j + 1] - B->rows[j];
    B->deg[t] = (float)s / (float)(uint32_t)(B->tiles[t + 1] - B->tiles[t]);
  }
  return B;
}

ORC_API void orc_bcsr_free(orc_bcsr* B) {
  if (!B) return;
  free(B->cuts);
  free(B->tiles);
  free(B->rows);
  free(B->cols);
  free(B->deg);
  free(B);
}

ORC_API float orc_bcsr_tile_deg(const orc_bcsr* B, uint32_t t) { return B->deg[t]; }
ORC_API uint32_t orc_bcsr_tile_nnz(const orc_bcsr* B, uint32_t t) {
  return B->rows[B->tiles[t + 1]] - B->rows[B->tiles[t]];
}
ORC_API uint32_t orc_bcsr_tile_height(const orc_bcsr* B, uint32_t t) {
  return (uint32_t)(B->tiles[t + 1] - B->tiles[t]);
}
/* global columns of local row r of tile t, written to out; returns count */
ORC_API uint32_t orc_bcsr_row(const orc_bcsr* B, uint32_t t, uint32_t r, uint32_t* out) {
  const uint64_t s = B->tiles[t] + r;
  const uint32_t b = B->rows[s], e = B->rows[s + 1];
  const uint32_t k = t % B->P;
  for (uint32_t x = b; x < e; ++x) out[x - b] = B->cols[x] + B->cuts[k];
  return e - b;
}

/* ------------------------------------------------------------------------ */
/* tasks (tasks.hpp:7-34)                                                    */
/* ------------------------------------------------------------------------ */
static float task_weight(const orc_bcsr* B, const uint32_t* t) {
  const uint32_t P = B->P;
  const uint32_t ab = t[0] * P + t[1], bc = t[1] * P + t[2], ac = t[0] * P + t[2];
  const float a = B->deg[bc], b = B->deg[ac];
  const float mx = (a < b) ? b : a; /* std::max */
  return (float)orc_bcsr_tile_nnz(B, ab) * mx;
}

/* order: 0 lexicographic, 1 weight descending (GPU build), 2 ascending (CPU build) */
ORC_API uint64_t orc_tasks(const orc_bcsr* B, int order, uint32_t* ijk) {
  const uint32_t P = B->P;
  uint64_t n = 0;
  for (uint32_t i = 0; i < P; ++i)
    for (uint32_t j = i; j < P; ++j)
      for (uint32_t k = j; k < P; ++k) { ijk[3 * n] = i; ijk[3 * n + 1] = j; ijk[3 * n + 2] = k; ++n; }
  if (order == 0) return n;
  /* stable insertion-merge sort by weight */
  float* w = (float*)malloc(4ull * n);
  uint64_t* idx = (uint64_t*)malloc(8ull * n);
  uint64_t* tmp = (uint64_t*)malloc(8ull * n);
  for (uint64_t q = 0; q < n; ++q) { w[q] = task_weight(B, ijk + 3 * q); idx[q] = q; }
  for (uint64_t width = 1; width < n; width *= 2) {
    for (uint64_t lo = 0; lo < n; lo += 2 * width) {
      uint64_t mid = lo + width < n ? lo + width : n, hi = lo + 2 * width < n ? lo + 2 * width : n;
      uint64_t a = lo, b = mid, o = lo;
      while (a < mid && b < hi) {
        const int take_b = order == 1 ? (w[idx[b]] > w[idx[a]]) : (w[idx[b]] < w[idx[a]]);
        tmp[o++] = take_b ? idx[b++] : idx[a++];
      }
      while (a < mid) tmp[o++] = idx[a++];
      while (b < hi) tmp[o++] = idx[b++];
    }
    memcpy(idx, tmp, 8ull * n);
  }
  uint32_t* out = (uint32_t*)malloc(12ull * n);
  for (uint64_t q = 0; q < n; ++q) memcpy(out + 3 * q, ijk + 3 * idx[q], 12);
  memcpy(ijk, out, 12ull * n);
  free(out);
  free(w);
  free(idx);
  free(tmp);
  return n;
}

/* ------------------------------------------------------------------------ */
/* count (bb_tc.cpp:118-139, 184-197, 255-273; cpu_*_kernel.cpp)              */
/* ------------------------------------------------------------------------ */
static long list_intersect(const orc_bcsr* B, uint64_t ab_src, uint64_t bc_src, uint64_t ac_src,
                           uint32_t f) {
  const uint32_t* rows = B->rows;
  const uint32_t* cols = B->cols;
  long c = 0;
  const uint32_t r_i = rows[ac_src + f + 1];
  for (uint32_t i = rows[ab_src + f]; i < rows[ab_src + f + 1]; ++i) {
    const uint32_t b = cols[i];
    uint32_t l_j = rows[bc_src + b], r_j = rows[bc_src + b + 1], l_i = rows[ac_src + f];
    while (l_i < r_i && l_j < r_j) {
      if (cols[l_i] == cols[l_j]) { ++l_i; ++l_j; ++c; }
      else if (cols[l_i] < cols[l_j]) {
        if (cols[l_j] > cols[r_i - 1]) break;
        ++l_i;
      } else {
        if (cols[l_i] > cols[r_j - 1]) break;
        ++l_j;
      }
    }
  }
  return c;
}

/* Rows f of band i in blocks of ORC_ROW_BLOCK consecutive rows, the blocks
 * b with b % stride == phase only (stride 1: the whole task, as the
 * reference counts it; larger strides time a systematic sample of every
 * task's rows that keeps the runs of consecutive rows a whole count reads). */
#define ORC_ROW_BLOCK 64u
#define FOR_SAMPLED_ROWS(f, h, stride, phase)                                              \
  for (uint32_t blk_ = (phase); (uint64_t)blk_ * ORC_ROW_BLOCK < (h); blk_ += (stride))     \
    for (uint32_t f = blk_ * ORC_ROW_BLOCK, fe_ = (f + ORC_ROW_BLOCK < (h) ? f + ORC_ROW_BLOCK : (h)); f < fe_; ++f)
static long task_count(const orc_bcsr* B, const uint32_t* t, uint32_t** hmap, uint64_t* hcap,
                       uint32_t stride, uint32_t phase) {
  const uint32_t P = B->P;
  const uint32_t ab = t[0] * P + t[1], bc = t[1] * P + t[2], ac = t[0] * P + t[2];
  if (B->deg[bc] == 0 || B->deg[ac] == 0 || B->deg[ab] == 0) return 0;
  const uint64_t ab_src = B->tiles[ab], bc_src = B->tiles[bc], ac_src = B->tiles[ac];
  const uint32_t h = (uint32_t)(B->tiles[ab + 1] - B->tiles[ab]);
  const uint32_t* rows = B->rows;
  const uint32_t* cols = B->cols;
  long c = 0;
  const float lim = 10000.f;
  if (B->deg[ac] > lim || B->deg[bc] > lim || B->deg[ab] > lim) {
    const uint32_t width = B->cuts[t[2] + 1] - B->cuts[t[2]];
    if (*hcap < width) {
      free(*hmap);
      *hmap = (uint32_t*)malloc(4ull * width);
      *hcap = width;
    }
    uint32_t* hm = *hmap;
    for (uint32_t x = 0; x < width; ++x) hm[x] = 0x7FFFFFFFu;
    FOR_SAMPLED_ROWS(f, h, stride, phase) {
      if (rows[ac_src + f + 1] == rows[ac_src + f]) continue;
      if (rows[ab_src + f + 1] - rows[ab_src + f] < 4) {
        c += list_intersect(B, ab_src, bc_src, ac_src, f);
        continue;
      }
      for (uint32_t x = rows[ac_src + f]; x < rows[ac_src + f + 1]; ++x) hm[cols[x]] = f;
      for (uint32_t x = rows[ab_src + f]; x < rows[ab_src + f + 1]; ++x) {
        const uint32_t b = cols[x];
        for (uint32_t y = rows[bc_src + b]; y < rows[bc_src + b + 1]; ++y) c += hm[cols[y]] == f;
      }
    }
  } else {
    FOR_SAMPLED_ROWS(f, h, stride, phase) {
      if (rows[ac_src + f + 1] == rows[ac_src + f]) continue;
      c += list_intersect(B, ab_src, bc_src, ac_src, f);
    }
  }
  return c;
}

/* Workers claim tasks from the back of the list (RunCpuKernels); with
 * stride > 1 only the rows f % stride == phase of each task. */
static uint64_t count_rows(const orc_bcsr* B, const uint32_t* ijk, uint64_t ntasks, int nthreads,
                           uint64_t* per_task, uint32_t stride, uint32_t phase);

ORC_API uint64_t orc_count(const orc_bcsr* B, const uint32_t* ijk, uint64_t ntasks, int nthreads,
                           uint64_t* per_task) {
  return count_rows(B, ijk, ntasks, nthreads, per_task, 1, 0);
}

ORC_API uint64_t orc_count_rows(const orc_bcsr* B, const uint32_t* ijk, uint64_t ntasks, int nthreads,
                                uint64_t* per_task, uint32_t stride, uint32_t phase) {
  if (stride == 0) stride = 1;
  return count_rows(B, ijk, ntasks, nthreads, per_task, stride, phase % stride);
}

static uint64_t count_rows(const orc_bcsr* B, const uint32_t* ijk, uint64_t ntasks, int nthreads,
                           uint64_t* per_task, uint32_t stride, uint32_t phase) {
  int64_t next = (int64_t)ntasks - 1;
  uint64_t total = 0;
  const int T = nthreads > 0 ? nthreads : omp_get_max_threads();
#pragma omp parallel num_threads(T) reduction(+ : total)
  {
    uint32_t* hmap = NULL;
    uint64_t hcap = 0;
    for (int64_t q = __atomic_fetch_sub(&next, 1, __ATOMIC_RELAXED); q >= 0;
         q = __atomic_fetch_sub(&next, 1, __ATOMIC_RELAXED)) {
      const long c = task_count(B, ijk + 3 * q, &hmap, &hcap, stride, phase);
      if (per_task) per_task[q] = (uint64_t)c;
      total += (uint64_t)c;
    }
    free(hmap);
  }
  return total;
}

/* ------------------------------------------------------------------------ */
/* Independent checks                                                        */
/* ------------------------------------------------------------------------ */
static uint32_t band_of(const uint32_t* cuts, uint32_t p, uint32_t v) {
  uint32_t lo = 0, hi = p; /* largest b with cuts[b] <= v */
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) / 2;
    if (cuts[mid] <= v) lo = mid;
    else hi = mid;
  }
  return lo;
}

/* Per-task counts straight from the definition: every triangle u<v<w of the
 * oriented graph adds one to cell (band u, band v, band w) of a p^3 array. */
ORC_API uint64_t orc_count_definition(const uint64_t* off, const uint32_t* col, uint32_t n,
                                      const uint32_t* cuts, uint32_t p, uint64_t* cube) {
  memset(cube, 0, 8ull * p * p * p);
  uint64_t total = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : total)
  for (int64_t u = 0; u < (int64_t)n; ++u) {
    const uint32_t bu = band_of(cuts, p, (uint32_t)u);
    for (uint64_t x = off[u]; x < off[u + 1]; ++x) {
      const uint32_t v = col[x], bv = band_of(cuts, p, v);
      uint64_t a = x + 1, b = off[v];
      while (a < off[u + 1] && b < off[v + 1]) {
        if (col[a] == col[b]) {
          const uint32_t bw = band_of(cuts, p, col[a]);
          __atomic_fetch_add(&cube[((uint64_t)bu * p + bv) * p + bw], 1ull, __ATOMIC_RELAXED);
          ++total;
          ++a;
          ++b;
        } else if (col[a] < col[b]) ++a;
        else ++b;
      }
    }
  }
  return total;
}

/* Triangles by brute force over the symmetric adjacency matrix (n <= 2048). */
ORC_API uint64_t orc_brute_force(const uint64_t* off, const uint32_t* col, uint32_t n) {
  if (n > 2048) return UINT64_MAX;
  unsigned char* A = (unsigned char*)calloc((size_t)n * n, 1);
  for (uint32_t u = 0; u < n; ++u)
    for (uint64_t x = off[u]; x < off[u + 1]; ++x) {
      A[(size_t)u * n + col[x]] = 1;
      A[(size_t)col[x] * n + u] = 1;
    }
  uint64_t t = 0;
  for (uint32_t a = 0; a < n; ++a)
    for (uint32_t b = a + 1; b < n; ++b)
      if (A[(size_t)a * n + b])
        for (uint32_t c = b + 1; c < n; ++c) t += A[(size_t)a * n + c] & A[(size_t)b * n + c];
  free(A);
  return t;
}
