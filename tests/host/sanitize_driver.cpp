// Host-side C-ABI exercise for the sanitizer test (tests/test_host_sanitizers.py):
// generate, save and load a graph, partition it with every partitioner, build
// layouts, count on the CPU, compute algorithmic bytes, round-trip .bcsr files,
// convert text edge lists (a dense and a sparse id space), count a wedge
// that merges a long bc row, and take an error path; everything is released, so
// AddressSanitizer's leak check sees no leak.
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>
#include "bbtc.h"
#define CK(x) do { bbtc_status s_ = (x); if (s_) { std::printf("fail %s: %s\n", #x, bbtc_last_error()); return 1; } } while (0)
int main(int argc, char** argv) {
  const char* dir = argc > 1 ? argv[1] : "/tmp";
  bbtc_hgraph* g = nullptr;
  CK(bbtc_hgraph_gen_rmat(12, 8, 0.57, 0.19, 0.19, 1, &g));
  std::string p = std::string(dir) + "/a.bbtc";
  CK(bbtc_hgraph_save(g, p.c_str()));
  bbtc_hgraph* g2 = nullptr;
  CK(bbtc_hgraph_load(p.c_str(), &g2));
  for (bbtc_part k : {BBTC_PART_UNI, BBTC_PART_N1D, BBTC_PART_PAL, BBTC_PART_RAC}) {
    for (uint32_t P : {1u, 3u, 8u}) {
      std::vector<uint32_t> cuts(P + 1);
      CK(bbtc_partition(g2, k, P, cuts.data()));
      bbtc_layout* L = nullptr;
      CK(bbtc_layout_build(g2, cuts.data(), P, &L));
      std::vector<bbtc_task> t(bbtc_task_count(P));
      uint64_t n = 0;
      CK(bbtc_tasks_build(L, BBTC_ORDER_DESC, t.data(), t.size(), &n));
      std::vector<uint64_t> per(n), bytes(n);
      uint64_t tot = 0, tb = 0, w = 0;
      CK(bbtc_count_cpu(L, t.data(), n, 4, per.data(), &tot));
      CK(bbtc_tasks_bytes(L, t.data(), n, 4, bytes.data(), &tb, &w));
      std::string b = std::string(dir) + "/a.bcsr";
      CK(bbtc_layout_save_bcsr(L, b.c_str()));
      bbtc_hgraph* g3 = nullptr; bbtc_layout* L3 = nullptr;
      CK(bbtc_layout_load_bcsr(b.c_str(), &g3, &L3));
      bbtc_layout_free(L3); bbtc_hgraph_free(g3);
      bbtc_tile_info ti;
      CK(bbtc_layout_tile(L, 0, P - 1, &ti));
      bbtc_layout_free(L);
      std::printf("part %d P %u triangles %llu\n", int(k), P, (unsigned long long)tot);
    }
  }
  // text ingest
  std::string tp = std::string(dir) + "/e.txt";
  FILE* f = std::fopen(tp.c_str(), "w");
  std::fprintf(f, "%% comment\n0 1\n1 2\n2 0\n2 3\n3 0\n1 3\n");
  std::fclose(f);
  CK(bbtc_convert_text(tp.c_str(), (tp + ".bbtc").c_str(), 1));
  bbtc_hgraph* gt = nullptr;
  CK(bbtc_hgraph_load_text(tp.c_str(), &gt));
  bbtc_hgraph_free(gt);
  // an id space far sparser than the edges (rows built with atomic slots)
  f = std::fopen(tp.c_str(), "w");
  std::fprintf(f, "0 5000000\n5000000 7\n7 0\n");
  std::fclose(f);
  CK(bbtc_hgraph_load_text(tp.c_str(), &gt));
  bbtc_hgraph_free(gt);
  // one ac row of 2 entries against a bc row of 300 (merged, not probed)
  {
    std::vector<uint64_t> off = {0, 2, 302};
    std::vector<uint32_t> col = {1, 2};
    for (uint32_t c = 2; c < 302; ++c) col.push_back(c);
    for (int i = 0; i < 300; ++i) off.push_back(302);
    bbtc_hgraph* gl = nullptr;
    CK(bbtc_hgraph_from_csr(302, 302, col.size(), off.data(), col.data(), &gl));
    uint32_t cuts[2];
    CK(bbtc_partition(gl, BBTC_PART_UNI, 1, cuts));
    bbtc_layout* L = nullptr;
    CK(bbtc_layout_build(gl, cuts, 1, &L));
    bbtc_task t{0, 0, 0};
    uint64_t per = 0, tot = 0;
    CK(bbtc_count_cpu(L, &t, 1, 2, &per, &tot));
    if (tot != 1) return 3;
    bbtc_layout_free(L);
    bbtc_hgraph_free(gl);
  }
  // error paths
  bbtc_hgraph* bad = nullptr;
  if (bbtc_hgraph_load("/nonexistent.bbtc", &bad) == BBTC_OK) return 2;
  bbtc_hgraph_free(g2);
  bbtc_hgraph_free(g);
  std::printf("done\n");
  return 0;
}
