// ref_harness.cpp - drives the reference's own CPU code (built from its
// sources, never copied) to print golden vectors as JSON.
//
// TEST INFRASTRUCTURE ONLY: built by oracle/Makefile into oracle/_ref/ and run
// only by tests/golden/make_golden.py in the container that has the
// reference sources. It is not part of the product.
//
// Usage: ref_harness <graph.bbtc> <uni|n1d|pal> <ncut> [cuts-only|tasks-parallel]
// Prints: {"n":..,"m":..,"p":..,"cuts":[..],"order_asc":[[i,j,k],..],
//          "tasks":[[i,j,k,count],..] (lexicographic),"total":..}
//
// It calls, in order, exactly what the reference CPU driver does
// (src/bbtc.cpp:12-74): utils::BlockGraph -> ReadGraph (sparsify, partition,
// cuts[p] = N+1, BlockCSR), tasks::TaskList + SortTasks, BBTc. Per-task
// counts come from a BBTc constructed over a one-task list (bb_tc.cpp:3).
// With tasks-parallel, the one-task BBTc objects run concurrently, one per
// OpenMP thread (each owns its hash maps; the block CSR is only read), and
// the whole-list count ("total_all") is skipped: for graphs whose tasks are
// too many to count one after another.
#include <cmath>

#include "bb_tc.h"
#include "block_csr.h"
#include "definitions.h"
#include "utils.hpp"
#include "tasks.hpp"

int main(int argc, const char** argv) {
  if (argc < 4) {
    std::cerr << "usage: ref_harness <graph.bbtc> <uni|n1d|pal|rac> <ncut> [cuts-only|tasks-parallel]\n";
    return 2;
  }
  std::string file = argv[1];
  std::string alg = argv[2];
  const int p = std::stoi(argv[3]);
  const bool cuts_only = argc > 4 && std::string(argv[4]) == "cuts-only";
  const bool parallel = argc > 4 && std::string(argv[4]) == "tasks-parallel";

  BlockCSR<Ordinal, Vertex> st = utils::BlockGraph<Ordinal, Vertex, double>(file, alg, -1, p);
  const Ordinal P = st.GetNoOfCuts();
  std::cout << "{\"n\":" << st.GetNoOfVertices() << ",\"m\":" << st.GetNoOfEdges()
            << ",\"p\":" << P << ",\"partitioner\":\"" << alg << "\",\"cuts\":[";
  for (Ordinal b = 0; b <= P; ++b) std::cout << (b ? "," : "") << st.GetCuts()[b];
  std::cout << "]";
  if (!cuts_only) {
    auto sorted = tasks::TaskList(P);
    tasks::SortTasks(sorted, st);
    std::cout << ",\"order_asc\":[";
    for (size_t q = 0; q < sorted.size(); ++q) {
      const size_t i = sorted[q][0] / P, j = sorted[q][0] % P, k = sorted[q][1] % P;
      std::cout << (q ? "," : "") << "[" << i << "," << j << "," << k << "]";
    }
    std::cout << "]";
    auto lex = tasks::TaskList(P);
    std::vector<unsigned long long> per(lex.size(), 0);
    auto one_task = [&](size_t q) {
      std::vector<std::vector<size_t>> one{lex[q]};
      BBTc bbtc(&st, one, P);
      bbtc.InitFlags();
      bbtc.Separator = 0;
      bbtc.ResetFlags();
      per[q] = bbtc.NoOfTriangles();
      bbtc.CleanUp();
    };
    if (parallel) {
#pragma omp parallel for schedule(dynamic, 1)
      for (size_t q = 0; q < lex.size(); ++q) one_task(q);
    } else {
      for (size_t q = 0; q < lex.size(); ++q) one_task(q);
    }
    unsigned long long total = 0;
    std::cout << ",\"tasks\":[";
    for (size_t q = 0; q < lex.size(); ++q) {
      total += per[q];
      const size_t i = lex[q][0] / P, j = lex[q][0] % P, k = lex[q][1] % P;
      std::cout << (q ? "," : "") << "[" << i << "," << j << "," << k << "," << per[q] << "]";
    }
    std::cout << "],\"total\":" << total;
    if (!parallel) {
      BBTc all(&st, lex, P);
      all.InitFlags();
      all.Separator = 0;
      all.ResetFlags();
      std::cout << ",\"total_all\":" << all.NoOfTriangles();
      all.CleanUp();
    }
  }
  std::cout << "}" << std::endl;
  st.CleanUp();
  return 0;
}
