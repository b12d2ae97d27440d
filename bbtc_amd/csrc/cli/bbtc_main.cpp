// bbtc - command-line drop-in for the reference's bbtc binary (GPU build).
//
// Same flags, defaults, validation messages and report lines as
// src/bbtc.cu:15-90 with utils::ParseArguments (include/utils/utils.hpp:
// 21-154) and hooks.hpp:13-52, driving libbbtc.so through its C-ABI:
//   load .bbtc (text input is converted first and saved as <path>.bbtc, as
//   utils.hpp:171-268 does) -> partition (uni|n1d|pal) -> block layout ->
//   task list sorted by weight, descending -> per repeat: [copy] + count.
// Differences, all additive or fail-safe:
//   * text conversion drops duplicate edges unless --keep-duplicates is
//     given (the counters accept simple graphs only); METIS .graph input is
//     refused;
//   * argument errors exit with status 2 (the reference exits 0);
//   * "Cuda Initial Memory Allocations and Stream Creation" is reported as
//     "HIP Initial Memory Allocations and Stream Creation";
//   * extra flags: --kernel auto|merge|bitmap|dense|rdense|lane|tuned, --threads n, --cuts-file path,
//     --per-task-out path, --keep-duplicates, --save-bcsr path (a .bcsr
//     --graph is read back with its own cuts).
#include <omp.h>
#include <strings.h>

#include <chrono>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "bbtc.h"

namespace {

struct Params {
  std::string file = "-";
  std::string part = "pal";
  int ngpu = 0;
  int nstream = 4;
  int border = 0;
  int ncut = 32;
  int repeat = 5;
  bool copy_first = false;
  std::string kernel = "auto";
  int threads = 0;
  std::string cuts_file;
  std::string per_task_out;
  bool keep_duplicates = false;
  std::string save_bcsr;
};

void usage() {
  std::cout << "bbTC: A Block Based Triangle Counting Algorithm (MI355X build)\n"
            << "syntax: bbtc --graph <path> [OPTION].. [--help]\n"
            << "Options:                                                           (Defaults:)\n"
            << "  --graph <path>        Path for the graph file.                   ()\n"
            << "                          Supported formats: edge lists, matrix\n"
            << "                          market files and binary CSR.\n"
            << "  --partitioner <str>   Name of the partitioner algorithm.         (pal)\n"
            << "                          Algorithms: [uni | n1d | pal]\n"
            << "                          uni: uniform, checkerboard partitioning\n"
            << "                          n1d: Nicol's 1D Partitioning\n"
            << "                          rac: refine a cut partitioning\n"
            << "                          pal: probe a load partitioning\n"
            << "  --ncut <int>          Number of cuts                             (32)\n"
            << "  --ngpu <int>          Number of GPUs to use                      (all)\n"
            << "                          Note: if 0 then only CPUs used.\n"
            << "  --nstream <int>       Number of streams per GPU                  (4)\n"
            << "                          Note: Cannot be 0.\n"
            << "  --border <int>        Work queue border for GPUs                 (0)\n"
            << "                          Note: if set to -1, GPUs process all.\n"
            << "  --copy-first          If set then copy graph to GPUs first       (false)\n"
            << "  --repeat <int>        Run bbTC algorithm '--repeat' times        (5)\n"
            << "  --kernel <str>        auto | merge | bitmap | dense | rdense | lane | tuned (auto)\n"
            << "  --threads <int>       Host worker threads                        (all)\n"
            << "  --cuts-file <path>    Read the p+1 cuts from a text file instead of partitioning\n"
            << "  --per-task-out <path> Write 'i j k count' for every task\n"
            << "  --keep-duplicates     Keep repeated edges when converting text input\n"
            << "  --save-bcsr <path>    Write the block layout (.bcsr) for later runs\n"
            << "\nExamples:\n"
            << "  bbtc --graph email-EuAll_adj.mtx --partitioner n1d --ncut 4\n"
            << "  bbtc --graph email-EuAll_adj.txt --ncut 4 --copy-first --border 4\n";
}

[[noreturn]] void bad(const std::string& msg) {
  std::cout << msg << std::endl;
  usage();
  std::exit(2);
}

int to_int(const char* s, const std::string& flag) {
  char* end = nullptr;
  const long v = std::strtol(s, &end, 10);
  if (!s[0] || *end || v < INT_MIN || v > INT_MAX) bad("Wrong value '" + std::string(s) + "' for " + flag);
  return int(v);
}

// utils.hpp:79-85
void check_int(const std::string& name, int v, int lo = 0, int hi = INT_MAX) {
  if (v <= lo || v >= hi) {
    std::ostringstream o;
    o << name << " has to be set in between: " << lo << " < " << name << " < " << hi;
    bad(o.str());
  }
}

bool file_exists(const std::string& p) {
  std::ifstream f(p.c_str());
  return f.good();
}

Params parse(int argc, char** argv) {
  Params p;
  if (argc < 3) {
    usage();
    std::exit(0);
  }
  int ndev = 0;
  bbtc_device_count(&ndev);
  p.ngpu = ndev;
  for (int i = 1; i < argc; ++i) {
    const char* a = argv[i];
    auto val = [&](const char* flag) -> const char* {
      if (i + 1 >= argc) bad(std::string("Missing value for ") + flag);
      return argv[++i];
    };
    if (!strcasecmp(a, "--graph")) p.file = val("--graph");
    else if (!strcasecmp(a, "--ngpu")) {
      const int t = to_int(val("--ngpu"), "--ngpu");
      check_int("--ngpu", t, -1, ndev + 1);
      p.ngpu = t;
    } else if (!strcasecmp(a, "--nstream")) {
      p.nstream = to_int(val("--nstream"), "--nstream");
      check_int("--nstream", p.nstream);
    } else if (!strcasecmp(a, "--border")) {
      p.border = to_int(val("--border"), "--border");
      check_int("--border", p.border, -2);
    } else if (!strcasecmp(a, "--repeat")) {
      p.repeat = to_int(val("--repeat"), "--repeat");
      check_int("--repeat", p.repeat);
    } else if (!strcasecmp(a, "--copy-first")) p.copy_first = true;
    else if (!strcasecmp(a, "--ncut")) {
      p.ncut = to_int(val("--ncut"), "--ncut");
      check_int("--ncut", p.ncut, 1);
    } else if (!strcasecmp(a, "--partitioner")) {
      p.part = val("--partitioner");
      if (p.part != "pal" && p.part != "n1d" && p.part != "uni" && p.part != "rac")
        bad("Wrong partitioning algorithm.");
    } else if (!strcasecmp(a, "--kernel")) {
      p.kernel = val("--kernel");
      if (p.kernel != "auto" && p.kernel != "merge" && p.kernel != "bitmap" && p.kernel != "dense" &&
          p.kernel != "rdense" && p.kernel != "lane" && p.kernel != "tuned")
        bad("Wrong kernel.");
    } else if (!strcasecmp(a, "--threads")) {
      p.threads = to_int(val("--threads"), "--threads");
      check_int("--threads", p.threads);
    } else if (!strcasecmp(a, "--cuts-file")) p.cuts_file = val("--cuts-file");
    else if (!strcasecmp(a, "--per-task-out")) p.per_task_out = val("--per-task-out");
    else if (!strcasecmp(a, "--keep-duplicates")) p.keep_duplicates = true;
    else if (!strcasecmp(a, "--save-bcsr")) p.save_bcsr = val("--save-bcsr");
    else if (!strcasecmp(a, "--help")) {
      usage();
      std::exit(0);
    } else bad("Wrong argument '" + std::string(a) + "' given.");
  }
  if (!file_exists(p.file)) bad("Graph doesn't exist");
  return p;
}

// hooks.hpp: region names and seconds with 4 decimals, printed at the end
struct Fields {
  std::vector<std::pair<std::string, double>> f;
  std::chrono::steady_clock::time_point t0;
  std::string name;
  void begin(const std::string& n) { name = n; t0 = std::chrono::steady_clock::now(); }
  void end() {
    const auto us = std::chrono::duration_cast<std::chrono::microseconds>(std::chrono::steady_clock::now() - t0).count();
    f.emplace_back(name, us / 1e6);
  }
  void print() const {
    for (auto& x : f) std::cout << x.first << " : " << std::fixed << std::setprecision(4) << x.second << std::endl;
  }
};

[[noreturn]] void die(bbtc_status s, const std::string& what) {
  std::cerr << "bbtc: " << what << ": " << bbtc_strerror(s) << ": " << bbtc_last_error() << std::endl;
  std::exit(1);
}

bool ends_with(const std::string& s, const std::string& t) {
  return s.size() >= t.size() && s.compare(s.size() - t.size(), t.size(), t) == 0;
}

}  // namespace

int main(int argc, char** argv) {
  Params P = parse(argc, argv);
  std::string path = P.file;
  const bool bcsr = ends_with(path, ".bcsr");  // utils.hpp:300-301
  if (bcsr) {
  } else if (file_exists(path + ".bbtc")) {  // utils.hpp:168-170: a sibling "<path>.bbtc" is preferred
    path += ".bbtc";
  } else if (!ends_with(path, ".bbtc")) {
    // utils.hpp:171-268: convert the text graph and save it beside the input
    std::cerr << "# Saving binary CSR graph.." << std::endl;
    const bbtc_status cs = bbtc_convert_text(path.c_str(), (path + ".bbtc").c_str(), P.keep_duplicates ? 0 : 1);
    if (cs) {
      std::cerr << "bbtc: converting " << path << ": " << bbtc_strerror(cs) << ": " << bbtc_last_error()
                << std::endl;
      return 1;
    }
    path += ".bbtc";
  }
  bbtc_hgraph* g = nullptr;
  bbtc_layout* L = nullptr;
  bbtc_status st = BBTC_OK;
  Fields F;
  uint32_t n = 0, ncols = 0;
  uint64_t m = 0;
  uint32_t p = uint32_t(P.ncut);
  if (bcsr) {
    // a stored block layout: its own cuts, no partitioning (block_csr.cpp:215-255)
    F.begin("Partitioning and Layout construction (s)");
    st = bbtc_layout_load_bcsr(path.c_str(), &g, &L);
    if (st) die(st, "reading " + path);
    F.end();
    bbtc_hgraph_info(g, &n, &ncols, &m);
    p = bbtc_layout_ncut(L);
    std::cerr << "# Graph read.. N: " << n << ", NNZ: " << m << std::endl;
  } else {
    st = bbtc_hgraph_load(path.c_str(), &g);
    if (st) die(st, "reading " + path);
    bbtc_hgraph_info(g, &n, &ncols, &m);
    std::cerr << "# Graph read.. N: " << n << ", NNZ: " << m << std::endl;

    F.begin("Partitioning and Layout construction (s)");
    std::vector<uint32_t> cuts(p + 1);
    if (!P.cuts_file.empty()) {
      std::ifstream cf(P.cuts_file);
      for (uint32_t b = 0; b <= p; ++b)
        if (!(cf >> cuts[b])) {
          std::cerr << "bbtc: --cuts-file needs " << p + 1 << " integers" << std::endl;
          return 1;
        }
    } else {
      const bbtc_part kind = P.part == "uni"   ? BBTC_PART_UNI
                             : P.part == "n1d" ? BBTC_PART_N1D
                             : P.part == "rac" ? BBTC_PART_RAC
                                               : BBTC_PART_PAL;
      st = bbtc_partition(g, kind, p, cuts.data());
      if (st) die(st, "partitioning");
    }
    st = bbtc_layout_build(g, cuts.data(), p, &L);
    if (st) die(st, "building the block layout");
    F.end();
  }
  if (!P.save_bcsr.empty()) {
    st = bbtc_layout_save_bcsr(L, P.save_bcsr.c_str());
    if (st) die(st, "writing " + P.save_bcsr);
  }

  uint64_t ntasks = 0;
  std::vector<bbtc_task> tasks(bbtc_task_count(p));
  st = bbtc_tasks_build(L, BBTC_ORDER_DESC, tasks.data(), tasks.size(), &ntasks);
  if (st) die(st, "building the task list");

  std::cout << "=====================================" << std::endl;
  std::cout << "Graph file: " << P.file << std::endl;
  std::cout << "Number of Vertices: " << n << std::endl;
  std::cout << "Number of Edges: " << m << std::endl;
  std::cout << "Number of Cuts: " << p << std::endl;
  std::cout << "Number of Tasks: " << ntasks << std::endl;
  std::cout << "Number of GPUs: " << P.ngpu << std::endl;
  std::cout << "Number of Workers: " << (P.threads > 0 ? P.threads : omp_get_max_threads()) << std::endl;
  std::cout << "=====================================" << std::endl;

  // bbtc.cu:54-62
  uint64_t separator = 0;
  if (P.border == -1 || uint64_t(P.border) > ntasks) separator = ntasks;
  else separator = uint64_t(P.border);

  const bbtc_kernel kern = P.kernel == "merge"    ? BBTC_KERNEL_MERGE
                           : P.kernel == "bitmap" ? BBTC_KERNEL_BITMAP
                           : P.kernel == "dense"  ? BBTC_KERNEL_DENSE
                           : P.kernel == "rdense" ? BBTC_KERNEL_RDENSE
                           : P.kernel == "lane"   ? BBTC_KERNEL_LANE
                                                  : BBTC_KERNEL_AUTO;
  // --kernel tuned: per-task forms measured once on the first upload
  std::vector<uint8_t> choice;
  bbtc_ctx* ctx = nullptr;
  F.begin("HIP Initial Memory Allocations and Stream Creation");
  if (P.ngpu > 0) {
    st = bbtc_ctx_create(0, nullptr, &ctx);
    if (st) die(st, "creating the device context");
  }
  F.end();

  std::vector<uint64_t> per(ntasks, 0);
  uint64_t total = 0;
  bbtc_dgraph* dg = nullptr;
  if (ctx && P.kernel == "tuned") {
    F.begin("Autotune kernel forms (s)");
    st = bbtc_graph_upload(ctx, L, &dg);
    if (st) die(st, "uploading the graph");
    choice.resize(std::max<uint64_t>(ntasks, 1));
    st = bbtc_autotune(ctx, dg, tasks.data(), ntasks, choice.data(), nullptr);
    if (st) die(st, "autotuning");
    F.end();
  }
  for (int r = 0; r < P.repeat; ++r) {
    if (dg) {  // as ResetFlags (bb_tc_helper.cpp:31-50), every repeat starts without device tiles
      bbtc_graph_release(dg);
      dg = nullptr;
    }
    if (ctx && P.copy_first) {
      F.begin("Copy tiles to device (s)");
      st = bbtc_graph_upload(ctx, L, &dg);
      if (st) die(st, "uploading the graph");
      F.end();
    }
    F.begin("Triangle Counting Execution Time (s)");
    if (ctx && !dg) {
      st = bbtc_graph_upload(ctx, L, &dg);
      if (st) die(st, "uploading the graph");
    }
    st = bbtc_count_hybrid(ctx, dg, L, tasks.data(), ntasks, separator, P.threads, P.nstream, kern,
                           choice.empty() ? nullptr : choice.data(), per.data(), &total, nullptr);
    if (st) die(st, "counting");
    if (ctx) bbtc_ctx_synchronize(ctx);
    F.end();
  }

  std::cout << "Number of Triangles: " << total << std::endl;
  F.print();

  if (!P.per_task_out.empty()) {
    std::ofstream o(P.per_task_out);
    for (uint64_t q = 0; q < ntasks; ++q)
      o << tasks[q].i << " " << tasks[q].j << " " << tasks[q].k << " " << per[q] << "\n";
  }
  if (dg) bbtc_graph_release(dg);
  if (ctx) bbtc_ctx_destroy(ctx);
  bbtc_layout_free(L);
  bbtc_hgraph_free(g);
  return 0;
}
