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
//   * extra flags: --kernel auto|merge|bitmap|dense|rdense|lane|rows|tuned, --threads n, --cuts-file path,
//     --per-task-out path, --keep-duplicates, --save-bcsr path (a .bcsr
//     --graph is read back with its own cuts), --forms-out / --forms-in path
//     (the autotuned per-task forms saved for, and reused by, later runs);
//   * extra modes (the reference counts one graph per run): --batch <list>
//     counts every graph a list file names in one batched launch, and
//     --serve <list> runs them as requests (with arrival times) through the
//     continuous-batching server.
#include <omp.h>
#include <strings.h>

#include <algorithm>
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
  bool device_layout = false;  // build the block layout on the GPU
  std::string kernel = "auto";
  int threads = 0;
  std::string cuts_file;
  std::string per_task_out;
  bool keep_duplicates = false;
  std::string save_bcsr;
  std::string forms_out, forms_in;  // autotuned forms written / reused
  bool ncut_given = false;
  std::string batch_list, serve_list;  // list files of the extra modes
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
            << "  --device-layout       Build the block layout on the GPU from the CSR (timed\n"
            << "                          with partitioning); it stays resident for every\n"
            << "                          repeat, so no tile copies follow\n"
            << "  --repeat <int>        Run bbTC algorithm '--repeat' times        (5)\n"
            << "  --kernel <str>        auto | merge | bitmap | dense | rdense | lane | rows | tuned (auto)\n"
            << "  --threads <int>       Host worker threads                        (all)\n"
            << "  --cuts-file <path>    Read the p+1 cuts from a text file instead of partitioning\n"
            << "  --per-task-out <path> Write 'i j k count' for every task\n"
            << "  --keep-duplicates     Keep repeated edges when converting text input\n"
            << "  --save-bcsr <path>    Write the block layout (.bcsr) for later runs\n"
            << "  --forms-out <path>    With --kernel tuned: write the measured per-task forms\n"
            << "  --forms-in <path>     Count with per-task forms written by --forms-out for\n"
            << "                          the same cuts (tuned speed, no autotuning)\n"
            << "  --batch <list>        Count every graph listed ('path [ncut]' per line) in one\n"
            << "                          batched launch; without an ncut column or --ncut the\n"
            << "                          cut count per graph comes from measured sample graphs\n"
            << "  --serve <list>        Serve the listed graphs ('path [arrival_s]' per line) as\n"
            << "                          requests through the continuous-batching server\n"
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
    else if (!strcasecmp(a, "--device-layout")) p.device_layout = true;
    else if (!strcasecmp(a, "--ncut")) {
      p.ncut = to_int(val("--ncut"), "--ncut");
      check_int("--ncut", p.ncut, 1);
      p.ncut_given = true;
    } else if (!strcasecmp(a, "--partitioner")) {
      p.part = val("--partitioner");
      if (p.part != "pal" && p.part != "n1d" && p.part != "uni" && p.part != "rac")
        bad("Wrong partitioning algorithm.");
    } else if (!strcasecmp(a, "--kernel")) {
      p.kernel = val("--kernel");
      if (p.kernel != "auto" && p.kernel != "merge" && p.kernel != "bitmap" && p.kernel != "dense" &&
          p.kernel != "rdense" && p.kernel != "lane" && p.kernel != "rows" && p.kernel != "tuned")
        bad("Wrong kernel.");
    } else if (!strcasecmp(a, "--threads")) {
      p.threads = to_int(val("--threads"), "--threads");
      check_int("--threads", p.threads);
    } else if (!strcasecmp(a, "--cuts-file")) p.cuts_file = val("--cuts-file");
    else if (!strcasecmp(a, "--per-task-out")) p.per_task_out = val("--per-task-out");
    else if (!strcasecmp(a, "--keep-duplicates")) p.keep_duplicates = true;
    else if (!strcasecmp(a, "--save-bcsr")) p.save_bcsr = val("--save-bcsr");
    else if (!strcasecmp(a, "--forms-out")) p.forms_out = val("--forms-out");
    else if (!strcasecmp(a, "--forms-in")) p.forms_in = val("--forms-in");
    else if (!strcasecmp(a, "--batch")) p.batch_list = val("--batch");
    else if (!strcasecmp(a, "--serve")) p.serve_list = val("--serve");
    else if (!strcasecmp(a, "--help")) {
      usage();
      std::exit(0);
    } else bad("Wrong argument '" + std::string(a) + "' given.");
  }
  if (!p.batch_list.empty() && !p.serve_list.empty()) bad("--batch and --serve are exclusive");
  if (p.device_layout && (p.ngpu == 0 || !p.batch_list.empty() || !p.serve_list.empty()))
    bad("--device-layout builds one graph's layout on a GPU (--ngpu 1, no --batch or --serve)");
  if (!p.forms_out.empty() || !p.forms_in.empty()) {
    if (!p.batch_list.empty() || !p.serve_list.empty()) bad("--forms-out and --forms-in count one graph");
    if (p.ngpu == 0) bad("--forms-out and --forms-in need a GPU (--ngpu 1)");
    if (!p.forms_out.empty() && p.kernel != "tuned") bad("--forms-out needs --kernel tuned");
    if (!p.forms_in.empty() && p.kernel != "auto" && p.kernel != "tuned")
      bad("--forms-in sets the forms; --kernel must be auto or tuned");
    if (!p.forms_in.empty() && !p.forms_out.empty()) bad("--forms-out and --forms-in are exclusive");
    if (!p.forms_in.empty() && !file_exists(p.forms_in)) bad("Forms file doesn't exist");
  }
  const std::string& list = !p.batch_list.empty() ? p.batch_list : p.serve_list;
  if (!list.empty()) {
    if (!file_exists(list)) bad("List file doesn't exist");
    if (p.ngpu == 0) bad("--batch and --serve need a GPU (--ngpu 1)");
    return p;
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

// Fingerprint of a graph for the forms file: n, m and a hash of its offsets
// and columns (chunks hashed in parallel, combined in chunk order).
struct GraphPrint {
  uint64_t n = 0, m = 0, hash = 0;
};

GraphPrint graph_print(const bbtc_hgraph* g) {
  GraphPrint f;
  uint32_t n = 0, nc = 0;
  bbtc_hgraph_info(g, &n, &nc, &f.m);
  f.n = n;
  const uint64_t* off = bbtc_hgraph_offsets(g);
  const uint32_t* col = bbtc_hgraph_cols(g);
  auto mix = [](uint64_t h, uint64_t x) {
    h ^= x + 0x9E3779B97F4A7C15ull + (h << 6) + (h >> 2);
    return h * 0xBF58476D1CE4E5B9ull;
  };
  constexpr uint64_t kChunk = 1ull << 20;
  const uint64_t nchunks = (f.m + kChunk - 1) / kChunk;
  std::vector<uint64_t> part(nchunks, 0);
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < int64_t(nchunks); ++c) {
    uint64_t h = uint64_t(c);
    const uint64_t e = std::min<uint64_t>(f.m, uint64_t(c + 1) * kChunk);
    for (uint64_t x = uint64_t(c) * kChunk; x < e; ++x) h = mix(h, col[x]);
    part[c] = h;
  }
  uint64_t h = mix(f.n, f.m);
  for (uint64_t x = 0; x <= f.n; ++x) h = mix(h, off[x]);
  for (uint64_t v : part) h = mix(h, v);
  f.hash = h;
  return f;
}

// Forms file: "bbtc-forms 2", then "<ncut> <ntasks> <n> <m> <hash>" (the
// graph's fingerprint), the ncut+1 cuts, and one choice per task in task
// order (form 1 merge, 2 bitmap, 3 dense, 4 rdense, 5 lane, 0 the default
// rule, plus 16 times the rows' halvings), all as text. The fingerprint and
// the cuts tie it to one graph and layout.
void write_forms(const std::string& path, const bbtc_hgraph* g, const bbtc_layout* L,
                 const std::vector<uint8_t>& choice, uint64_t ntasks) {
  const uint32_t p = bbtc_layout_ncut(L);
  const uint32_t* cuts = bbtc_layout_cuts(L);
  const GraphPrint f = graph_print(g);
  std::ofstream o(path);
  o << "bbtc-forms 2\n" << p << " " << ntasks << " " << f.n << " " << f.m << " " << f.hash << "\n";
  for (uint32_t b = 0; b <= p; ++b) o << cuts[b] << (b == p ? "\n" : " ");
  for (uint64_t q = 0; q < ntasks; ++q) o << int(choice[q]) << (q + 1 == ntasks ? "\n" : " ");
  if (!o) {
    std::cerr << "bbtc: cannot write " << path << std::endl;
    std::exit(1);
  }
}

std::vector<uint8_t> read_forms(const std::string& path, const bbtc_hgraph* g, const bbtc_layout* L,
                                uint64_t ntasks) {
  auto fail = [&](const std::string& why) {
    std::cerr << "bbtc: forms file " << path << ": " << why << std::endl;
    std::exit(1);
  };
  std::ifstream f(path);
  std::string magic;
  int version = 0;
  uint64_t p = 0, nt = 0, n = 0, m = 0, hash = 0;
  if (!(f >> magic >> version) || magic != "bbtc-forms" || (version != 1 && version != 2))
    fail("not a bbtc forms file");
  if (version == 1) fail("written without a graph fingerprint; write it again with --forms-out");
  if (!(f >> p >> nt >> n >> m >> hash)) fail("truncated header");
  if (p != bbtc_layout_ncut(L) || nt != ntasks) fail("written for another number of cuts or tasks");
  const GraphPrint gp = graph_print(g);
  if (n != gp.n || m != gp.m || hash != gp.hash) fail("written for another graph");
  const uint32_t* cuts = bbtc_layout_cuts(L);
  for (uint64_t b = 0; b <= p; ++b) {
    uint64_t c = 0;
    if (!(f >> c) || c != cuts[b]) fail("written for other cuts");
  }
  std::vector<uint8_t> choice(std::max<uint64_t>(ntasks, 1), 0);
  for (uint64_t q = 0; q < ntasks; ++q) {
    int c = -1;
    if (!(f >> c) || c < 0 || (c & 7) > 5 || (c >> 4) > 2 || (c & 8))
      fail("bad or missing form for task " + std::to_string(q));
    choice[q] = uint8_t(c);
  }
  return choice;
}

bbtc_part part_kind(const std::string& s) {
  return s == "uni" ? BBTC_PART_UNI : s == "n1d" ? BBTC_PART_N1D : s == "rac" ? BBTC_PART_RAC : BBTC_PART_PAL;
}

bbtc_kernel kernel_kind(const std::string& s) {
  return s == "merge"    ? BBTC_KERNEL_MERGE
         : s == "bitmap" ? BBTC_KERNEL_BITMAP
         : s == "dense"  ? BBTC_KERNEL_DENSE
         : s == "rdense" ? BBTC_KERNEL_RDENSE
         : s == "lane"   ? BBTC_KERNEL_LANE
         : s == "rows"   ? BBTC_KERNEL_ROWS
                         : BBTC_KERNEL_AUTO;
}

// ---- --batch / --serve -------------------------------------------------------

struct ListEntry {
  std::string path;
  double value = 0;  // ncut (--batch) or arrival seconds (--serve)
  bool has_value = false;
};

// 'path [number]' per line; blank lines and lines starting with '#' skipped
std::vector<ListEntry> read_list(const std::string& file) {
  std::ifstream in(file);
  std::vector<ListEntry> out;
  std::string line;
  while (std::getline(in, line)) {
    std::istringstream ls(line);
    ListEntry e;
    if (!(ls >> e.path) || e.path[0] == '#') continue;
    if (ls >> e.value) e.has_value = true;
    out.push_back(e);
  }
  if (out.empty()) {
    std::cerr << "bbtc: " << file << " names no graph" << std::endl;
    std::exit(1);
  }
  return out;
}

// A .bbtc file (or its sibling <path>.bbtc), else text converted in memory
// (duplicate edges dropped).
bbtc_hgraph* load_listed(const std::string& path) {
  bbtc_hgraph* g = nullptr;
  bbtc_status st;
  if (ends_with(path, ".bcsr")) {
    std::cerr << "bbtc: " << path << ": block layouts (.bcsr) are not accepted in a list" << std::endl;
    std::exit(1);
  }
  if (ends_with(path, ".bbtc")) st = bbtc_hgraph_load(path.c_str(), &g);
  else if (file_exists(path + ".bbtc")) st = bbtc_hgraph_load((path + ".bbtc").c_str(), &g);
  else st = bbtc_hgraph_load_text(path.c_str(), &g);
  if (st) die(st, "reading " + path);
  return g;
}

uint64_t edges_of(const bbtc_hgraph* g) {
  uint32_t n = 0, nc = 0;
  uint64_t m = 0;
  bbtc_hgraph_info(g, &n, &nc, &m);
  return m;
}

// Cut count per graph: the list's column (use_column), else --ncut, else a
// rule from measured device work on sample graphs spread over the sizes
// (summed clocks for a batch, whose graphs share one launch; launch time for
// requests served one by one: by_time).
std::vector<uint32_t> cut_counts(bbtc_ctx* ctx, const std::vector<bbtc_hgraph*>& g,
                                 const std::vector<ListEntry>& e, const Params& P, bool use_column,
                                 bool by_time) {
  const size_t n = g.size();
  std::vector<uint32_t> ncut(n, uint32_t(P.ncut));
  bool need_rule = false;
  for (size_t x = 0; x < n; ++x) {
    if (use_column && e[x].has_value) {
      if (e[x].value < 1 || e[x].value > 65535) {
        std::cerr << "bbtc: bad cut count for " << e[x].path << std::endl;
        std::exit(1);
      }
      ncut[x] = uint32_t(e[x].value);
    } else if (!P.ncut_given) {
      need_rule = true;
    }
  }
  if (!need_rule) return ncut;
  std::vector<size_t> by_m(n);
  for (size_t x = 0; x < n; ++x) by_m[x] = x;
  std::stable_sort(by_m.begin(), by_m.end(), [&](size_t a, size_t b) { return edges_of(g[a]) < edges_of(g[b]); });
  std::vector<const bbtc_hgraph*> sample;
  const size_t step = std::max<size_t>(1, n / 64);
  for (size_t x = 0; x < n; x += step) sample.push_back(g[by_m[x]]);
  const uint32_t cands[] = {1, 2, 4, 8};
  uint32_t rule[64];
  const bbtc_status st = bbtc_tune_ncut(ctx, sample.data(), sample.size(), cands, 4, part_kind(P.part),
                                        by_time ? BBTC_TUNE_BY_TIME : 0, rule, nullptr);
  if (st) die(st, "choosing cut counts");
  for (size_t x = 0; x < n; ++x)
    if (!(use_column && e[x].has_value)) ncut[x] = bbtc_rule_ncut(rule, edges_of(g[x]));
  return ncut;
}

double percentile(std::vector<double> v, double q) {
  if (v.empty()) return 0;
  std::sort(v.begin(), v.end());
  const double pos = q / 100.0 * double(v.size() - 1);
  const size_t lo = size_t(pos), hi = std::min(v.size() - 1, lo + 1);
  return v[lo] + (v[hi] - v[lo]) * (pos - double(lo));
}

int run_batch(const Params& P) {
  const std::vector<ListEntry> e = read_list(P.batch_list);
  Fields F;
  bbtc_ctx* ctx = nullptr;
  F.begin("HIP Initial Memory Allocations and Stream Creation");
  bbtc_status st = bbtc_ctx_create(0, nullptr, &ctx);
  if (st) die(st, "creating the device context");
  F.end();
  std::vector<bbtc_hgraph*> g;
  for (auto& x : e) g.push_back(load_listed(x.path));
  F.begin("Partitioning and Layout construction (s)");
  const std::vector<uint32_t> ncut = cut_counts(ctx, g, e, P, true, false);
  std::vector<bbtc_layout*> L(g.size(), nullptr);
  st = bbtc_layouts_build(g.data(), g.size(), ncut.data(), part_kind(P.part), L.data());
  if (st) die(st, "building the block layouts");
  F.end();
  F.begin("Batch upload and plan (s)");
  bbtc_batch* b = nullptr;
  st = bbtc_batch_create(ctx, L.data(), L.size(), P.kernel == "tuned" ? BBTC_KERNEL_AUTO : kernel_kind(P.kernel),
                         BBTC_BATCH_SMALL_GRAPHS_FIRST, &b);
  if (st) die(st, "creating the batch");
  if (P.kernel == "tuned" && (st = bbtc_batch_autotune(b, nullptr))) die(st, "autotuning the batch");
  if ((st = bbtc_batch_capture(b))) die(st, "capturing the batch launch");
  F.end();
  uint64_t m = 0;
  for (auto* x : g) m += edges_of(x);
  std::cout << "=====================================" << std::endl;
  std::cout << "Graph list: " << P.batch_list << std::endl;
  std::cout << "Number of Graphs: " << g.size() << std::endl;
  std::cout << "Number of Edges: " << m << std::endl;
  std::cout << "Number of Tasks: " << bbtc_batch_ntasks(b) << std::endl;
  std::cout << "Number of GPUs: " << P.ngpu << std::endl;
  std::cout << "=====================================" << std::endl;
  std::vector<uint64_t> per(g.size(), 0);
  std::vector<double> lat(g.size(), 0);
  for (int r = 0; r < P.repeat; ++r) {
    F.begin("Batch Counting Execution Time (s)");
    if ((st = bbtc_batch_launch(b))) die(st, "launching the batch");
    if ((st = bbtc_batch_results(b, per.data(), nullptr, lat.data()))) die(st, "reading the batch results");
    F.end();
  }
  uint64_t total = 0;
  for (size_t x = 0; x < g.size(); ++x) {
    std::cout << "Graph " << x << " " << e[x].path << ": cuts " << ncut[x] << ", triangles " << per[x]
              << ", done after " << std::fixed << std::setprecision(3) << lat[x] << " ms" << std::endl;
    total += per[x];
  }
  std::cout << "Number of Triangles: " << total << std::endl;
  F.print();
  bbtc_batch_destroy(b);
  for (auto* l : L) bbtc_layout_free(l);
  for (auto* x : g) bbtc_hgraph_free(x);
  bbtc_ctx_destroy(ctx);
  return 0;
}

int run_serve(const Params& P) {
  std::vector<ListEntry> e = read_list(P.serve_list);
  // the server takes arrivals in time order
  std::vector<size_t> order(e.size());
  for (size_t x = 0; x < e.size(); ++x) order[x] = x;
  std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return e[a].value < e[b].value; });
  Fields F;
  bbtc_ctx* ctx = nullptr;
  F.begin("HIP Initial Memory Allocations and Stream Creation");
  bbtc_status st = bbtc_ctx_create(0, nullptr, &ctx);
  if (st) die(st, "creating the device context");
  bbtc_server* srv = nullptr;
  if ((st = bbtc_server_create(ctx, nullptr, &srv))) die(st, "creating the server");
  F.end();
  std::vector<bbtc_hgraph*> g;
  for (auto& x : e) g.push_back(load_listed(x.path));
  F.begin("Partitioning and Layout construction (s)");
  const std::vector<uint32_t> ncut = cut_counts(ctx, g, e, P, false, true);
  std::vector<bbtc_layout*> L(g.size(), nullptr);
  st = bbtc_layouts_build(g.data(), g.size(), ncut.data(), part_kind(P.part), L.data());
  if (st) die(st, "building the block layouts");
  std::vector<std::vector<bbtc_task>> tasks(g.size());
  for (size_t x = 0; x < g.size(); ++x) {
    tasks[x].resize(bbtc_task_count(ncut[x]));
    uint64_t got = 0;
    if ((st = bbtc_tasks_build(L[x], BBTC_ORDER_DESC, tasks[x].data(), tasks[x].size(), &got)))
      die(st, "building a task list");
    tasks[x].resize(got);
  }
  F.end();
  uint64_t m = 0;
  for (auto* x : g) m += edges_of(x);
  std::cout << "=====================================" << std::endl;
  std::cout << "Request list: " << P.serve_list << std::endl;
  std::cout << "Number of Requests: " << g.size() << std::endl;
  std::cout << "Number of Edges: " << m << std::endl;
  std::cout << "Number of GPUs: " << P.ngpu << std::endl;
  std::cout << "=====================================" << std::endl;
  const bbtc_kernel kern = P.kernel == "tuned" ? BBTC_KERNEL_AUTO : kernel_kind(P.kernel);
  std::vector<uint64_t> id(g.size(), 0);
  std::vector<bbtc_server_result> res(g.size());
  F.begin("Serving makespan (s)");
  if ((st = bbtc_server_reset_clock(srv))) die(st, "resetting the server clock");
  for (size_t x : order)
    if ((st = bbtc_server_submit(srv, L[x], tasks[x].data(), tasks[x].size(), kern, e[x].value, nullptr, &id[x])))
      die(st, "submitting " + e[x].path);
  if ((st = bbtc_server_wait_idle(srv))) die(st, "serving");
  F.end();
  std::vector<bbtc_server_result> out(g.size());
  uint64_t got = 0;
  if ((st = bbtc_server_poll(srv, out.data(), out.size(), 0, &got))) die(st, "collecting results");
  uint64_t total = 0;
  std::vector<double> lat;
  for (size_t x = 0; x < g.size(); ++x) {
    const bbtc_server_result* r = nullptr;
    for (uint64_t k = 0; k < got; ++k)
      if (out[k].id == id[x]) r = &out[k];
    if (!r || r->status != BBTC_OK) {
      std::cerr << "bbtc: request " << x << " (" << e[x].path << ") failed" << std::endl;
      return 1;
    }
    const double ms = (r->done_s - r->arrival_s) * 1e3;
    lat.push_back(ms);
    total += r->total;
    std::cout << "Request " << x << " " << e[x].path << ": cuts " << ncut[x] << ", triangles " << r->total
              << ", arrival " << std::fixed << std::setprecision(3) << r->arrival_s << " s, latency " << ms
              << " ms" << std::endl;
  }
  std::cout << "Number of Triangles: " << total << std::endl;
  F.print();
  std::cout << "Latency p50 (ms) : " << std::fixed << std::setprecision(4) << percentile(lat, 50) << std::endl;
  std::cout << "Latency p99 (ms) : " << std::fixed << std::setprecision(4) << percentile(lat, 99) << std::endl;
  bbtc_server_destroy(srv);
  for (auto* l : L) bbtc_layout_free(l);
  for (auto* x : g) bbtc_hgraph_free(x);
  bbtc_ctx_destroy(ctx);
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  Params P = parse(argc, argv);
  if (!P.batch_list.empty()) return run_batch(P);
  if (!P.serve_list.empty()) return run_serve(P);
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
  bbtc_dgraph* dg = nullptr;
  bbtc_ctx* ctx = nullptr;
  bbtc_status st = BBTC_OK;
  Fields F;
  uint32_t n = 0, ncols = 0;
  uint64_t m = 0;
  uint32_t p = uint32_t(P.ncut);
  if (bcsr && P.device_layout) {
    std::cerr << "bbtc: --device-layout builds the layout from a CSR graph, not from a .bcsr file" << std::endl;
    return 1;
  }
  // bbtc.cu:54-62: tasks [border, n) are shared with the host workers
  auto separator_of = [&](uint64_t ntasks) -> uint64_t {
    if (P.border == -1 || uint64_t(P.border) > ntasks) return ntasks;
    return uint64_t(P.border);
  };
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

    if (P.device_layout) {
      // the layout is built on the device: its context comes first
      F.begin("HIP Initial Memory Allocations and Stream Creation");
      st = bbtc_ctx_create(0, nullptr, &ctx);
      if (st) die(st, "creating the device context");
      F.end();
    }
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
      st = bbtc_partition(g, part_kind(P.part), p, cuts.data());
      if (st) die(st, "partitioning");
    }
    if (P.device_layout) {
      // the host arrays only where host code reads them: the host tail of
      // the task queue and the .bcsr writer
      const bool host = separator_of(bbtc_task_count(p)) < bbtc_task_count(p) || !P.save_bcsr.empty();
      st = bbtc_graph_build(ctx, g, cuts.data(), p, host ? BBTC_BUILD_HOST_ARRAYS : 0u, &L, &dg);
      if (st) die(st, "building the block layout on the device");
      std::cerr << "# Layout built on the device" << (host ? " (host arrays copied back)" : "") << std::endl;
    } else {
      st = bbtc_layout_build(g, cuts.data(), p, &L);
      if (st) die(st, "building the block layout");
    }
    F.end();
  }
  if (!P.save_bcsr.empty()) {
    st = bbtc_layout_save_bcsr(L, P.save_bcsr.c_str());
    if (st) die(st, "writing " + P.save_bcsr);
  }

  // With a GPU the device takes the heaviest tasks from the front of a
  // descending list (tasks.hpp:25-28); with none, the order of the
  // reference's CPU build (ascending, tasks.hpp:29-32), so the workers, who
  // claim from the back, start with the heaviest tasks.
  uint64_t ntasks = 0;
  std::vector<bbtc_task> tasks(bbtc_task_count(p));
  st = bbtc_tasks_build(L, P.ngpu > 0 ? BBTC_ORDER_DESC : BBTC_ORDER_ASC, tasks.data(), tasks.size(),
                        &ntasks);
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

  const uint64_t separator = separator_of(ntasks);

  const bbtc_kernel kern = kernel_kind(P.kernel);
  // --kernel tuned: per-task forms measured once on the first upload
  std::vector<uint8_t> choice;
  if (!ctx) {
    F.begin("HIP Initial Memory Allocations and Stream Creation");
    if (P.ngpu > 0) {
      st = bbtc_ctx_create(0, nullptr, &ctx);
      if (st) die(st, "creating the device context");
    }
    F.end();
  }

  std::vector<uint64_t> per(ntasks, 0);
  uint64_t total = 0;
  if (ctx && !P.forms_in.empty()) {
    choice = read_forms(P.forms_in, g, L, ntasks);
  } else if (ctx && P.kernel == "tuned") {
    F.begin("Autotune kernel forms (s)");
    if (!dg) {
      st = bbtc_graph_upload(ctx, L, &dg);
      if (st) die(st, "uploading the graph");
    }
    choice.resize(std::max<uint64_t>(ntasks, 1));
    st = bbtc_autotune(ctx, dg, tasks.data(), ntasks, choice.data(), nullptr);
    if (st) die(st, "autotuning");
    F.end();
    if (!P.forms_out.empty()) write_forms(P.forms_out, g, L, choice, ntasks);
  }
  for (int r = 0; r < P.repeat; ++r) {
    // as ResetFlags (bb_tc_helper.cpp:31-50), every repeat starts without
    // device tiles; a layout built on the device stays there
    if (dg && !P.device_layout) {
      bbtc_graph_release(dg);
      dg = nullptr;
    }
    if (ctx && P.copy_first && !dg) {
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
