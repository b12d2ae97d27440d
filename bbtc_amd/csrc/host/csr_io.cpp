// .bbtc (PIGO-CSR-v2) reader/writer and CSR validation.
//
// File layout, as written by pigo::CSR::save (PIGO csr.impl.hpp:440-473) and
// read by read_bin_ (csr.impl.hpp:475-512):
//   "PIGO-CSR-v2" (11 bytes, no terminator)
//   u8 L_size, u8 O_size
//   n (L), m (O), nrows (L), ncols (L)
//   offsets[n+1] (O each), endpoints[m] (L each)
// The reference requires L = O = 4 (csr.impl.hpp:485-486); this reader takes
// L = 4 with O = 4 or 8. Like sarma::Matrix (csr_matrix.hpp:139-145) a row
// count field larger than n pads the offsets with empty rows.
#include <omp.h>

#include <cstdio>
#include <vector>

#include "graph.hpp"

namespace bbtc {

static const char kHeader[] = "PIGO-CSR-v2";
static const size_t kHeaderLen = 11;

void validate_csr(const Csr& g) {
  if (g.off.size() != size_t(g.n) + 1) fail(BBTC_ERR_FORMAT, "offset array size != n+1");
  if (g.off[0] != 0) fail(BBTC_ERR_FORMAT, "offsets[0] != 0");
  if (g.off[g.n] != g.m) fail(BBTC_ERR_FORMAT, "offsets[n] != m");
  if (uint64_t(g.n) + 1 > 0xFFFFFFFFull) fail(BBTC_ERR_FORMAT, "n+1 must fit in 32 bits");
  const uint64_t limit = std::max<uint64_t>(g.n, g.ncols);
  int bad = 0;
  int64_t bad_row = -1;
#pragma omp parallel for schedule(dynamic, 4096) reduction(| : bad)
  for (int64_t u = 0; u < int64_t(g.n); ++u) {
    const uint64_t b = g.off[u], e = g.off[u + 1];
    if (e < b || e > g.m) { bad |= 1; continue; }
    uint64_t prev = uint64_t(u);  // every column must exceed the row id
    for (uint64_t x = b; x < e; ++x) {
      const uint64_t c = g.col[x];
      if (c <= prev || c >= limit) {
        bad |= (c >= limit) ? 2 : (c <= uint64_t(u) ? 4 : 8);
#pragma omp critical
        bad_row = u;
        break;
      }
      prev = c;
    }
  }
  if (bad & 1) fail(BBTC_ERR_FORMAT, "offsets decrease or exceed m");
  if (bad & 2) fail(BBTC_ERR_FORMAT, "column id out of range (row " + std::to_string(bad_row) + ")");
  if (bad & 4)
    fail(BBTC_ERR_FORMAT, "entry on or below the diagonal (row " + std::to_string(bad_row) +
                              "): input must be the upper triangle");
  if (bad & 8)
    fail(BBTC_ERR_FORMAT, "row " + std::to_string(bad_row) +
                              " is not strictly increasing (duplicates or unsorted)");
}

namespace {
struct File {
  FILE* f = nullptr;
  ~File() { if (f) std::fclose(f); }
};

void read_exact(FILE* f, void* dst, size_t bytes, const std::string& what) {
  char* p = static_cast<char*>(dst);
  while (bytes) {
    size_t chunk = std::min<size_t>(bytes, size_t(1) << 30);
    size_t got = std::fread(p, 1, chunk, f);
    if (got != chunk) fail(BBTC_ERR_FORMAT, "truncated .bbtc file while reading " + what);
    p += got;
    bytes -= got;
  }
}

void write_exact(FILE* f, const void* src, size_t bytes) {
  const char* p = static_cast<const char*>(src);
  while (bytes) {
    size_t chunk = std::min<size_t>(bytes, size_t(1) << 30);
    if (std::fwrite(p, 1, chunk, f) != chunk) fail(BBTC_ERR_IO, "short write");
    p += chunk;
    bytes -= chunk;
  }
}
}  // namespace

Csr load_bbtc(const std::string& path) {
  File fh;
  fh.f = std::fopen(path.c_str(), "rb");
  if (!fh.f) fail(BBTC_ERR_IO, "cannot open graph file: " + path);
  std::fseek(fh.f, 0, SEEK_END);
  const long long fsize = std::ftell(fh.f);
  std::fseek(fh.f, 0, SEEK_SET);

  char hdr[kHeaderLen];
  if (fsize < (long long)kHeaderLen + 2) fail(BBTC_ERR_FORMAT, "file too short for a .bbtc header");
  read_exact(fh.f, hdr, kHeaderLen, "header");
  if (std::memcmp(hdr, kHeader, kHeaderLen) != 0)
    fail(BBTC_ERR_FORMAT, "not a PIGO-CSR-v2 file (bad magic)");
  uint8_t lo[2];
  read_exact(fh.f, lo, 2, "template sizes");
  const unsigned L = lo[0], O = lo[1];
  if (L != 4) fail(BBTC_ERR_FORMAT, "unsupported label size " + std::to_string(L) + " (need 4)");
  if (O != 4 && O != 8)
    fail(BBTC_ERR_FORMAT, "unsupported offset size " + std::to_string(O) + " (need 4 or 8)");

  uint32_t n = 0, nrows = 0, ncols = 0;
  uint64_t m = 0;
  read_exact(fh.f, &n, 4, "n");
  if (O == 4) { uint32_t m32; read_exact(fh.f, &m32, 4, "m"); m = m32; }
  else read_exact(fh.f, &m, 8, "m");
  read_exact(fh.f, &nrows, 4, "nrows");
  read_exact(fh.f, &ncols, 4, "ncols");

  const long long need = (long long)kHeaderLen + 2 + 12 + O + (long long)O * (n + 1ll) + 4ll * (long long)m;
  if (fsize < need)
    fail(BBTC_ERR_FORMAT, "truncated .bbtc: header promises " + std::to_string(need) +
                              " bytes, file has " + std::to_string(fsize));

  const uint32_t rows = std::max(n, nrows);
  if (uint64_t(rows) + 1 > 0xFFFFFFFFull) fail(BBTC_ERR_FORMAT, "too many rows");
  Csr g;
  g.n = rows;
  g.ncols = ncols;
  g.m = m;
  g.off.resize(size_t(rows) + 1);
  if (O == 8) {
    read_exact(fh.f, g.off.data(), 8ull * (n + 1ull), "offsets");
  } else {
    Buf<uint32_t> tmp(size_t(n) + 1);
    read_exact(fh.f, tmp.data(), 4ull * (n + 1ull), "offsets");
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i <= int64_t(n); ++i) g.off[i] = tmp[i];
  }
  for (uint64_t r = uint64_t(n) + 1; r <= rows; ++r) g.off[r] = g.off[n];
  g.col.resize(m);
  if (m) read_exact(fh.f, g.col.data(), 4ull * m, "endpoints");
  validate_csr(g);
  return g;
}

void save_bbtc(const Csr& g, const std::string& path) {
  File fh;
  fh.f = std::fopen(path.c_str(), "wb");
  if (!fh.f) fail(BBTC_ERR_IO, "cannot create " + path);
  const uint8_t L = 4, O = (g.m < (1ull << 32)) ? 4 : 8;
  write_exact(fh.f, kHeader, kHeaderLen);
  write_exact(fh.f, &L, 1);
  write_exact(fh.f, &O, 1);
  const uint32_t n = g.n;
  write_exact(fh.f, &n, 4);
  if (O == 4) { uint32_t m32 = uint32_t(g.m); write_exact(fh.f, &m32, 4); }
  else write_exact(fh.f, &g.m, 8);
  write_exact(fh.f, &n, 4);          // nrows
  write_exact(fh.f, &g.ncols, 4);    // ncols
  if (O == 8) {
    write_exact(fh.f, g.off.data(), 8ull * (n + 1ull));
  } else {
    Buf<uint32_t> tmp(size_t(n) + 1);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i <= int64_t(n); ++i) tmp[i] = uint32_t(g.off[i]);
    write_exact(fh.f, tmp.data(), 4ull * (n + 1ull));
  }
  if (g.m) write_exact(fh.f, g.col.data(), 4ull * g.m);
  if (std::fflush(fh.f) != 0) fail(BBTC_ERR_IO, "flush failed for " + path);
}

Csr copy_csr(uint32_t n, uint32_t ncols, uint64_t m, const uint64_t* off, const uint32_t* col) {
  if (!off || (m && !col)) fail(BBTC_ERR_INVALID, "null CSR array");
  Csr g;
  g.n = n;
  g.ncols = ncols;
  g.m = m;
  g.off.resize(size_t(n) + 1);
  std::memcpy(g.off.data(), off, 8ull * (n + 1ull));
  g.col.resize(m);
  if (m) std::memcpy(g.col.data(), col, 4ull * m);
  validate_csr(g);
  return g;
}

}  // namespace bbtc
