// Shared host-side types for the block-based triangle counter.
//
// Everything here is plain C++17 with OpenMP; no HIP types appear in this
// header so the host library builds with g++ alone.
#pragma once

#include <algorithm>
#include <chrono>
#include <cstddef>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <new>
#include <stdexcept>
#include <string>
#include <utility>

#include "bbtc.h"

namespace bbtc {

// Error carried through the C++ layer and turned into a bbtc_status at the
// C-ABI boundary. Nothing in this library calls exit().
struct Error : std::runtime_error {
  bbtc_status code;
  Error(bbtc_status c, const std::string& msg) : std::runtime_error(msg), code(c) {}
};

[[noreturn]] inline void fail(bbtc_status c, const std::string& msg) { throw Error(c, msg); }

// Uninitialised, 256-byte aligned array of trivially copyable T. Large graph
// arrays (billions of entries) must not pay for std::vector's zero fill.
// adopt() hands it memory of another allocator (page-locked host memory from
// the device library) with the function that frees it; pinned() says so.
template <class T>
class Buf {
 public:
  using FreeFn = void (*)(void*);
  Buf() = default;
  explicit Buf(size_t n) { resize(n); }
  Buf(const Buf&) = delete;
  Buf& operator=(const Buf&) = delete;
  Buf(Buf&& o) noexcept : p_(o.p_), n_(o.n_), free_(o.free_) { o.p_ = nullptr; o.n_ = 0; o.free_ = nullptr; }
  Buf& operator=(Buf&& o) noexcept {
    if (this != &o) {
      release();
      p_ = o.p_; n_ = o.n_; free_ = o.free_;
      o.p_ = nullptr; o.n_ = 0; o.free_ = nullptr;
    }
    return *this;
  }
  ~Buf() { release(); }

  void adopt(T* p, size_t n, FreeFn free_fn) {
    release();
    p_ = p;
    n_ = n;
    free_ = free_fn;
  }
  bool pinned() const { return free_ != nullptr; }

  void resize(size_t n) {
    release();
    if (n == 0) return;
    void* q = nullptr;
    size_t bytes = ((n * sizeof(T)) + 255) & ~size_t(255);
    if (posix_memalign(&q, 256, bytes) != 0 || !q)
      fail(BBTC_ERR_NOMEM, "host allocation of " + std::to_string(bytes) + " bytes failed");
    p_ = static_cast<T*>(q);
    n_ = n;
  }
  void fill(const T& v) { for (size_t i = 0; i < n_; ++i) p_[i] = v; }
  T* data() { return p_; }
  const T* data() const { return p_; }
  size_t size() const { return n_; }
  T& operator[](size_t i) { return p_[i]; }
  const T& operator[](size_t i) const { return p_[i]; }
  T* begin() { return p_; }
  T* end() { return p_ + n_; }
  const T* begin() const { return p_; }
  const T* end() const { return p_ + n_; }

 private:
  void release() {
    if (free_) {
      if (p_) free_(p_);
    } else {
      std::free(p_);
    }
    p_ = nullptr;
    n_ = 0;
    free_ = nullptr;
  }
  T* p_ = nullptr;
  size_t n_ = 0;
  FreeFn free_ = nullptr;
};

// Degree-ordered, upper-triangular CSR with sorted rows: the content of a
// .bbtc file once loaded. Offsets are held as 64-bit so graphs with more than
// 2^32 edges need no special path; column ids are 32-bit like the file's L=4.
struct Csr {
  uint32_t n = 0;      // rows (vertices)
  uint32_t ncols = 0;  // column count field of the file (== n for files we write)
  uint64_t m = 0;      // stored entries (undirected edges)
  Buf<uint64_t> off;   // n + 1
  Buf<uint32_t> col;   // m
};

inline double now_s() {
  using clk = std::chrono::steady_clock;
  return std::chrono::duration<double>(clk::now().time_since_epoch()).count();
}

}  // namespace bbtc
