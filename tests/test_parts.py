"""The first-phase item parts of a launch (device/descriptors.hpp): every
item of a launch belongs to exactly one position of one part, and a claim of
kItemBatch positions never leaves its chunk, for one chunk per part and for
fixed chunks. Checked by compiling the header's
constexpr mapping on the host."""
import os
import subprocess

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))

CHECK = r"""
#include <cstdio>
#include <vector>
#include "device/descriptors.hpp"
using namespace bbtc;
int main() {
  const unsigned long long lens[] = {0, 1, 7, 8, 9, 63, 64, 65, 4095, 4096, 4097, 8191, 32767, 32768, 32769,
                                     100000, 3214342};
  for (unsigned long long len : lens)
    for (unsigned long long chunk : {0ull, (unsigned long long)kListChunkItems, 8ull * kItemBatch})
    for (unsigned long long base : {0ull, 5ull, 123456ull}) {
      const unsigned long long C = part_chunk(len, chunk);
      const unsigned shift = part_shift(len, chunk);
      if (C % kItemBatch) return 4;
      std::vector<int> seen(len, 0);
      unsigned long long total = 0;
      for (unsigned q = 0; q < unsigned(kQueues); ++q) {
        const unsigned long long n = part_positions(len, q, C);
        total += n;
        for (unsigned long long x = 0; x < n; x += kItemBatch) {
          const unsigned long long take = (x + kItemBatch < n ? x + kItemBatch : n) - x;
          const unsigned long long b = part_item(base, q, x, C, shift);
          for (unsigned long long y = 0; y < take; ++y) {
            const unsigned long long it = part_item(base, q, x + y, C, shift);
            if (it != b + y || it < base || it >= base + len) return 1;
            seen[it - base]++;
          }
        }
      }
      if (total != len) return 2;
      for (unsigned long long i = 0; i < len; ++i)
        if (seen[i] != 1) return 3;
    }
  std::puts("ok");
  return 0;
}
"""


def test_parts_cover_every_item_once(tmp_path):
    src = tmp_path / "parts.cpp"
    src.write_text(CHECK)
    exe = tmp_path / "parts"
    inc = os.path.join(ROOT, "bbtc_amd", "csrc")
    r = subprocess.run(["g++", "-std=c++17", "-O2", "-I", inc, str(src), "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip() == "ok", (r.returncode, r.stdout)
