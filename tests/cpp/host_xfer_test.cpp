// Host-side check of wgpu-sigops_amd/csrc/host_xfer.hpp without a GPU: the fork-join pool runs
// every thread index once per call, for_range covers each index once, and packing a staging image
// in parts split across threads (as upload() does) reproduces every segment byte for byte, with
// rebased u64 offsets, and writes nothing past the image.
#include <atomic>
#include <cstdio>
#include <random>
#include <vector>

#include "host_xfer.hpp"

using namespace sigops_host;

static int fail(const char* what) {
  printf("FAIL %s\n", what);
  return 1;
}

int main() {
  Pool pool(7);
  for (int r = 0; r < 2000; r++) {
    std::atomic<int> cnt[7];
    for (auto& x : cnt) x = 0;
    pool.run([&](unsigned t) { cnt[t]++; });
    for (auto& x : cnt)
      if (x != 1) return fail("pool.run");
  }
  for (size_t n : {0ul, 1ul, 100ul, 65536ul, 1000003ul}) {
    std::vector<std::atomic<int>> hit(n);
    pool.for_range(n, [&](size_t a, size_t b) {
      for (size_t i = a; i < b; i++) hit[i]++;
    });
    for (size_t i = 0; i < n; i++)
      if (hit[i] != 1) return fail("for_range");
  }
  std::mt19937_64 g(1);
  for (int trial = 0; trial < 40; trial++) {
    const int ns = 1 + (int)(g() % 5);
    std::vector<std::vector<uint8_t>> src(ns);
    std::vector<Up> segs;
    for (int k = 0; k < ns; k++) {
      size_t bytes = (g() % 3 == 0) ? 0 : (size_t)(g() % (20u << 20));
      const bool offsets = k == ns - 1 && bytes;
      if (offsets) bytes = bytes / 8 * 8;
      src[k].resize(bytes);
      for (auto& b : src[k]) b = (uint8_t)g();
      segs.push_back({nullptr, src[k].data(), bytes, offsets ? 12345ull : 0ull});
    }
    std::vector<size_t> pos;
    const size_t total = layout(segs, pos);
    std::vector<uint8_t> stage(total + 64, 0xAB);
    const unsigned nt = pool.size();
    for (size_t lo = 0; lo < total; lo += XFER_PART) {
      const size_t hi = std::min(total, lo + XFER_PART), span = hi - lo;
      pool.run([&](unsigned t) {
        const size_t a = lo + (span * t / nt) / 64 * 64;
        const size_t b = t + 1 == nt ? hi : lo + (span * (t + 1) / nt) / 64 * 64;
        pack_range(segs, pos, stage.data(), a, b);
      });
    }
    for (int k = 0; k < ns; k++) {
      if (pos[k] % XFER_ALIGN) return fail("alignment");
      if (segs[k].rebase == 0) {
        if (segs[k].bytes && memcmp(stage.data() + pos[k], src[k].data(), segs[k].bytes)) return fail("pack");
      } else {
        for (size_t i = 0; i < segs[k].bytes / 8; i++) {
          uint64_t v, w;
          memcpy(&v, src[k].data() + 8 * i, 8);
          memcpy(&w, stage.data() + pos[k] + 8 * i, 8);
          if (w != v - 12345ull) return fail("rebase");
        }
      }
    }
    if (stage[total] != 0xAB) return fail("overrun");
  }
  printf("ok\n");
  return 0;
}
