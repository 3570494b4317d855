// CPU test of the streamed batch's host side (wgpu-sigops_amd/csrc/stream_host.hpp), built with
// AddressSanitizer and UBSan by tests/test_stream_host.py: random mixed batches (and all-Ed25519
// ones) are packed chunk by chunk; a stand-in for the kernel checks every tile against the caller's
// arrays (bucket, items, messages at their offsets) and writes results derived from the inputs;
// unpacking must return every item's result at its caller index. Prints "ok".
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "stream_host.hpp"

using namespace sigops_host;

static int fails = 0;
#define CHECK(c)                                                  \
  do {                                                            \
    if (!(c)) {                                                   \
      fprintf(stderr, "check failed at line %d: %s\n", __LINE__, #c); \
      if (++fails > 10) exit(1);                                  \
    }                                                             \
  } while (0)

static void run(size_t n, size_t chunk, bool all_ed, unsigned seed) {
  std::mt19937_64 rng(seed);
  std::vector<uint8_t> sch(n), sigs(n * 64), aux(n * 32);
  std::vector<uint64_t> off(n + 1, 0);
  for (size_t i = 0; i < n; i++) {
    sch[i] = all_ed ? 2 : (uint8_t)(rng() % 3);
    const uint64_t len = sch[i] == 2 ? (rng() % 4 == 0 ? rng() % 6000 : rng() % 300) : 0;
    off[i + 1] = off[i] + len;
  }
  std::vector<uint8_t> msgs(off[n] + 1);
  for (auto& b : msgs) b = (uint8_t)rng();
  for (auto& b : sigs) b = (uint8_t)rng();
  for (auto& b : aux) b = (uint8_t)rng();
  uint32_t order[ST_BUCKETS];
  order[0] = ST_R1;
  order[1] = ST_K1;
  for (uint32_t b = 0; b < ST_ED_BINS; b++) order[2 + b] = ST_ED0 + ST_ED_BINS - 1 - b;
  StreamGeom g = stream_geometry(n, chunk);
  // the processing order, in 3 parts
  std::vector<uint8_t> bkt(n);
  std::vector<uint32_t> sorted(n);
  stream_sort([](auto fn) { for (unsigned p = 0; p < 3; p++) fn(p); }, 3, n, all_ed ? nullptr : sch.data(),
              off.data(), order, bkt.data(), sorted.data());
  for (size_t k = g.nchunks; k-- > 0;) stream_message_bytes(g, k, sorted.data(), off.data());
  stream_message_layout(g);
  std::vector<StreamDesc> desc(g.slots);
  std::vector<uint8_t> in(g.slots * ST_IN_CHUNKS * ST_LANES * 16), outb(g.slots * ST_OUT_CHUNKS * ST_LANES * 16);
  std::vector<uint8_t> st(g.slots * ST_LANES), hm(g.msg_base[g.nchunks] + 16);
  std::vector<uint32_t> mrel(g.slots * ST_LANES), mlen(g.slots * ST_LANES), perm(g.slots * ST_LANES);
  StreamBufs B{desc.data(), in.data(), mrel.data(), mlen.data(), hm.data(), outb.data(), st.data(), perm.data()};
  std::vector<uint32_t> tiles(g.nchunks);
  {
    std::vector<int> once(n, 0);
    uint32_t last_rank = 0;
    for (size_t j = 0; j < n; j++) {
      const size_t i = sorted[j];
      once[i]++;
      CHECK(bkt[i] == stream_bucket(all_ed ? 2 : sch[i], off[i + 1] - off[i]));
      uint32_t rank = 0;
      while (order[rank] != bkt[i]) rank++;
      CHECK(rank >= last_rank);
      if (j && rank == last_rank) CHECK(sorted[j - 1] < i);  // stable within a bucket
      last_rank = rank;
    }
    for (size_t i = 0; i < n; i++) CHECK(once[i] == 1);
  }
  // pack the chunks in a scrambled order (threads pack them in any order)
  std::vector<size_t> ks(g.nchunks);
  for (size_t k = 0; k < g.nchunks; k++) ks[k] = k;
  std::shuffle(ks.begin(), ks.end(), rng);
  for (size_t k : ks)
    tiles[k] = stream_pack_chunk(g, k, sorted.data(), bkt.data(), sigs.data(), aux.data(), msgs.data(), off.data(), B);
  // the kernel's view: every slot of a chunk, its items, their inputs and messages
  std::vector<int> seen(n, 0);
  for (size_t k = 0; k < g.nchunks; k++) {
    CHECK(tiles[k] <= g.tpc);
    uint32_t last_rank = 0;
    size_t pos = k * g.chunk;  // the chunk's items, in slot and lane order, are its range of the sorted order
    for (size_t j = 0; j < g.tpc; j++) {
      const size_t t = k * g.tpc + j;
      const uint32_t c = desc[t].info & 0xFF, b = desc[t].info >> 8;
      CHECK((j < tiles[k]) == (c != 0));
      if (!c) continue;
      CHECK(c <= 64 && b < ST_BUCKETS);
      uint32_t rank = 0;
      while (order[rank] != b) rank++;
      CHECK(rank >= last_rank);  // tiles of a chunk in processing order
      last_rank = rank;
      CHECK(desc[t].m0 % 128 == 0 && desc[t].m0 + desc[t].mbytes + 4 <= g.msg_base[k + 1]);
      CHECK(desc[t].m0 >= g.msg_base[k]);
      for (uint32_t l = 0; l < c; l++) {
        const size_t i = perm[t * 64 + l];
        CHECK(i == sorted[pos++]);
        seen[i]++;
        CHECK(stream_bucket(sch[i], off[i + 1] - off[i]) == b);
        CHECK(st[t * 64 + l] == ST_CANARY);
        for (int q = 0; q < 4; q++) CHECK(!memcmp(&in[((t * 6 + q) * 64 + l) * 16], &sigs[i * 64 + q * 16], 16));
        for (int q = 0; q < 2; q++) CHECK(!memcmp(&in[((t * 6 + 4 + q) * 64 + l) * 16], &aux[i * 32 + q * 16], 16));
        CHECK(mlen[t * 64 + l] == off[i + 1] - off[i]);
        CHECK(mrel[t * 64 + l] + mlen[t * 64 + l] <= desc[t].mbytes);
        CHECK(!memcmp(&hm[desc[t].m0 + mrel[t * 64 + l]], &msgs[off[i]], mlen[t * 64 + l]));
        // stand-in results: out = the signature, status = low byte of the item index
        for (int q = 0; q < 4; q++)
          memcpy(&outb[((t * 4 + q) * 64 + l) * 16], &in[((t * 6 + q) * 64 + l) * 16], 16);
        st[t * 64 + l] = (uint8_t)i;
      }
    }
  }
  for (size_t i = 0; i < n; i++) CHECK(seen[i] == 1);
  std::vector<uint8_t> out(n * 64, 0xAA), status(n, 0xAA);
  for (size_t k = 0; k < g.nchunks; k++)
    stream_unpack_tiles(k * g.tpc, k * g.tpc + tiles[k], B, out.data(), status.data());
  for (size_t i = 0; i < n; i++) {
    CHECK(status[i] == (uint8_t)i);
    if (sch[i] == 2) {
      for (int q = 0; q < 64; q++) CHECK(out[i * 64 + q] == 0);
    } else {
      CHECK(!memcmp(&out[i * 64], &sigs[i * 64], 64));
    }
  }
  // one lane again (the careful pass's fix-up)
  if (n) {
    const size_t t = 0, l = 0;
    const size_t i = perm[t * 64 + l];
    memset(&outb[((t * 4 + 0) * 64 + l) * 16], 0x5C, 16);
    st[t * 64 + l] = 0x44;
    stream_unpack_lane(t * 64 + l, B, out.data(), status.data());
    CHECK(status[i] == 0x44 && out[i * 64] == 0x5C);
  }
}

int main() {
  run(1, 64, false, 1);
  run(63, 64, false, 2);
  run(1000, 256, false, 3);
  run(5000, 1024, true, 4);
  run(40000, 4096, false, 5);
  run(70000, 16384, false, 6);
  if (fails) return 1;
  printf("ok\n");
  return 0;
}
