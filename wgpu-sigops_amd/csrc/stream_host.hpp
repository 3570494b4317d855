// stream_host.hpp -- host side of the streamed batch: the layout a persistent kernel drains while
// the host is still filling it, and the packing of caller arrays into it, chunk by chunk.
//
// A mixed batch arrives in the caller's (pageable) memory. Instead of copying it all to HBM and
// then launching (transfers and compute one after the other), the host sorts it by bucket
// (secp256k1, secp256r1, Ed25519 by SHA-512 block count) and packs it into pinned, device-mapped
// memory in chunks of that order, publishing each chunk as it is packed; the kernel, already
// running, takes tiles of published chunks and reads them over PCIe, so the packing, the transfers
// and the compute overlap, and the device needs no bucketing pass. Layout, per tile slot t (a
// chunk owns `tpc` consecutive slots, the unused ones empty):
//   desc[t]                  {info = bucket << 8 | item count (0: empty), message bytes, message start, tag}
//   in[(t * 6 + k) * 64 + l] 16-byte chunk k of lane l's item: signature (k = 0..3), digest or key (4..5)
//   mrel / mlen[t * 64 + l]  the item's message: start relative to the tile's, and length
//   msgs                     the tiles' messages back to back, each tile's region 128-byte aligned
//   out[(t * 4 + k) * 64 + l], st[t * 64 + l]: results, written by the kernel
// so one wavefront instruction reads or writes 64 consecutive 16-byte chunks (1 KB) of a tile.
// perm[t * 64 + l] is the caller's index of the item (host only), for unpacking the results.
// Host code only; no GPU code here.
#pragma once
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <vector>

namespace sigops_host {

constexpr uint32_t ST_LANES = 64;
constexpr uint32_t ST_IN_CHUNKS = 6;
constexpr uint32_t ST_OUT_CHUNKS = 4;
constexpr uint32_t ST_K1 = 0, ST_R1 = 1, ST_ED0 = 2;
constexpr uint32_t ST_ED_BINS = 40;
constexpr uint32_t ST_BUCKETS = 2 + ST_ED_BINS;
constexpr uint8_t ST_CANARY = 0xEE;

struct StreamDesc {
  uint32_t info;    // bucket << 8 | count
  uint32_t mbytes;  // message bytes of the tile (its region in msgs, before alignment padding)
  uint32_t m0;      // start of the tile's region in msgs (128-byte aligned)
  uint32_t tag;     // the kernel sets the position's done word to seq << 32 | tag
};
static_assert(sizeof(StreamDesc) == 16, "one 16-byte load per descriptor");

struct StreamGeom {
  size_t n = 0, chunk = 0, nchunks = 0, tpc = 0, slots = 0;
  // message bytes reserved per chunk: its bytes plus up to 15 bytes of padding per tile slot,
  // rounded up to 128 bytes
  std::vector<uint64_t> msg_base;  // nchunks + 1
};

// item bucket: 0 secp256k1, 1 secp256r1, 2 + (SHA-512 blocks of R || A || M) - 1 for Ed25519 (40 = "40 or
// more"), as the device bucketing computes it
inline uint32_t stream_bucket(uint8_t scheme, uint64_t len) {
  if (scheme == 0) return ST_K1;
  if (scheme == 1) return ST_R1;
  uint64_t nb = (64 + len + 17 + 127) >> 7;
  if (nb > ST_ED_BINS) nb = ST_ED_BINS;
  return ST_ED0 + (uint32_t)nb - 1;
}

// Chunks of `chunk` items (a multiple of 64) of the sorted order hold at most chunk / 64 full tiles
// plus one partial tile per bucket. Every region a chunk owns starts on a 128-byte boundary (tiles
// per chunk a multiple of 8, message regions rounded up to 128 bytes), so no cache line holds
// bytes of two chunks. The message regions need the sorted order: stream_message_layout.
inline StreamGeom stream_geometry(size_t n, size_t chunk) {
  StreamGeom g;
  g.n = n;
  g.chunk = chunk;
  g.nchunks = (n + chunk - 1) / chunk;
  g.tpc = (chunk / ST_LANES + ST_BUCKETS + 7) & ~(size_t)7;
  g.slots = g.nchunks * g.tpc;
  g.msg_base.assign(g.nchunks + 1, 0);
  return g;
}

// Message bytes of chunk k (its items' lengths in the sorted order, plus up to 131 bytes of padding
// per tile slot, rounded up to 128 bytes), then the regions' starts: call
// stream_message_bytes for every chunk (in any order or in parallel), then stream_message_layout.
inline void stream_message_bytes(StreamGeom& g, size_t k, const uint32_t* sorted, const uint64_t* msg_off) {
  uint64_t bytes = 0;
  for (size_t j = k * g.chunk; j < std::min(g.n, (k + 1) * g.chunk); j++)
    bytes += msg_off[sorted[j] + 1] - msg_off[sorted[j]];
  g.msg_base[k + 1] = (bytes + 132 * g.tpc + 127) & ~(uint64_t)127;
}

inline void stream_message_layout(StreamGeom& g) {
  for (size_t k = 0; k < g.nchunks; k++) g.msg_base[k + 1] += g.msg_base[k];
}

// pinned buffers of a streamed batch (host addresses; the device reaches them through the mapping)
struct StreamBufs {
  StreamDesc* desc;
  uint8_t* in;  // 16-byte chunks
  uint32_t* mrel;
  uint32_t* mlen;
  uint8_t* msgs;
  uint8_t* out;  // 16-byte chunks
  uint8_t* st;
  uint32_t* perm;
};

// The batch in processing order: bkt[i] = bucket of caller item i, sorted[j] = caller index of the
// j-th item, the buckets in the order `order` (processing rank -> bucket), stable within a bucket.
// A counting sort in `parts` ranges of the items, run by par(fn), which calls fn(p) for every p in
// [0, parts) (in parallel or not): chunks then cut the sorted order, so every bucket sits in one
// run of consecutive chunks, as in a whole-batch bucketing (few partial tiles, and the waves that
// run at the same time run the same scheme's code). scheme may be NULL (every item Ed25519).
template <class Par>
void stream_sort(Par par, unsigned parts, size_t n, const uint8_t* scheme, const uint64_t* msg_off,
                 const uint32_t* order, uint8_t* bkt, uint32_t* sorted) {
  std::vector<uint32_t> hist((size_t)parts * ST_BUCKETS, 0);
  par([&](unsigned p) {
    uint32_t* h = &hist[(size_t)p * ST_BUCKETS];
    for (size_t i = n * p / parts; i < n * (p + 1) / parts; i++) {
      bkt[i] = (uint8_t)stream_bucket(scheme ? scheme[i] : 2, msg_off[i + 1] - msg_off[i]);
      h[bkt[i]]++;
    }
  });
  // where part p's items of bucket q start: buckets in processing order, parts in index order
  std::vector<uint32_t> base((size_t)parts * ST_BUCKETS);
  uint32_t at = 0;
  for (uint32_t r = 0; r < ST_BUCKETS; r++) {
    const uint32_t q = order[r];
    for (unsigned p = 0; p < parts; p++) {
      base[(size_t)p * ST_BUCKETS + q] = at;
      at += hist[(size_t)p * ST_BUCKETS + q];
    }
  }
  par([&](unsigned p) {
    uint32_t* b = &base[(size_t)p * ST_BUCKETS];
    for (size_t i = n * p / parts; i < n * (p + 1) / parts; i++) sorted[b[bkt[i]]++] = (uint32_t)i;
  });
}

// One item's inputs as the packing reads them
struct StreamItem {
  const uint8_t* sig;  // 64 bytes
  const uint8_t* aux;  // 32 bytes: digest or public key
  const uint8_t* msg;  // len bytes (Ed25519)
  uint64_t len;
};

// Packs one tile at slot t: c (1..64) items of bucket q, lane l's from item(l); their messages go
// back to back from msgs + m0. Statuses of the packed lanes are set to ST_CANARY so a lane the
// kernel did not write shows. Returns the tile's message bytes.
template <class ItemOf>
inline uint64_t stream_pack_tile(StreamBufs& B, size_t t, uint32_t q, uint32_t c, uint64_t m0, uint32_t tag,
                                 ItemOf item) {
  uint8_t* in = B.in + t * (ST_IN_CHUNKS * ST_LANES * 16);
  uint64_t rel = 0;
  for (uint32_t l = 0; l < c; l++) {
    const StreamItem it = item(l);
    for (uint32_t cc = 0; cc < 4; cc++) memcpy(in + (cc * ST_LANES + l) * 16, it.sig + cc * 16, 16);
    for (uint32_t cc = 0; cc < 2; cc++) memcpy(in + ((4 + cc) * ST_LANES + l) * 16, it.aux + cc * 16, 16);
    B.st[t * ST_LANES + l] = ST_CANARY;
    B.mrel[t * ST_LANES + l] = (uint32_t)rel;
    B.mlen[t * ST_LANES + l] = (uint32_t)it.len;
    if (it.len) memcpy(B.msgs + m0 + rel, it.msg, it.len);
    rel += it.len;
  }
  B.desc[t] = StreamDesc{(q << 8) | c, (uint32_t)rel, (uint32_t)m0, tag};
  return rel;
}

// Bytes a tile's message region takes: 128-byte aligned, with room for the hash's reads past the
// last message byte (at most one word), so no cache line and no read spans two tiles
inline uint64_t stream_region_bytes(uint64_t msg_bytes) { return (msg_bytes + 4 + 127) & ~(uint64_t)127; }

// Packs chunk k (items sorted[k * chunk ..] of the processing order) into its slots; returns the
// number of tiles that hold items (the first ones of the chunk's slots), which the kernel marks
// done one by one. A tile holds up to 64 consecutive items of one bucket.
inline uint32_t stream_pack_chunk(const StreamGeom& g, size_t k, const uint32_t* sorted, const uint8_t* bkt,
                                  const uint8_t* sigs, const uint8_t* aux, const uint8_t* msgs,
                                  const uint64_t* msg_off, StreamBufs& B) {
  const size_t b = k * g.chunk, e = std::min(g.n, b + g.chunk);
  const size_t t0 = k * g.tpc;
  uint32_t tiles = 0;
  uint64_t mcur = g.msg_base[k];
  for (size_t j = b; j < e;) {
    // a tile: up to 64 items of the bucket of item j
    const uint32_t q = bkt[sorted[j]];
    uint32_t c = 0;
    while (c < ST_LANES && j + c < e && bkt[sorted[j + c]] == q) c++;
    const size_t t = t0 + tiles++;
    for (uint32_t l = 0; l < c; l++) B.perm[t * ST_LANES + l] = sorted[j + l];
    const uint64_t rel = stream_pack_tile(B, t, q, c, mcur, (uint32_t)t, [&](uint32_t l) {
      const size_t i = sorted[j + l];
      return StreamItem{sigs + i * 64, aux + i * 32, msgs + msg_off[i], msg_off[i + 1] - msg_off[i]};
    });
    mcur += stream_region_bytes(rel);
    j += c;
  }
  for (size_t t = t0 + tiles; t < t0 + g.tpc; t++) B.desc[t] = StreamDesc{0, 0, 0, (uint32_t)t};
  return tiles;
}

// Copies the results of tile slots [t_begin, t_end) to the caller's arrays: out (n x 64; zeros for
// Ed25519 items) and status (n).
inline void stream_unpack_tiles(size_t t_begin, size_t t_end, const StreamBufs& B, uint8_t* out, uint8_t* status) {
  for (size_t t = t_begin; t < t_end; t++) {
    const uint32_t c = B.desc[t].info & 0xFF;
    const bool ed = (B.desc[t].info >> 8) >= ST_ED0;
    const uint8_t* o = B.out + t * (ST_OUT_CHUNKS * ST_LANES * 16);
    for (uint32_t l = 0; l < c; l++) {
      const size_t i = B.perm[t * ST_LANES + l];
      if (ed) {
        memset(out + i * 64, 0, 64);
      } else {
        for (uint32_t k = 0; k < ST_OUT_CHUNKS; k++) memcpy(out + i * 64 + k * 16, o + (k * ST_LANES + l) * 16, 16);
      }
      status[i] = B.st[t * ST_LANES + l];
    }
  }
}

// Copies one lane's result (after the careful pass rewrote it)
inline void stream_unpack_lane(size_t slot_lane, const StreamBufs& B, uint8_t* out, uint8_t* status) {
  const size_t t = slot_lane / ST_LANES, l = slot_lane % ST_LANES;
  const size_t i = B.perm[slot_lane];
  const uint8_t* o = B.out + t * (ST_OUT_CHUNKS * ST_LANES * 16);
  for (uint32_t k = 0; k < ST_OUT_CHUNKS; k++) memcpy(out + i * 64 + k * 16, o + (k * ST_LANES + l) * 16, 16);
  status[i] = B.st[slot_lane];
}

}  // namespace sigops_host
