// sha256.hpp -- SHA-256 (FIPS 180-4), one message per lane: Fuel's Message::new and addresses.
//
// Fuel hashes the data a signature covers with SHA-256 (`Message::new(data)`, the step before the
// recovery path: src/tests/ed25519_eddsa.rs:19, src/benchmarks/secp256k1_ecdsa.rs:162) and derives
// an address as SHA-256 of the 64-byte public key X || Y (the step after it). Both run here on
// the device, so data stays resident from raw bytes to recovered addresses.
// The padded message is generated word by word (message words from memory at any byte offset,
// then 0x80, zeros and the 64-bit bit length), 16 rounds per loop trip over a circular schedule.
#pragma once
#include "field.hpp"

namespace sigops {

DEV u32 rotr32(u32 x, int n) { return (x >> n) | (x << (32 - n)); }

DEV void sha256_round(u32 (&s)[8], int j, u32 kw) {
  const u32 a = s[(8 - j) & 7], b = s[(9 - j) & 7], c = s[(10 - j) & 7];
  u32& d = s[(11 - j) & 7];
  const u32 e = s[(12 - j) & 7], f = s[(13 - j) & 7], g = s[(14 - j) & 7];
  u32& hh = s[(15 - j) & 7];
  const u32 S1 = rotr32(e, 6) ^ rotr32(e, 11) ^ rotr32(e, 25);
  const u32 ch = (e & f) ^ (~e & g);
  const u32 t1 = hh + S1 + ch + kw;
  const u32 S0 = rotr32(a, 2) ^ rotr32(a, 13) ^ rotr32(a, 22);
  const u32 maj = (a & b) | (c & (a | b));
  d += t1;
  hh = t1 + S0 + maj;  // becomes the new "a": the register names rotate by one per round
}

DEV void sha256_compress(u32 h[8], u32 w[16]) {
  u32 s[8];
#pragma unroll
  for (int i = 0; i < 8; i++) s[i] = h[i];
#pragma unroll
  for (int j = 0; j < 16; j++) sha256_round(s, j & 7, SHA256_K[j] + w[j]);
#pragma unroll 1
  for (int t0 = 16; t0 < 64; t0 += 16) {
#pragma unroll
    for (int j = 0; j < 16; j++) {
      const u32 w15 = w[(j + 1) & 15], w2 = w[(j + 14) & 15];
      const u32 s0 = rotr32(w15, 7) ^ rotr32(w15, 18) ^ (w15 >> 3);
      const u32 s1 = rotr32(w2, 17) ^ rotr32(w2, 19) ^ (w2 >> 10);
      w[j] += s0 + w[(j + 9) & 15] + s1;
      sha256_round(s, j & 7, SHA256_K[t0 + j] + w[j]);
    }
  }
#pragma unroll
  for (int i = 0; i < 8; i++) h[i] += s[i];
}

// 4 little-endian bytes at byte offset off of a 4-byte-aligned buffer (reads the aligned word
// below off and the next one; the buffer carries slack after its last message)
template <class Mem>
DEV u32 load4_le(const Mem& mem, u64 off) {
  const u64 wa = off >> 2;
  const u32 sh = (u32)(off & 3u) * 8u;
  const u32 d0 = mem.word(wa);
  return sh ? ((d0 >> sh) | (mem.word(wa + 1) << (32 - sh))) : d0;
}

// SHA-256(msg[0..len)), msg at byte offset off; out: the 32 digest bytes as 8 little-endian words
template <class Mem>
DEV void sha256_msg(const Mem& mem, u64 off, u32 len, u32 out[8]) {
  u32 h[8];
#pragma unroll
  for (int i = 0; i < 8; i++) h[i] = SHA256_IV[i];
  const u64 bits = (u64)len * 8u;
  const u32 nblk = (u32)(((u64)len + 1 + 8 + 63) >> 6);
  for (u32 b = 0; b < nblk; b++) {
    u32 w[16];
#pragma unroll
    for (int q = 0; q < 16; q++) {
      const u64 o = (u64)b * 64 + 4 * q;  // byte offset in the padded message
      u32 le = 0;
      if (o < len) le = load4_le(mem, off + o);
      const i64 keep = (i64)len - (i64)o;  // message bytes in this word
      if (keep < 4) {
        const u32 m = keep <= 0 ? 0u : ((1u << (8 * keep)) - 1u);
        le &= m;
        if (keep >= 0) le |= 0x80u << (8 * keep);
      }
      u32 word = bswap32(le);
      if (b == nblk - 1 && q == 14) word = (u32)(bits >> 32);
      if (b == nblk - 1 && q == 15) word = (u32)bits;
      w[q] = word;
    }
    sha256_compress(h, w);
  }
#pragma unroll
  for (int i = 0; i < 8; i++) out[i] = bswap32(h[i]);
}

// SHA-256 of 64 bytes held in registers as 16 little-endian words (a public key X || Y as stored):
// one data block and one constant padding block (0x80, zeros, bit length 512)
DEV void sha256_64(const u32 in[16], u32 out[8]) {
  u32 h[8], w[16];
#pragma unroll
  for (int i = 0; i < 8; i++) h[i] = SHA256_IV[i];
#pragma unroll
  for (int q = 0; q < 16; q++) w[q] = bswap32(in[q]);
  sha256_compress(h, w);
#pragma unroll
  for (int q = 0; q < 16; q++) w[q] = q == 0 ? 0x80000000u : (q == 15 ? 512u : 0u);
  sha256_compress(h, w);
#pragma unroll
  for (int i = 0; i < 8; i++) out[i] = bswap32(h[i]);
}

}  // namespace sigops
