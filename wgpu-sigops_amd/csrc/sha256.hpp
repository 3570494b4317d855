// sha256.hpp -- SHA-256 (FIPS 180-4), one message per lane: Fuel's Message::new and addresses.
//
// Fuel hashes the data a signature covers with SHA-256 (`Message::new(data)`, the step before the
// recovery path: src/tests/ed25519_eddsa.rs:19, src/benchmarks/secp256k1_ecdsa.rs:162) and derives
// an address as SHA-256 of the 64-byte public key X || Y (the step after it). Both run here on
// the device, so data stays resident from raw bytes to recovered addresses.
// A padded block is generated word by word from memory at any byte offset (17 aligned word loads,
// or 5 aligned 16-byte loads from a 16-byte-aligned buffer, and funnel shifts; then 0x80, zeros
// and the 64-bit bit length); the 64 rounds run 16 per loop trip over a circular schedule. The
// 3-input bitwise functions (Ch, Maj and the xor of three rotations) are one v_bitop3_b32 each on
// gfx950.
#pragma once
#include "field.hpp"

namespace sigops {

DEV u32 rotr32(u32 x, int n) { return (x >> n) | (x << (32 - n)); }

// bitop3 immediates are truth tables over the inputs src0 = 0xF0, src1 = 0xCC, src2 = 0xAA
DEV u32 xor3_32(u32 a, u32 b, u32 c) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
#else
  return a ^ b ^ c;
#endif
}

DEV u32 sha_ch(u32 e, u32 f, u32 g) {  // f where e is set, else g
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_bitop3_b32(e, f, g, 0xCA);
#else
  return (e & f) ^ (~e & g);
#endif
}

DEV u32 sha_maj(u32 a, u32 b, u32 c) {  // bitwise majority
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_bitop3_b32(a, b, c, 0xE8);
#else
  return (a & b) | (c & (a | b));
#endif
}

DEV void sha256_round(u32 (&s)[8], int j, u32 kw) {
  const u32 a = s[(8 - j) & 7], b = s[(9 - j) & 7], c = s[(10 - j) & 7];
  u32& d = s[(11 - j) & 7];
  const u32 e = s[(12 - j) & 7], f = s[(13 - j) & 7], g = s[(14 - j) & 7];
  u32& hh = s[(15 - j) & 7];
  const u32 S1 = xor3_32(rotr32(e, 6), rotr32(e, 11), rotr32(e, 25));
  const u32 t1 = hh + S1 + sha_ch(e, f, g) + kw;
  const u32 S0 = xor3_32(rotr32(a, 2), rotr32(a, 13), rotr32(a, 22));
  d += t1;
  hh = t1 + S0 + sha_maj(a, b, c);  // becomes the new "a": the register names rotate by one per round
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
      const u32 s0 = xor3_32(rotr32(w15, 7), rotr32(w15, 18), w15 >> 3);
      const u32 s1 = xor3_32(rotr32(w2, 17), rotr32(w2, 19), w2 >> 10);
      w[j] += s0 + w[(j + 9) & 15] + s1;
      sha256_round(s, j & 7, SHA256_K[t0 + j] + w[j]);
    }
  }
#pragma unroll
  for (int i = 0; i < 8; i++) h[i] += s[i];
}

// a block whose schedule is a constant, given as K[t] + W[t] (the padding block of a message
// whose length is a multiple of 64 bytes): rounds only
DEV void sha256_compress_kw(u32 h[8], const u32 (&kw)[64]) {
  u32 s[8];
#pragma unroll
  for (int i = 0; i < 8; i++) s[i] = h[i];
#pragma unroll
  for (int t = 0; t < 64; t++) sha256_round(s, t & 7, kw[t]);
#pragma unroll
  for (int i = 0; i < 8; i++) h[i] += s[i];
}

// blocks of the padded message of len bytes: the bytes, 0x80, zeros to 56 mod 64, the bit length
DEV u32 sha256_nblocks(u32 len) { return (u32)(((u64)len + 1 + 8 + 63) >> 6); }

// Schedule words 0..15 of block b (of nblk) of the padded message of len bytes, from raw[0..16]:
// the 17 aligned little-endian words from the one holding the block's first byte, whose byte
// offset within that word gives sh = 8 * offset. Bytes past the message are masked off, whatever
// was loaded for them.
DEV void sha256_block_assemble(const u32 raw[17], u32 sh, u32 len, u32 b, u32 nblk, u32 w[16]) {
  const i64 rem64 = (i64)len - (i64)b * 64;  // message bytes from the block start
  const i32 rem = rem64 > 64 ? 64 : (rem64 < -1 ? -1 : (i32)rem64);
  const u64 bits = (u64)len * 8u;
#pragma unroll
  for (int q = 0; q < 16; q++) {
    u32 le = funnel32(raw[q + 1], raw[q], sh);  // bytes 4q .. 4q+3 of the block, little-endian
    const i32 keep = rem - 4 * q;               // message bytes in this word
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
}

// Block b of the message msg[0 .. len) at byte offset off of a 4-byte-aligned buffer, one word
// load per aligned word. Loads are clamped to the aligned word holding the message's last byte
// (the word at off for an empty message), so they never reach more than 3 bytes past the message
// and need no branch.
template <class Mem>
DEV void sha256_block_words(const Mem& mem, u64 off, u32 len, u32 b, u32 nblk, u32 w[16]) {
  const u64 base = off + (u64)b * 64;
  const u64 wa = base >> 2;
  const u64 lim = (off + len - (len != 0)) >> 2;
  u32 raw[17];
#pragma unroll
  for (int k = 0; k < 17; k++) raw[k] = mem.word(wa + k < lim ? wa + k : lim);
  sha256_block_assemble(raw, (u32)(base & 3u) * 8u, len, b, nblk, w);
}

// The same from a 16-byte-aligned buffer with five 16-byte loads per lane instead of 17 word loads
// (a third of the load instructions when each lane hashes its own message). Loads are clamped to
// the 16-byte chunk holding the message's last byte (the chunk at off for an empty message), so
// they reach at most 15 bytes past the message. mem.quad(c, r) reads the four words of chunk c.
// The block's words are placed with lane masks; a version with conditional selects measured
// twice as slow on messages of mixed lengths.
template <class Mem>
DEV void sha256_block_words_x4(const Mem& mem, u64 off, u32 len, u32 b, u32 nblk, u32 w[16]) {
  const u64 base = off + (u64)b * 64;
  const u64 c0 = base >> 4;
  const u64 cl = (off + len - (len != 0)) >> 4;
  u32 r[20];
#pragma unroll
  for (int k = 0; k < 5; k++) mem.quad(c0 + k < cl ? c0 + k : cl, r + 4 * k);
  // the block starts at word q0 = 0..3 of the first chunk: raw[j] = r[q0 + j], by lane masks
  const u32 q0 = (u32)(base >> 2) & 3u;
  const u32 m1 = 0u - (q0 & 1u), m2 = 0u - ((q0 >> 1) & 1u);
  u32 raw[17];
#pragma unroll
  for (int j = 0; j < 17; j++) {
    const u32 even = (r[j + 2] & m2) | (r[j] & ~m2);
    const u32 odd = (r[j + 3] & m2) | (r[j + 1] & ~m2);
    raw[j] = (odd & m1) | (even & ~m1);
  }
  sha256_block_assemble(raw, (u32)(base & 3u) * 8u, len, b, nblk, w);
}

// SHA-256(msg[0..len)), msg at byte offset off; out: the 32 digest bytes as 8 little-endian words
// (X4: the buffer is 16-byte aligned and read with 16-byte loads)
template <bool X4 = false, class Mem>
DEV void sha256_msg(const Mem& mem, u64 off, u32 len, u32 out[8]) {
  u32 h[8];
#pragma unroll
  for (int i = 0; i < 8; i++) h[i] = SHA256_IV[i];
  const u32 nblk = sha256_nblocks(len);
  for (u32 b = 0; b < nblk; b++) {
    u32 w[16];
    if constexpr (X4)
      sha256_block_words_x4(mem, off, len, b, nblk, w);
    else
      sha256_block_words(mem, off, len, b, nblk, w);
    sha256_compress(h, w);
  }
#pragma unroll
  for (int i = 0; i < 8; i++) out[i] = bswap32(h[i]);
}

// SHA-256 of 64 bytes held in registers as 16 little-endian words (a public key X || Y as stored):
// one data block, then the constant padding block (0x80, zeros, bit length 512)
DEV void sha256_64(const u32 in[16], u32 out[8]) {
  u32 h[8], w[16];
#pragma unroll
  for (int i = 0; i < 8; i++) h[i] = SHA256_IV[i];
#pragma unroll
  for (int q = 0; q < 16; q++) w[q] = bswap32(in[q]);
  sha256_compress(h, w);
  sha256_compress_kw(h, SHA256_PAD64_KW);
#pragma unroll
  for (int i = 0; i < 8; i++) out[i] = bswap32(h[i]);
}

}  // namespace sigops
