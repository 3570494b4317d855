// sha512.hpp -- SHA-512 (FIPS 180-4) of R || A || M for Ed25519, one message per lane.
//
// The reference hashes exactly 96 bytes in one block (src/wgsl/sha512.wgsl:114-194, padding
// hard-coded at :122-123) because its messages are fixed 32-byte digests; here M has any length,
// so the padded preimage is generated word by word: 64 header bytes (R || A) from registers, the
// message from memory, then 0x80, zeros and the 128-bit bit length. Lanes of a wavefront with
// fewer blocks would go idle while the longest message finishes, which is why the kernels hash
// Ed25519 items grouped by block count.
#pragma once
#include "field.hpp"

namespace sigops {

DEV u64 rotr64(u64 x, int n) { return (x >> n) | (x << (64 - n)); }
DEV u64 bswap64(u64 x) { return __builtin_bswap64(x); }

// 16 rounds per loop trip with the 16-word schedule kept as a circular buffer in registers; the
// round constants are read with uniform (scalar) loads. Unrolling all 80 rounds lets the compiler
// hoist the whole message schedule and doubles the register footprint of the verifier.
DEV void sha512_round(u64 (&s)[8], int j, u64 kw) {
  const u64 a = s[(8 - j) & 7], b = s[(9 - j) & 7], c = s[(10 - j) & 7];
  u64& d = s[(11 - j) & 7];
  const u64 e = s[(12 - j) & 7], f = s[(13 - j) & 7], g = s[(14 - j) & 7];
  u64& hh = s[(15 - j) & 7];
  const u64 S1 = rotr64(e, 14) ^ rotr64(e, 18) ^ rotr64(e, 41);
  const u64 ch = (e & f) ^ (~e & g);
  const u64 t1 = hh + S1 + ch + kw;
  const u64 S0 = rotr64(a, 28) ^ rotr64(a, 34) ^ rotr64(a, 39);
  const u64 maj = (a & b) | (c & (a | b));
  d += t1;
  hh = t1 + S0 + maj;  // becomes the new "a": the register names rotate by one per round
}

DEV void sha512_compress(u64 h[8], u64 w[16]) {
  u64 s[8];
#pragma unroll
  for (int i = 0; i < 8; i++) s[i] = h[i];
#pragma unroll
  for (int j = 0; j < 16; j++) sha512_round(s, j & 7, SHA512_K[j] + w[j]);
#pragma unroll 1
  for (int t0 = 16; t0 < 80; t0 += 16) {
#pragma unroll
    for (int j = 0; j < 16; j++) {
      const u64 w15 = w[(j + 1) & 15], w2 = w[(j + 14) & 15];
      const u64 s0 = rotr64(w15, 1) ^ rotr64(w15, 8) ^ (w15 >> 7);
      const u64 s1 = rotr64(w2, 19) ^ rotr64(w2, 61) ^ (w2 >> 6);
      w[j] += s0 + w[(j + 9) & 15] + s1;
      sha512_round(s, j & 7, SHA512_K[t0 + j] + w[j]);
    }
  }
#pragma unroll
  for (int i = 0; i < 8; i++) h[i] += s[i];
}

// the digest as 16 little-endian u32 words (its 64 bytes in order)
DEV void sha512_out(const u64 h[8], u32 out[16]) {
#pragma unroll
  for (int i = 0; i < 8; i++) {
    const u64 le = bswap64(h[i]);
    out[2 * i] = (u32)le;
    out[2 * i + 1] = (u32)(le >> 32);
  }
}

// blocks of the padded preimage of total bytes: the bytes, 0x80, zeros, the 128-bit bit length
DEV u32 sha512_nblocks(u64 total) { return (u32)((total + 1 + 16 + 127) >> 7); }

// Schedule words 0..15 of block b (of nblk) of the padded preimage hdr[0..HL) || msg[0..len),
// hdr given as 16 raw u32 words (only the first HL/4 are used), msg at byte offset msg_off of a
// 4-byte-aligned buffer. The block's message bytes come from one window of 33 aligned words,
// assembled by funnel shifts (the shift is the same for every block: 128 and HL are multiples of
// 4). Window loads are clamped to the words that hold message bytes (the word at msg_off for an
// empty message), so no load branches or leaves the message; bytes past its end are masked off.
template <int HL, class Mem>
DEV void sha512_block_words(const u32 hdr[16], const Mem& mem, u64 msg_off, u32 len, u32 b, u32 nblk, u64 w[16]) {
  const u64 total = (u64)HL + len;
  // byte offset of preimage byte b * 128 in the buffer (below msg_off in block 0)
  const i64 S = (i64)msg_off + (i64)b * 128 - HL;
  const i64 wa = S >> 2;  // floor
  const u32 sh = (u32)(msg_off & 3u) * 8u;
  const i64 lo = (i64)(msg_off >> 2), hi = (i64)((msg_off + len - (len != 0)) >> 2);
  u32 raw[33];
#pragma unroll
  for (int k = 0; k < 33; k++) {
    const i64 i = wa + k;
    raw[k] = mem.word((u64)(i < lo ? lo : (i > hi ? hi : i)));
  }
#pragma unroll
  for (int q = 0; q < 16; q++) {
    const u64 o = (u64)b * 128 + 8 * q;  // byte offset in the preimage
    u64 word;
    if (o < (u64)HL) {
      const int hq = (int)(o >> 3) & 7;
      word = bswap64((u64)hdr[2 * hq] | ((u64)hdr[2 * hq + 1] << 32));
    } else {
      const u64 mo = o - HL;
      u64 le = (u64)funnel32(raw[2 * q + 1], raw[2 * q], sh) |
               ((u64)funnel32(raw[2 * q + 2], raw[2 * q + 1], sh) << 32);
      const i64 keep = (i64)len - (i64)mo;  // message bytes in this word
      if (keep < 8) {
        const u64 m = keep <= 0 ? 0ull : ((1ull << (8 * keep)) - 1ull);
        le &= m;
        if (keep >= 0) le |= 0x80ull << (8 * keep);
      }
      word = bswap64(le);
      if (b == nblk - 1 && q == 15) word = total * 8ull;
    }
    w[q] = word;
  }
}

// SHA-512(hdr[0..HL) || msg[0..len)) for a header of HL = 32 or 64 bytes, given as 16 raw u32
// words (only the first HL/4 are used). Output: the 64 digest bytes as 16 little-endian u32 words.
template <int HL, class Mem>
DEV void sha512_hdr(const u32 hdr[16], const Mem& mem, u64 msg_off, u32 len, u32 out[16]) {
  static_assert(HL == 32 || HL == 64, "header is 32 or 64 bytes");
  u64 h[8];
#pragma unroll
  for (int i = 0; i < 8; i++) h[i] = SHA512_IV[i];
  const u32 nblk = sha512_nblocks((u64)HL + len);
  for (u32 b = 0; b < nblk; b++) {
    u64 w[16];
    sha512_block_words<HL>(hdr, mem, msg_off, len, b, nblk, w);
    sha512_compress(h, w);
  }
  sha512_out(h, out);
}

// SHA-512(R || A || M): hdr = R then A (16 raw words)
template <class Mem>
DEV void sha512_ed(const u32 hdr[16], const Mem& mem, u64 msg_off, u32 len, u32 out[16]) {
  sha512_hdr<64>(hdr, mem, msg_off, len, out);
}

}  // namespace sigops
