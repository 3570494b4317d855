// gen.hpp -- seeded synthetic workload generation on the device (signing), one item per lane.
//
// Benchmarks and serving experiments need millions of valid signatures; this signs them on the GPU
// with the engine's own field and curve code. Item i draws all of its randomness from splitmix64
// streams keyed by (seed, i, purpose tag), so the output is independent of the launch shape and is
// reproduced byte for byte by the CPU reference generator used in the tests.
//   secp: sk uniform in [1, n-1] (rejection), digest 32 uniform bytes, nonce uniform in [1, n-1],
//         R = kG, r = R.x (x >= n rejected: fuel encodes only the y parity), s = (z + r sk)/k,
//         low-s normalisation with the parity flipped, parity in bit 7 of byte 32 (fuel format).
//   ed25519: 32-byte seed -> RFC 8032 key and signature; message bytes from the item stream.
//   corruption: with probability ppm/1e6 one uniform bit of r/R, s or the digest/message flips.
#pragma once
#include "ed25519.hpp"

namespace sigops {

enum : u64 { TAG_SK = 1, TAG_DIGEST = 2, TAG_NONCE = 3, TAG_CORRUPT = 4, TAG_SEED = 5, TAG_MSG = 6 };

DEV u64 splitmix_next(u64& s) {
  u64 z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

DEV u64 stream_init(u64 seed, u64 i, u64 tag) {
  u64 s = seed ^ (i * 0x9E3779B97F4A7C15ull) ^ (tag * 0xD1B54A32D192ED03ull);
  (void)splitmix_next(s);
  return s;
}

// 32 stream bytes as 8 little-endian words (byte b of the stream = byte b of the result)
DEV void stream_words32(u64 seed, u64 i, u64 tag, u32 w[8]) {
  u64 s = stream_init(seed, i, tag);
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const u64 x = splitmix_next(s);
    w[2 * q] = (u32)x;
    w[2 * q + 1] = (u32)(x >> 32);
  }
}

// uniform scalar in [1, n-1]: the 32 stream bytes read big-endian, retried with tag + 16 * attempt
DEV void draw_scalar_be(u64 seed, u64 i, u64 tag, const u32 (&n)[8], u32 raw[8], u32 val[8]) {
  for (u64 attempt = 0;; attempt++) {
    stream_words32(seed, i, tag + 16 * attempt, raw);
    words_from_be_bytes(raw, val);
    if (!words_is_zero(val) && words_lt(val, n)) return;
  }
}

// fixed-base k*G (k in [1, n-1]) with the LDS/global G table, affine plain words out
template <class C, class GTab>
DEV void secp_mul_g(const u32 k[8], const GTab& gt, u32 x_out[8], u32 y_out[8]) {
  typedef typename C::Fp F;
  const u32(&n)[8] = C::n32();
  const bool neg = (k[0] & 1u) == 0;
  u32 kk[8], t[8], c[8];
  words_sub(n, k, t);
#pragma unroll
  for (int q = 0; q < 8; q++) kk[q] = neg ? t[q] : k[q];
#pragma unroll
  for (int q = 0; q < 8; q++) c[q] = (kk[q] >> 1) | (q < 7 ? kk[q + 1] << 31 : 0x80000000u);
  bool exc = false;
  fe gx, gy;
  jac acc;
  gt.load(byte8(c, 31) - 128, gx, gy);
  acc.X = gx;
  acc.Y = neg ? fneg<F>(gy) : gy;
  acc.Z = fe_from_const(F::ONE);
  for (int j = 30; j >= 0; j--) {
#pragma unroll 1
    for (int d = 0; d < 8; d++) acc = jdbl<C>(acc);
    const u32 b = byte8(c, j);
    const bool dn = (b < 128) != neg;
    gt.load(((b < 128) ? ~b : b) & 127u, gx, gy);
    acc = jmadd<C>(acc, gx, dn ? fneg<F>(gy) : gy, exc);
  }
  const fe zi = finv<F>(acc.Z);
  const fe zi2 = mont_sqr<F>(zi);
  fe_to_words(from_mont<F>(mont_mul<F>(acc.X, zi2)), x_out);
  fe_to_words(from_mont<F>(mont_mul<F>(acc.Y, mont_mul<F>(zi2, zi))), y_out);
}

DEV void flip_bit_words(u32* w, u32 bit) {  // bit index over the bytes of a little-endian word array
  w[bit >> 5] ^= 1u << (bit & 31);
}

// one secp item: sig (16 raw words), digest (8 raw words), pub (16 raw words X||Y BE), corrupted flag
template <class C, class GTab>
DEV void gen_secp_item(u64 seed, u64 i, u32 ppm, const GTab& gt, u32 sig[16], u32 dig[8], u32 pub[16],
                       u32& corrupted) {
  typedef typename C::Fn Fn;
  const u32(&n)[8] = C::n32();
  u32 skraw[8], sk[8];
  draw_scalar_be(seed, i, TAG_SK, n, skraw, sk);
  stream_words32(seed, i, TAG_DIGEST, dig);
  u32 px[8], py[8];
  secp_mul_g<C>(sk, gt, px, py);
  words_to_be_bytes(px, pub);
  words_to_be_bytes(py, pub + 8);
  u32 z[8];
  words_from_be_bytes(dig, z);
  if (!words_lt(z, n)) {
    u32 t[8];
    words_sub(z, n, t);
#pragma unroll
    for (int q = 0; q < 8; q++) z[q] = t[q];
  }
  for (u64 attempt = 0;; attempt++) {
    u32 kraw[8], k[8], rx[8], ry[8];
    draw_scalar_be(seed, i, TAG_NONCE + 256 * attempt, n, kraw, k);
    secp_mul_g<C>(k, gt, rx, ry);
    if (!words_lt(rx, n) || words_is_zero(rx)) continue;
    const fe km = to_mont<Fn>(fe_from_words(k));
    const fe rm = to_mont<Fn>(fe_from_words(rx));
    const fe dm = to_mont<Fn>(fe_from_words(sk));
    const fe zm = to_mont<Fn>(fe_from_words(z));
    const fe sm = mont_mul<Fn>(fadd<Fn>(mont_mul<Fn>(rm, dm), zm), finv<Fn>(km));
    u32 s[8];
    fe_to_words(from_mont<Fn>(sm), s);
    if (words_is_zero(s)) continue;
    u32 odd = ry[0] & 1u;
    // low-s: s > n/2  <=>  2s > n  (n odd)
    u32 half[8];
#pragma unroll
    for (int q = 0; q < 8; q++) half[q] = (n[q] >> 1) | (q < 7 ? n[q + 1] << 31 : 0u);
    if (words_lt(half, s)) {
      u32 t[8];
      words_sub(n, s, t);
#pragma unroll
      for (int q = 0; q < 8; q++) s[q] = t[q];
      odd ^= 1u;
    }
    words_to_be_bytes(rx, sig);
    words_to_be_bytes(s, sig + 8);
    sig[8] |= odd << 7;
    break;
  }
  u64 cs = stream_init(seed, i, TAG_CORRUPT);
  const u64 w = splitmix_next(cs);
  corrupted = (w % 1000000ull) < ppm ? 1u : 0u;
  if (corrupted) {
    const u64 w2 = splitmix_next(cs);
    const u32 field = (u32)(w2 % 3), bit = (u32)((w2 >> 8) % 256);
    u32* target = field == 0 ? sig : field == 1 ? sig + 8 : dig;
    flip_bit_words(target, bit);
  }
}

// ---- Ed25519 ----
template <class BTab>
DEV void ed_mul_b(const u32 k[8], const BTab& bt, u32 enc[8]) {  // compress(k * B), k < l
  const bool neg = (k[0] & 1u) == 0;
  u32 kk[8], t[8], c[8];
  words_sub(ED_L32, k, t);
#pragma unroll
  for (int q = 0; q < 8; q++) kk[q] = neg ? t[q] : k[q];
#pragma unroll
  for (int q = 0; q < 8; q++) c[q] = (kk[q] >> 1) | (q < 7 ? kk[q + 1] << 31 : 0x80000000u);
  ext acc = ext_identity();
  for (int j = 31; j >= 0; j--) {
    if (j < 31) {
#pragma unroll 1
      for (int d = 0; d < 7; d++) acc = edbl(acc, false);
      acc = edbl(acc, true);
    }
    const u32 b = byte8(c, j);
    const bool dn = (b < 128) != neg;
    fe ypx, ymx, t2d;
    bt.load(((b < 128) ? ~b : b) & 127u, ypx, ymx, t2d);
    if (dn) {
      const fe tmp = ypx;
      ypx = ymx;
      ymx = tmp;
      t2d = fneg<Fe>(t2d);
    }
    acc = eadd_niels(acc, ypx, ymx, t2d, true);
  }
  const fe zi = finv<Fe>(acc.Z);
  const fe xa = from_mont<Fe>(mont_mul<Fe>(acc.X, zi));
  fe_to_words(from_mont<Fe>(mont_mul<Fe>(acc.Y, zi)), enc);
  enc[7] ^= (xa.v[0] & 1u) << 31;
}

// 64 little-endian digest bytes (16 words) mod l -> 8 words
DEV void wide_mod_l(const u32 h[16], u32 k[8]) {
  const fe hi = mont_mul<Fl>(fe_from_words(h + 8), fe_from_const(L_C256));
  const fe lo = mont_mul<Fl>(fe_from_words(h), fe_from_const(L_C0));
  fe_to_words(fcanon<Fl>(fadd<Fl>(hi, lo)), k);
}

DEV void scalar_mod_l(const u32 a[8], u32 out[8]) {  // a < 2^256
  fe_to_words(fcanon<Fl>(mont_mul<Fl>(fe_from_words(a), fe_from_const(L_C0))), out);
}

// message bytes of item i (written by a kernel of their own, before signing reads them)
template <class PutByte>
DEV void gen_ed_message(u64 seed, u64 i, u64 off, u32 len, PutByte& put) {
  u64 s = stream_init(seed, i, TAG_MSG);
  for (u32 p = 0; p < len; p += 8) {
    const u64 x = splitmix_next(s);
    for (u32 j = 0; j < 8 && p + j < len; j++) put(off + p + j, (u32)(x >> (8 * j)) & 0xFFu);
  }
}

// keys and signature of item i over its message (already in mem); a message corruption is
// returned as (byte offset, xor mask) for the caller to apply after every lane has hashed
template <class BTab, class Mem>
DEV void gen_ed_item(u64 seed, u64 i, u32 ppm, u64 off, u32 len, const BTab& bt, const Mem& mem, u32 sig[16],
                     u32 pk[8], u32& corrupted, u64& msg_flip_off, u32& msg_flip_mask) {
  msg_flip_mask = 0;
  msg_flip_off = 0;
  u32 seedw[8];
  stream_words32(seed, i, TAG_SEED, seedw);
  // h = SHA-512(seed): the seed is the 32-byte "header", no message
  u32 h[16];
  {
    u32 hdr[16];
#pragma unroll
    for (int q = 0; q < 8; q++) { hdr[q] = seedw[q]; hdr[8 + q] = 0; }
    sha512_hdr<32>(hdr, mem, 0, 0, h);
  }
  u32 a[8];
#pragma unroll
  for (int q = 0; q < 8; q++) a[q] = h[q];
  a[0] &= ~7u;
  a[7] &= 0x7FFFFFFFu;
  a[7] |= 0x40000000u;
  u32 amod[8];
  scalar_mod_l(a, amod);
  ed_mul_b(amod, bt, pk);
  // r = SHA-512(prefix || M) mod l
  u32 hr[16], r[8];
  {
    u32 hdr[16];
#pragma unroll
    for (int q = 0; q < 8; q++) { hdr[q] = h[8 + q]; hdr[8 + q] = 0; }
    sha512_hdr<32>(hdr, mem, off, len, hr);
  }
  wide_mod_l(hr, r);
  u32 Renc[8];
  ed_mul_b(r, bt, Renc);
  // k = SHA-512(R || A || M) mod l
  u32 hk[16], k[8];
  {
    u32 hdr[16];
#pragma unroll
    for (int q = 0; q < 8; q++) { hdr[q] = Renc[q]; hdr[8 + q] = pk[q]; }
    sha512_hdr<64>(hdr, mem, off, len, hk);
  }
  wide_mod_l(hk, k);
  // S = r + k*a mod l
  const fe km = to_mont<Fl>(fe_from_words(k));
  const fe am = to_mont<Fl>(fe_from_words(amod));
  const fe rm = to_mont<Fl>(fe_from_words(r));
  u32 S[8];
  fe_to_words(from_mont<Fl>(fadd<Fl>(mont_mul<Fl>(km, am), rm)), S);
#pragma unroll
  for (int q = 0; q < 8; q++) { sig[q] = Renc[q]; sig[8 + q] = S[q]; }
  u64 cs = stream_init(seed, i, TAG_CORRUPT);
  const u64 w = splitmix_next(cs);
  corrupted = (w % 1000000ull) < ppm ? 1u : 0u;
  if (corrupted) {
    const u64 w2 = splitmix_next(cs);
    u32 field = (u32)(w2 % 3);
    if (field == 2 && len == 0) field = (u32)((w2 >> 40) % 2);
    if (field == 2) {
      const u64 bit = (w2 >> 8) % (8ull * len);
      msg_flip_off = off + (bit >> 3);
      msg_flip_mask = 1u << (bit & 7);
    } else {
      flip_bit_words(field == 0 ? sig : sig + 8, (u32)((w2 >> 8) % 256));
    }
  }
}

}  // namespace sigops
