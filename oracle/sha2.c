/*
 * sha2.c -- SHA-512 and SHA-256 (FIPS 180-4) for the oracle (test infrastructure only).
 *
 * SHA-512 replaces, as a checker, the reference's one-block sha512_96
 * (src/wgsl/sha512.wgsl:114-194; IV and round constants :4-56) and is general in the length.
 * SHA-256 is what fuel-crypto's Message::new applies to message bytes
 * (src/tests/ed25519_eddsa.rs:19, src/benchmarks/secp256k1_ecdsa.rs:162).
 * The round constants are computed here from their definition (fractional parts of cube
 * roots of the first primes) instead of being written out, and the tests check them against
 * hashlib on many lengths.
 */
#include <math.h>
#include <string.h>
#include "oracle_internal.h"

static uint64_t K512[80], H512[8];
static uint32_t K256[64], H256[8];
static int consts_ready = 0;

/* floor(frac(x) * 2^bits) for x = root(prime), computed exactly with integer Newton steps */
static uint64_t icbrt_frac64(uint64_t p) {
  /* find floor(cbrt(p) * 2^64) mod 2^64 via integer cube root of p * 2^192 */
  unsigned __int128 lo = 0, hi = (unsigned __int128)1 << 72; /* cbrt(p*2^192) < 2^(64+8) */
  /* binary search on y with y^3 <= p * 2^192 ; compare using 256-bit arithmetic in pieces */
  while (hi - lo > 1) {
    unsigned __int128 mid = lo + (hi - lo) / 2;
    /* compute mid^3 and compare to p << 192 without overflow: mid < 2^72 -> mid^3 < 2^216 */
    /* represent as base-2^64 limbs */
    uint64_t m0 = (uint64_t)mid, m1 = (uint64_t)(mid >> 64);
    /* sq = mid^2 (up to 144 bits) */
    unsigned __int128 a = (unsigned __int128)m0 * m0;
    unsigned __int128 b = (unsigned __int128)m0 * m1 * 2;
    unsigned __int128 cc = (unsigned __int128)m1 * m1;
    uint64_t s0 = (uint64_t)a;
    unsigned __int128 t = (a >> 64) + (uint64_t)b;
    uint64_t s1 = (uint64_t)t;
    t = (t >> 64) + (b >> 64) + cc;
    uint64_t s2 = (uint64_t)t;
    /* cube = sq * mid */
    uint64_t c[4] = {0, 0, 0, 0};
    uint64_t sq[3] = {s0, s1, s2}, mm[2] = {m0, m1};
    for (int i = 0; i < 3; i++) {
      unsigned __int128 carry = 0;
      for (int j = 0; j < 2 && i + j < 4; j++) {
        carry += (unsigned __int128)sq[i] * mm[j] + c[i + j];
        c[i + j] = (uint64_t)carry;
        carry >>= 64;
      }
      if (i + 2 < 4) c[i + 2] += (uint64_t)carry;
    }
    /* target = p << 192 -> limbs {0,0,0,p} */
    int le;
    if (c[3] != p) le = c[3] < p;
    else le = (c[2] | c[1] | c[0]) == 0;
    if (le) lo = mid; else hi = mid;
  }
  return (uint64_t)lo; /* low 64 bits = fractional part */
}
static uint64_t isqrt_frac64(uint64_t p) {
  unsigned __int128 lo = 0, hi = (unsigned __int128)1 << 70;
  while (hi - lo > 1) {
    unsigned __int128 mid = lo + (hi - lo) / 2;
    uint64_t m0 = (uint64_t)mid, m1 = (uint64_t)(mid >> 64);
    unsigned __int128 a = (unsigned __int128)m0 * m0;
    unsigned __int128 b = (unsigned __int128)m0 * m1 * 2;
    unsigned __int128 cc = (unsigned __int128)m1 * m1;
    unsigned __int128 t = (a >> 64) + (uint64_t)b;
    uint64_t s1 = (uint64_t)t;
    t = (t >> 64) + (b >> 64) + cc;
    uint64_t s2 = (uint64_t)t;
    /* compare {s2,s1,s0} to p << 128 */
    int le;
    if (s2 != p) le = s2 < p;
    else le = (s1 | (uint64_t)a) == 0;
    if (le) lo = mid; else hi = mid;
  }
  return (uint64_t)lo;
}

static void init_consts(void) {
  if (consts_ready) return;
  uint64_t primes[80];
  int cnt = 0;
  for (uint64_t x = 2; cnt < 80; x++) {
    int isp = 1;
    for (uint64_t d = 2; d * d <= x; d++)
      if (x % d == 0) { isp = 0; break; }
    if (isp) primes[cnt++] = x;
  }
  for (int i = 0; i < 80; i++) K512[i] = icbrt_frac64(primes[i]);
  for (int i = 0; i < 8; i++) H512[i] = isqrt_frac64(primes[i]);
  for (int i = 0; i < 64; i++) K256[i] = (uint32_t)(K512[i] >> 32);
  for (int i = 0; i < 8; i++) H256[i] = (uint32_t)(H512[i] >> 32);
  __atomic_store_n(&consts_ready, 1, __ATOMIC_RELEASE);
}

#define ROR64(x, n) (((x) >> (n)) | ((x) << (64 - (n))))
#define ROR32(x, n) (((x) >> (n)) | ((x) << (32 - (n))))

static void sha512_block(uint64_t h[8], const uint8_t* blk) {
  uint64_t w[80];
  for (int i = 0; i < 16; i++) {
    uint64_t v = 0;
    for (int j = 0; j < 8; j++) v = (v << 8) | blk[i * 8 + j];
    w[i] = v;
  }
  for (int i = 16; i < 80; i++) {
    uint64_t s0 = ROR64(w[i - 15], 1) ^ ROR64(w[i - 15], 8) ^ (w[i - 15] >> 7);
    uint64_t s1 = ROR64(w[i - 2], 19) ^ ROR64(w[i - 2], 61) ^ (w[i - 2] >> 6);
    w[i] = w[i - 16] + s0 + w[i - 7] + s1;
  }
  uint64_t a = h[0], b = h[1], c = h[2], d = h[3], e = h[4], f = h[5], g = h[6], hh = h[7];
  for (int i = 0; i < 80; i++) {
    uint64_t S1 = ROR64(e, 14) ^ ROR64(e, 18) ^ ROR64(e, 41);
    uint64_t ch = (e & f) ^ (~e & g);
    uint64_t t1 = hh + S1 + ch + K512[i] + w[i];
    uint64_t S0 = ROR64(a, 28) ^ ROR64(a, 34) ^ ROR64(a, 39);
    uint64_t maj = (a & b) ^ (a & c) ^ (b & c);
    uint64_t t2 = S0 + maj;
    hh = g; g = f; f = e; e = d + t1; d = c; c = b; b = a; a = t1 + t2;
  }
  h[0] += a; h[1] += b; h[2] += c; h[3] += d; h[4] += e; h[5] += f; h[6] += g; h[7] += hh;
}

void or_sha512(const uint8_t* msg, size_t len, uint8_t out[64]) {
  init_consts();
  uint64_t h[8];
  memcpy(h, H512, sizeof h);
  size_t off = 0;
  for (; off + 128 <= len; off += 128) sha512_block(h, msg + off);
  uint8_t blk[256];
  memset(blk, 0, sizeof blk);
  size_t rem = len - off;
  if (rem) memcpy(blk, msg + off, rem);
  blk[rem] = 0x80;
  size_t nb = (rem + 1 + 16 <= 128) ? 1 : 2;
  unsigned __int128 bits = (unsigned __int128)len * 8;
  for (int j = 0; j < 16; j++) blk[nb * 128 - 1 - j] = (uint8_t)(bits >> (8 * j));
  for (size_t b = 0; b < nb; b++) sha512_block(h, blk + b * 128);
  for (int i = 0; i < 8; i++)
    for (int j = 0; j < 8; j++) out[i * 8 + j] = (uint8_t)(h[i] >> (56 - 8 * j));
}

static void sha256_block(uint32_t h[8], const uint8_t* blk) {
  uint32_t w[64];
  for (int i = 0; i < 16; i++)
    w[i] = ((uint32_t)blk[4 * i] << 24) | ((uint32_t)blk[4 * i + 1] << 16) |
           ((uint32_t)blk[4 * i + 2] << 8) | blk[4 * i + 3];
  for (int i = 16; i < 64; i++) {
    uint32_t s0 = ROR32(w[i - 15], 7) ^ ROR32(w[i - 15], 18) ^ (w[i - 15] >> 3);
    uint32_t s1 = ROR32(w[i - 2], 17) ^ ROR32(w[i - 2], 19) ^ (w[i - 2] >> 10);
    w[i] = w[i - 16] + s0 + w[i - 7] + s1;
  }
  uint32_t a = h[0], b = h[1], c = h[2], d = h[3], e = h[4], f = h[5], g = h[6], hh = h[7];
  for (int i = 0; i < 64; i++) {
    uint32_t S1 = ROR32(e, 6) ^ ROR32(e, 11) ^ ROR32(e, 25);
    uint32_t ch = (e & f) ^ (~e & g);
    uint32_t t1 = hh + S1 + ch + K256[i] + w[i];
    uint32_t S0 = ROR32(a, 2) ^ ROR32(a, 13) ^ ROR32(a, 22);
    uint32_t maj = (a & b) ^ (a & c) ^ (b & c);
    uint32_t t2 = S0 + maj;
    hh = g; g = f; f = e; e = d + t1; d = c; c = b; b = a; a = t1 + t2;
  }
  h[0] += a; h[1] += b; h[2] += c; h[3] += d; h[4] += e; h[5] += f; h[6] += g; h[7] += hh;
}

void or_sha256(const uint8_t* msg, size_t len, uint8_t out[32]) {
  init_consts();
  uint32_t h[8];
  memcpy(h, H256, sizeof h);
  size_t off = 0;
  for (; off + 64 <= len; off += 64) sha256_block(h, msg + off);
  uint8_t blk[128];
  memset(blk, 0, sizeof blk);
  size_t rem = len - off;
  if (rem) memcpy(blk, msg + off, rem);
  blk[rem] = 0x80;
  size_t nb = (rem + 1 + 8 <= 64) ? 1 : 2;
  uint64_t bits = (uint64_t)len * 8;
  for (int j = 0; j < 8; j++) blk[nb * 64 - 1 - j] = (uint8_t)(bits >> (8 * j));
  for (size_t b = 0; b < nb; b++) sha256_block(h, blk + b * 64);
  for (int i = 0; i < 8; i++)
    for (int j = 0; j < 4; j++) out[i * 4 + j] = (uint8_t)(h[i] >> (24 - 8 * j));
}
