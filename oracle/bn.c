/*
 * bn.c -- 256-bit modular arithmetic for the oracle (test infrastructure only).
 *
 * Plain Montgomery arithmetic over any odd 256-bit modulus with 4 x 64-bit limbs and
 * unsigned __int128 products. It replaces, for checking purposes, the reference's limb code:
 * bigint add/sub/compare (src/wgsl/bigint.wgsl:25-141), Barrett ff_mul (src/wgsl/ff.wgsl:169-198),
 * Montgomery mont_mul (src/wgsl/mont.wgsl:5-70), modpow (src/wgsl/mont.wgsl:80-99) and
 * ff_inverse (src/wgsl/ff.wgsl:60-120). Every modulus here is prime, so inversion is Fermat.
 * Pinned by tests/test_oracle_field.py against Python integers.
 */
#include "oracle_internal.h"
#include <string.h>

void u256_from_be(u256* r, const uint8_t b[32]) {
  for (int i = 0; i < 4; i++) {
    uint64_t w = 0;
    for (int j = 0; j < 8; j++) w = (w << 8) | b[(3 - i) * 8 + j];
    r->v[i] = w;
  }
}
void u256_to_be(uint8_t b[32], const u256* a) {
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 8; j++) b[(3 - i) * 8 + j] = (uint8_t)(a->v[i] >> (56 - 8 * j));
}
void u256_from_le(u256* r, const uint8_t b[32]) {
  for (int i = 0; i < 4; i++) {
    uint64_t w = 0;
    for (int j = 7; j >= 0; j--) w = (w << 8) | b[i * 8 + j];
    r->v[i] = w;
  }
}
void u256_to_le(uint8_t b[32], const u256* a) {
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 8; j++) b[i * 8 + j] = (uint8_t)(a->v[i] >> (8 * j));
}
int u256_cmp(const u256* a, const u256* b) {
  for (int i = 3; i >= 0; i--) {
    if (a->v[i] > b->v[i]) return 1;
    if (a->v[i] < b->v[i]) return -1;
  }
  return 0;
}
int u256_is_zero(const u256* a) { return (a->v[0] | a->v[1] | a->v[2] | a->v[3]) == 0; }
uint64_t u256_add(u256* r, const u256* a, const u256* b) {
  unsigned __int128 c = 0;
  for (int i = 0; i < 4; i++) {
    c += (unsigned __int128)a->v[i] + b->v[i];
    r->v[i] = (uint64_t)c;
    c >>= 64;
  }
  return (uint64_t)c;
}
uint64_t u256_sub(u256* r, const u256* a, const u256* b) {
  uint64_t borrow = 0;
  for (int i = 0; i < 4; i++) {
    unsigned __int128 d = (unsigned __int128)a->v[i] - b->v[i] - borrow;
    r->v[i] = (uint64_t)d;
    borrow = (uint64_t)(d >> 64) ? 1 : 0;
  }
  return borrow;
}
int u256_bit(const u256* a, int i) { return (int)((a->v[i >> 6] >> (i & 63)) & 1); }

/* -m^{-1} mod 2^64 by Newton iteration */
static uint64_t neg_inv64(uint64_t m0) {
  uint64_t x = 1;
  for (int i = 0; i < 7; i++) x *= 2 - m0 * x;
  return (uint64_t)0 - x;
}

void mod_init(modctx* c, const u256* m) {
  c->m = *m;
  c->m0inv = neg_inv64(m->v[0]);
  /* R mod m = 2^256 mod m, by doubling 1 256 times */
  u256 x = {{1, 0, 0, 0}};
  for (int i = 0; i < 512; i++) {
    if (i == 256) c->one = x;
    uint64_t carry = u256_add(&x, &x, &x);
    if (carry || u256_cmp(&x, m) >= 0) u256_sub(&x, &x, m);
  }
  c->r2 = x;
}

void mod_add(const modctx* c, u256* r, const u256* a, const u256* b) {
  uint64_t carry = u256_add(r, a, b);
  if (carry || u256_cmp(r, &c->m) >= 0) u256_sub(r, r, &c->m);
}
void mod_sub(const modctx* c, u256* r, const u256* a, const u256* b) {
  if (u256_sub(r, a, b)) u256_add(r, r, &c->m);
}
void mod_neg(const modctx* c, u256* r, const u256* a) {
  if (u256_is_zero(a)) { *r = *a; return; }
  u256_sub(r, &c->m, a);
}

/* CIOS Montgomery product: a*b*2^-256 mod m, for a*b < m*2^256 */
void mod_mul(const modctx* c, u256* r, const u256* a, const u256* b) {
  uint64_t t[6] = {0, 0, 0, 0, 0, 0};
  for (int i = 0; i < 4; i++) {
    unsigned __int128 carry = 0;
    for (int j = 0; j < 4; j++) {
      carry += (unsigned __int128)a->v[j] * b->v[i] + t[j];
      t[j] = (uint64_t)carry;
      carry >>= 64;
    }
    carry += t[4];
    t[4] = (uint64_t)carry;
    t[5] = (uint64_t)(carry >> 64);
    uint64_t mm = t[0] * c->m0inv;
    carry = (unsigned __int128)mm * c->m.v[0] + t[0];
    carry >>= 64;
    for (int j = 1; j < 4; j++) {
      carry += (unsigned __int128)mm * c->m.v[j] + t[j];
      t[j - 1] = (uint64_t)carry;
      carry >>= 64;
    }
    carry += t[4];
    t[3] = (uint64_t)carry;
    t[4] = t[5] + (uint64_t)(carry >> 64);
  }
  u256 res = {{t[0], t[1], t[2], t[3]}};
  if (t[4] || u256_cmp(&res, &c->m) >= 0) u256_sub(&res, &res, &c->m);
  *r = res;
}

void mod_to(const modctx* c, u256* r, const u256* a) { mod_mul(c, r, a, &c->r2); }
void mod_from(const modctx* c, u256* r, const u256* a) {
  u256 one = {{1, 0, 0, 0}};
  mod_mul(c, r, a, &one);
}
/* a (plain, any 256-bit value) reduced mod m */
void mod_reduce256(const modctx* c, u256* r, const u256* a) {
  *r = *a;
  while (u256_cmp(r, &c->m) >= 0) u256_sub(r, r, &c->m);
}
/* 512-bit little-endian limbs lo (v[0..3]) + hi (v[4..7]) reduced mod m, plain domain */
void mod_reduce512(const modctx* c, u256* r, const uint64_t x[8]) {
  u256 lo = {{x[0], x[1], x[2], x[3]}}, hi = {{x[4], x[5], x[6], x[7]}};
  u256 hr, lr, t;
  mod_reduce256(c, &hr, &hi);
  mod_reduce256(c, &lr, &lo);
  /* hi*2^256 mod m = mont(hi, R^2) ; then plain add */
  mod_mul(c, &t, &hr, &c->r2);
  mod_add(c, r, &t, &lr);
}
/* Montgomery-domain exponentiation, e is a plain exponent */
void mod_pow(const modctx* c, u256* r, const u256* a, const u256* e) {
  u256 acc = c->one;
  for (int i = 255; i >= 0; i--) {
    mod_mul(c, &acc, &acc, &acc);
    if (u256_bit(e, i)) mod_mul(c, &acc, &acc, a);
  }
  *r = acc;
}
void mod_inv(const modctx* c, u256* r, const u256* a) {
  u256 e = c->m, two = {{2, 0, 0, 0}};
  u256_sub(&e, &e, &two);
  mod_pow(c, r, a, &e);
}
int mod_eq(const u256* a, const u256* b) { return u256_cmp(a, b) == 0; }
