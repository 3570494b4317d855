/* oracle_internal.h -- shared internals of the oracle (test infrastructure only). */
#ifndef SIGOPS_ORACLE_INTERNAL_H
#define SIGOPS_ORACLE_INTERNAL_H
#include <stddef.h>
#include <stdint.h>
#include "oracle.h"

typedef struct { uint64_t v[4]; } u256; /* little-endian 64-bit limbs */

typedef struct {
  u256 m;
  uint64_t m0inv; /* -m^-1 mod 2^64 */
  u256 r2;        /* 2^512 mod m */
  u256 one;       /* 2^256 mod m (Montgomery 1) */
} modctx;

void u256_from_be(u256* r, const uint8_t b[32]);
void u256_to_be(uint8_t b[32], const u256* a);
void u256_from_le(u256* r, const uint8_t b[32]);
void u256_to_le(uint8_t b[32], const u256* a);
int u256_cmp(const u256* a, const u256* b);
int u256_is_zero(const u256* a);
uint64_t u256_add(u256* r, const u256* a, const u256* b);
uint64_t u256_sub(u256* r, const u256* a, const u256* b);
int u256_bit(const u256* a, int i);

void mod_init(modctx* c, const u256* m);
void mod_add(const modctx* c, u256* r, const u256* a, const u256* b);
void mod_sub(const modctx* c, u256* r, const u256* a, const u256* b);
void mod_neg(const modctx* c, u256* r, const u256* a);
void mod_mul(const modctx* c, u256* r, const u256* a, const u256* b);
void mod_to(const modctx* c, u256* r, const u256* a);
void mod_from(const modctx* c, u256* r, const u256* a);
void mod_reduce256(const modctx* c, u256* r, const u256* a);
void mod_reduce512(const modctx* c, u256* r, const uint64_t x[8]);
void mod_pow(const modctx* c, u256* r, const u256* a, const u256* e);
void mod_inv(const modctx* c, u256* r, const u256* a);
int mod_eq(const u256* a, const u256* b);

/* moduli table, index as in or_mod_* */
const modctx* or_modulus(int idx);

/* secp curves */
typedef struct {
  const modctx* fp;
  const modctx* fn;
  u256 a, b;   /* Montgomery domain */
  u256 gx, gy; /* Montgomery domain */
  u256 sqrt_exp; /* (p+1)/4 */
} curve_t;
const curve_t* or_curve(int which);

typedef struct { u256 x, y, z; int inf; } jpoint; /* Jacobian, Montgomery domain */
void jp_add(const curve_t* C, jpoint* r, const jpoint* p, const jpoint* q);
void jp_dbl(const curve_t* C, jpoint* r, const jpoint* p);
void jp_mul(const curve_t* C, jpoint* r, const jpoint* p, const u256* k);
void jp_mul2(const curve_t* C, jpoint* r, const jpoint* p, const u256* u1, const u256* u2);
int jp_to_affine(const curve_t* C, u256* x, u256* y, const jpoint* p); /* plain domain; 0 if inf */

uint64_t splitmix64(uint64_t* s);
#endif
