/*
 * ed25519.c -- Ed25519 verification with ed25519-dalek 2.1 `VerifyingKey::verify` semantics, plus
 * RFC 8032 key generation and signing, for the oracle (test infrastructure only).
 *
 * Reference: the CPU restatement src/curve_algos/ed25519_eddsa.rs:16-184 (sqrt_ratio_i :156-184,
 * decompress :79-101, hash :109-124, compression :126-133) and the GPU flow
 * src/wgsl/ed25519_eddsa.wgsl:6-63, src/wgsl/ed25519_utils.wgsl:27-118,
 * src/wgsl/main/ed25519_eddsa_main_0.wgsl:23-86 / _5.wgsl:20-46.
 * Acceptance rules are dalek's (the tests' oracle is `verifying_key.verify`,
 * src/tests/ed25519_eddsa.rs:26), which the reference WGSL does not fully implement (no s < l check,
 * src/wgsl/main/ed25519_eddsa_main_0.wgsl:51-54):
 *   - s must be canonical (s < l), else reject;
 *   - A must decompress: y = bytes with bit 255 cleared, taken mod p (non-canonical y accepted),
 *     x from sqrt_ratio_i(y^2-1, d y^2+1), negated when the sign bit is set (also for x = 0);
 *   - k = SHA-512(R_bytes || A_bytes || M) mod l over the bytes as given;
 *   - accept iff compress([s]B - [k]A) == R_bytes (cofactorless, byte comparison).
 * Group law: the unified extended-coordinates addition of RFC 8032 section 5.1.4, complete on the
 * whole curve because d is a non-square.
 */
#include <string.h>
#include "oracle_internal.h"

typedef struct { u256 X, Y, Z, T; } ept; /* Montgomery domain mod p */

static const modctx* Fp;
static const modctx* Fl;
static u256 D, D2, SQRTM1, P58, ONE;
static ept BASE;
static volatile int ed_ready = 0;

static void ed_add(ept* r, const ept* p, const ept* q) {
  u256 a, b, c, d, e, f, g, h, t1, t2;
  mod_sub(Fp, &t1, &p->Y, &p->X);
  mod_sub(Fp, &t2, &q->Y, &q->X);
  mod_mul(Fp, &a, &t1, &t2);
  mod_add(Fp, &t1, &p->Y, &p->X);
  mod_add(Fp, &t2, &q->Y, &q->X);
  mod_mul(Fp, &b, &t1, &t2);
  mod_mul(Fp, &c, &p->T, &D2);
  mod_mul(Fp, &c, &c, &q->T);
  mod_mul(Fp, &d, &p->Z, &q->Z);
  mod_add(Fp, &d, &d, &d);
  mod_sub(Fp, &e, &b, &a);
  mod_sub(Fp, &f, &d, &c);
  mod_add(Fp, &g, &d, &c);
  mod_add(Fp, &h, &b, &a);
  mod_mul(Fp, &r->X, &e, &f);
  mod_mul(Fp, &r->Y, &g, &h);
  mod_mul(Fp, &r->T, &e, &h);
  mod_mul(Fp, &r->Z, &f, &g);
}

static void ed_mul(ept* r, const ept* p, const u256* k) {
  ept acc = {{{0}}, ONE, ONE, {{0}}};
  for (int i = 255; i >= 0; i--) {
    ed_add(&acc, &acc, &acc);
    if (u256_bit(k, i)) ed_add(&acc, &acc, p);
  }
  *r = acc;
}

static int is_negative(const u256* xm) { /* parity of the canonical value */
  u256 x;
  mod_from(Fp, &x, xm);
  return (int)(x.v[0] & 1);
}

/* dalek FieldElement::sqrt_ratio_i; returns was_nonzero_square */
static int sqrt_ratio_i(u256* out, const u256* u, const u256* v) {
  u256 v2, v3, v7, uv3, uv7, r, check, negu, negui, rprime, t;
  mod_mul(Fp, &v2, v, v);
  mod_mul(Fp, &v3, &v2, v);
  mod_mul(Fp, &v7, &v3, &v3);
  mod_mul(Fp, &v7, &v7, v);
  mod_mul(Fp, &uv3, u, &v3);
  mod_mul(Fp, &uv7, u, &v7);
  mod_pow(Fp, &t, &uv7, &P58);
  mod_mul(Fp, &r, &uv3, &t);
  mod_mul(Fp, &check, &r, &r);
  mod_mul(Fp, &check, &check, v);
  mod_neg(Fp, &negu, u);
  mod_mul(Fp, &negui, &negu, &SQRTM1);
  int correct = mod_eq(&check, u);
  int flipped = mod_eq(&check, &negu);
  int flipped_i = mod_eq(&check, &negui);
  mod_mul(Fp, &rprime, &SQRTM1, &r);
  if (flipped || flipped_i) r = rprime;
  if (is_negative(&r)) mod_neg(Fp, &r, &r);
  *out = r;
  return correct || flipped;
}

static int decompress(ept* P, const uint8_t in[32]) {
  uint8_t b[32];
  memcpy(b, in, 32);
  int sign = b[31] >> 7;
  b[31] &= 0x7f;
  u256 y, ym, yy, u, v, x;
  u256_from_le(&y, b);
  mod_reduce256(Fp, &y, &y);
  mod_to(Fp, &ym, &y);
  mod_mul(Fp, &yy, &ym, &ym);
  mod_sub(Fp, &u, &yy, &ONE);
  mod_mul(Fp, &v, &yy, &D);
  mod_add(Fp, &v, &v, &ONE);
  if (!sqrt_ratio_i(&x, &u, &v)) return 0;
  if (sign) mod_neg(Fp, &x, &x);
  P->X = x;
  P->Y = ym;
  P->Z = ONE;
  mod_mul(Fp, &P->T, &x, &ym);
  return 1;
}

static void compress(uint8_t out[32], const ept* P) {
  u256 zi, x, y, xp, yp;
  mod_inv(Fp, &zi, &P->Z);
  mod_mul(Fp, &x, &P->X, &zi);
  mod_mul(Fp, &y, &P->Y, &zi);
  mod_from(Fp, &xp, &x);
  mod_from(Fp, &yp, &y);
  u256_to_le(out, &yp);
  out[31] ^= (uint8_t)((xp.v[0] & 1) << 7);
}

static void init_ed(void) {
  if (ed_ready) return;
  Fp = or_modulus(4);
  Fl = or_modulus(5);
  ONE = Fp->one;
  u256 t, a, b;
  /* d = -121665 / 121666 */
  u256 c1 = {{121665, 0, 0, 0}}, c2 = {{121666, 0, 0, 0}};
  mod_to(Fp, &a, &c1);
  mod_neg(Fp, &a, &a);
  mod_to(Fp, &b, &c2);
  mod_inv(Fp, &t, &b);
  mod_mul(Fp, &D, &a, &t);
  mod_add(Fp, &D2, &D, &D);
  /* sqrt(-1) = 2^((p-1)/4) */
  u256 two = {{2, 0, 0, 0}}, e = Fp->m, one = {{1, 0, 0, 0}}, twom;
  u256_sub(&e, &e, &one);
  for (int k = 0; k < 2; k++)
    for (int i = 0; i < 4; i++) e.v[i] = (e.v[i] >> 1) | (i < 3 ? e.v[i + 1] << 63 : 0);
  mod_to(Fp, &twom, &two);
  mod_pow(Fp, &SQRTM1, &twom, &e);
  /* (p-5)/8 */
  u256 five = {{5, 0, 0, 0}};
  P58 = Fp->m;
  u256_sub(&P58, &P58, &five);
  for (int k = 0; k < 3; k++)
    for (int i = 0; i < 4; i++) P58.v[i] = (P58.v[i] >> 1) | (i < 3 ? P58.v[i + 1] << 63 : 0);
  /* base point: y = 4/5, x even */
  u256 four = {{4, 0, 0, 0}}, f4, f5, y, yp;
  mod_to(Fp, &f4, &four);
  mod_to(Fp, &f5, &five);
  mod_inv(Fp, &t, &f5);
  mod_mul(Fp, &y, &f4, &t);
  mod_from(Fp, &yp, &y);
  uint8_t enc[32];
  u256_to_le(enc, &yp);
  decompress(&BASE, enc);
  __atomic_store_n(&ed_ready, 1, __ATOMIC_RELEASE);
}

/* initialise the constant tables once, before any thread can call in */
__attribute__((constructor)) static void ed_init_ctor(void) { init_ed(); }

int or_ed25519_decompress(const uint8_t in[32], uint8_t out_xy[64]) {
  init_ed();
  ept P;
  memset(out_xy, 0, 64);
  if (!decompress(&P, in)) return 0;
  u256 x, y;
  mod_from(Fp, &x, &P.X);
  mod_from(Fp, &y, &P.Y);
  u256_to_le(out_xy, &x);
  u256_to_le(out_xy + 32, &y);
  return 1;
}

static void hash_mod_l(u256* k, const uint8_t h[64]) {
  uint64_t x[8];
  for (int i = 0; i < 8; i++) {
    uint64_t w = 0;
    for (int j = 7; j >= 0; j--) w = (w << 8) | h[i * 8 + j];
    x[i] = w;
  }
  mod_reduce512(Fl, k, x);
}

void or_reduce512_l(const uint8_t in[64], uint8_t out[32]) {
  init_ed();
  u256 k;
  hash_mod_l(&k, in);
  u256_to_le(out, &k);
}

/* SHA-512 over the concatenation of up to three pieces */
#include <stdlib.h>
static void sha512_3(uint8_t out[64], const uint8_t* a, size_t la, const uint8_t* b, size_t lb,
                     const uint8_t* c, size_t lc) {
  size_t n = la + lb + lc;
  uint8_t stackbuf[4096];
  uint8_t* buf = n <= sizeof stackbuf ? stackbuf : (uint8_t*)malloc(n);
  if (la) memcpy(buf, a, la);
  if (lb) memcpy(buf + la, b, lb);
  if (lc) memcpy(buf + la + lb, c, lc);
  or_sha512(buf, n, out);
  if (buf != stackbuf) free(buf);
}

/* order divides 8: [8]P is the identity (X = 0, Y = Z) */
static int small_order(const ept* P) {
  ept Q = *P;
  for (int i = 0; i < 3; i++) ed_add(&Q, &Q, &Q);
  return u256_is_zero(&Q.X) && mod_eq(&Q.Y, &Q.Z);
}

int or_ed25519_is_small_order(const uint8_t enc[32]) {
  init_ed();
  ept P;
  return decompress(&P, enc) && small_order(&P);
}

/* verify (strict = 0) and dalek 2.1 verify_strict (strict = 1), which also requires that R
   decompresses and that neither R nor A has small order; both compare compressed R bytes */
static int verify_impl(const uint8_t sig[64], const uint8_t pk[32], const uint8_t* msg, size_t len, int strict) {
  init_ed();
  u256 s;
  u256_from_le(&s, sig + 32);
  if (u256_cmp(&s, &Fl->m) >= 0) return 0;
  ept A;
  if (!decompress(&A, pk)) return 0;
  if (strict) {
    ept R;
    if (!decompress(&R, sig)) return 0;
    if (small_order(&R) || small_order(&A)) return 0;
  }
  uint8_t h[64];
  sha512_3(h, sig, 32, pk, 32, msg, len);
  u256 k;
  hash_mod_l(&k, h);
  /* -A */
  mod_neg(Fp, &A.X, &A.X);
  mod_neg(Fp, &A.T, &A.T);
  ept sB, kA, P;
  ed_mul(&sB, &BASE, &s);
  ed_mul(&kA, &A, &k);
  ed_add(&P, &sB, &kA);
  uint8_t enc[32];
  compress(enc, &P);
  return memcmp(enc, sig, 32) == 0;
}

int or_ed25519_verify(const uint8_t sig[64], const uint8_t pk[32], const uint8_t* msg, size_t len) {
  return verify_impl(sig, pk, msg, len, 0);
}

int or_ed25519_verify_strict(const uint8_t sig[64], const uint8_t pk[32], const uint8_t* msg, size_t len) {
  return verify_impl(sig, pk, msg, len, 1);
}

static void expand_seed(const uint8_t seed[32], u256* a, uint8_t prefix[32]) {
  uint8_t h[64];
  or_sha512(seed, 32, h);
  h[0] &= 248;
  h[31] &= 127;
  h[31] |= 64;
  u256_from_le(a, h);
  memcpy(prefix, h + 32, 32);
}

void or_ed25519_pubkey(const uint8_t seed[32], uint8_t out_pk[32]) {
  init_ed();
  u256 a;
  uint8_t prefix[32];
  expand_seed(seed, &a, prefix);
  ept A;
  ed_mul(&A, &BASE, &a);
  compress(out_pk, &A);
}

void or_ed25519_sign(const uint8_t seed[32], const uint8_t* msg, size_t len, uint8_t out_sig[64]) {
  init_ed();
  u256 a, r, k, am, rm, km, sm, s;
  uint8_t prefix[32], pk[32], h[64], Renc[32];
  expand_seed(seed, &a, prefix);
  ept A, R;
  ed_mul(&A, &BASE, &a);
  compress(pk, &A);
  sha512_3(h, prefix, 32, msg, len, NULL, 0);
  hash_mod_l(&r, h);
  ed_mul(&R, &BASE, &r);
  compress(Renc, &R);
  sha512_3(h, Renc, 32, pk, 32, msg, len);
  hash_mod_l(&k, h);
  /* S = r + k*a mod l */
  u256 ared;
  mod_reduce256(Fl, &ared, &a);
  mod_to(Fl, &am, &ared);
  mod_to(Fl, &km, &k);
  mod_to(Fl, &rm, &r);
  mod_mul(Fl, &sm, &km, &am);
  mod_add(Fl, &sm, &sm, &rm);
  mod_from(Fl, &s, &sm);
  memcpy(out_sig, Renc, 32);
  u256_to_le(out_sig + 32, &s);
}
