/*
 * secp.c -- secp256k1 / secp256r1 group law, ECDSA public-key recovery and fuel-format signing
 * for the oracle (test infrastructure only).
 *
 * Recovery follows the reference flow (src/wgsl/secp256k1_ecdsa.wgsl:7-130,
 * src/curve_algos/secp256k1_ecdsa.rs:66-128, src/curve_algos/secp256r1_ecdsa.rs:16-117):
 * decode the recid parity from bit 7 of s's first byte (src/wgsl/signature.wgsl:6-21,
 * src/tests/mod.rs:151-163), lift x = r to R = (r, y) with the requested parity, then
 * Q = (-z/r)*G + (s/r)*R. Acceptance rules are those of the reference's CPU path rather than of
 * its WGSL (which has no range checks, src/wgsl/secp256k1_ecdsa.wgsl:29-37): fuel-crypto 0.49 over libsecp256k1
 * (secp256k1_ecdsa_recoverable_signature_parse_compact / secp256k1_ecdsa_sig_recover) for k1 and
 * fuel-crypto's p256 backend over ecdsa 0.16 recover_from_prehash for r1:
 *   r in [1, n-1], s in [1, n-1] (recid bit cleared), z = digest mod n,
 *   R.x = r (fuel never sets the x-reduced recid bit), R must exist, Q != infinity.
 * The group law is textbook Jacobian arithmetic with every exceptional case branched out, so it is
 * complete; it is pinned in tests/test_oracle.py and tests/test_oracle_pins.py by affine Python
 * arithmetic and OpenSSL.
 */
#include <string.h>
#include "oracle_internal.h"

static void hex_to_u256(u256* r, const char* hex) {
  uint8_t b[32];
  for (int i = 0; i < 32; i++) {
    unsigned v = 0;
    for (int j = 0; j < 2; j++) {
      char ch = hex[2 * i + j];
      v <<= 4;
      if (ch >= '0' && ch <= '9') v |= (unsigned)(ch - '0');
      else if (ch >= 'a' && ch <= 'f') v |= (unsigned)(ch - 'a' + 10);
      else if (ch >= 'A' && ch <= 'F') v |= (unsigned)(ch - 'A' + 10);
    }
    b[i] = (uint8_t)v;
  }
  u256_from_be(r, b);
}

static modctx MODS[6];
static curve_t CURVES[2];
static volatile int ready = 0;

static void init_all(void) {
  if (ready) return;
  static const char* mhex[6] = {
      "FFFFFFFFFFFFFFFFFFFFFFFFFFFFFFFFFFFFFFFFFFFFFFFFFFFFFFFEFFFFFC2F", /* secp256k1 p */
      "FFFFFFFFFFFFFFFFFFFFFFFFFFFFFFFEBAAEDCE6AF48A03BBFD25E8CD0364141", /* secp256k1 n */
      "FFFFFFFF00000001000000000000000000000000FFFFFFFFFFFFFFFFFFFFFFFF", /* P-256 p */
      "FFFFFFFF00000000FFFFFFFFFFFFFFFFBCE6FAADA7179E84F3B9CAC2FC632551", /* P-256 n */
      "7FFFFFFFFFFFFFFFFFFFFFFFFFFFFFFFFFFFFFFFFFFFFFFFFFFFFFFFFFFFFFED", /* 2^255 - 19 */
      "1000000000000000000000000000000014DEF9DEA2F79CD65812631A5CF5D3ED", /* l */
  };
  for (int i = 0; i < 6; i++) {
    u256 m;
    hex_to_u256(&m, mhex[i]);
    mod_init(&MODS[i], &m);
  }
  static const char* chex[2][4] = {
      {"0000000000000000000000000000000000000000000000000000000000000000",
       "0000000000000000000000000000000000000000000000000000000000000007",
       "79BE667EF9DCBBAC55A06295CE870B07029BFCDB2DCE28D959F2815B16F81798",
       "483ADA7726A3C4655DA4FBFC0E1108A8FD17B448A68554199C47D08FFB10D4B8"},
      {"FFFFFFFF00000001000000000000000000000000FFFFFFFFFFFFFFFFFFFFFFFC",
       "5AC635D8AA3A93E7B3EBBD55769886BC651D06B0CC53B0F63BCE3C3E27D2604B",
       "6B17D1F2E12C4247F8BCE6E563A440F277037D812DEB33A0F4A13945D898C296",
       "4FE342E2FE1A7F9B8EE7EB4A7C0F9E162BCE33576B315ECECBB6406837BF51F5"}};
  for (int c = 0; c < 2; c++) {
    curve_t* C = &CURVES[c];
    C->fp = &MODS[2 * c];
    C->fn = &MODS[2 * c + 1];
    u256 t;
    hex_to_u256(&t, chex[c][0]); mod_to(C->fp, &C->a, &t);
    hex_to_u256(&t, chex[c][1]); mod_to(C->fp, &C->b, &t);
    hex_to_u256(&t, chex[c][2]); mod_to(C->fp, &C->gx, &t);
    hex_to_u256(&t, chex[c][3]); mod_to(C->fp, &C->gy, &t);
    /* (p+1)/4 */
    u256 one = {{1, 0, 0, 0}}, e;
    u256_add(&e, &C->fp->m, &one);
    for (int k = 0; k < 2; k++) { /* shift right by 2 */
      for (int i = 0; i < 4; i++) e.v[i] = (e.v[i] >> 1) | (i < 3 ? e.v[i + 1] << 63 : 0);
    }
    C->sqrt_exp = e;
  }
  __atomic_store_n(&ready, 1, __ATOMIC_RELEASE);
}

/* build the constant tables once, before any thread can call in */
__attribute__((constructor)) static void secp_init_ctor(void) { init_all(); }

const modctx* or_modulus(int idx) { init_all(); return &MODS[idx]; }
const curve_t* or_curve(int which) { init_all(); return &CURVES[which]; }

void jp_dbl(const curve_t* C, jpoint* r, const jpoint* p) {
  const modctx* F = C->fp;
  if (p->inf || u256_is_zero(&p->y)) { memset(r, 0, sizeof *r); r->inf = 1; return; }
  u256 XX, YY, YYYY, ZZ, S, M, t, X3, Y3, Z3;
  mod_mul(F, &XX, &p->x, &p->x);
  mod_mul(F, &YY, &p->y, &p->y);
  mod_mul(F, &YYYY, &YY, &YY);
  mod_mul(F, &ZZ, &p->z, &p->z);
  mod_mul(F, &S, &p->x, &YY);
  mod_add(F, &S, &S, &S);
  mod_add(F, &S, &S, &S);                     /* S = 4 X Y^2 */
  mod_add(F, &M, &XX, &XX);
  mod_add(F, &M, &M, &XX);                    /* 3 X^2 */
  mod_mul(F, &t, &ZZ, &ZZ);
  mod_mul(F, &t, &t, &C->a);
  mod_add(F, &M, &M, &t);                     /* M = 3X^2 + a Z^4 */
  mod_mul(F, &X3, &M, &M);
  mod_sub(F, &X3, &X3, &S);
  mod_sub(F, &X3, &X3, &S);                   /* X3 = M^2 - 2S */
  mod_sub(F, &t, &S, &X3);
  mod_mul(F, &Y3, &M, &t);
  mod_add(F, &t, &YYYY, &YYYY);
  mod_add(F, &t, &t, &t);
  mod_add(F, &t, &t, &t);                     /* 8 Y^4 */
  mod_sub(F, &Y3, &Y3, &t);
  mod_mul(F, &Z3, &p->y, &p->z);
  mod_add(F, &Z3, &Z3, &Z3);
  r->x = X3; r->y = Y3; r->z = Z3; r->inf = 0;
}

void jp_add(const curve_t* C, jpoint* r, const jpoint* p, const jpoint* q) {
  const modctx* F = C->fp;
  if (p->inf) { *r = *q; return; }
  if (q->inf) { *r = *p; return; }
  u256 Z1Z1, Z2Z2, U1, U2, S1, S2, H, R, H2, H3, U1H2, t, X3, Y3, Z3;
  mod_mul(F, &Z1Z1, &p->z, &p->z);
  mod_mul(F, &Z2Z2, &q->z, &q->z);
  mod_mul(F, &U1, &p->x, &Z2Z2);
  mod_mul(F, &U2, &q->x, &Z1Z1);
  mod_mul(F, &S1, &p->y, &q->z);
  mod_mul(F, &S1, &S1, &Z2Z2);
  mod_mul(F, &S2, &q->y, &p->z);
  mod_mul(F, &S2, &S2, &Z1Z1);
  if (mod_eq(&U1, &U2)) {
    if (mod_eq(&S1, &S2)) { jp_dbl(C, r, p); return; }
    memset(r, 0, sizeof *r);
    r->inf = 1;
    return;
  }
  mod_sub(F, &H, &U2, &U1);
  mod_sub(F, &R, &S2, &S1);
  mod_mul(F, &H2, &H, &H);
  mod_mul(F, &H3, &H2, &H);
  mod_mul(F, &U1H2, &U1, &H2);
  mod_mul(F, &X3, &R, &R);
  mod_sub(F, &X3, &X3, &H3);
  mod_sub(F, &X3, &X3, &U1H2);
  mod_sub(F, &X3, &X3, &U1H2);
  mod_sub(F, &t, &U1H2, &X3);
  mod_mul(F, &Y3, &R, &t);
  mod_mul(F, &t, &S1, &H3);
  mod_sub(F, &Y3, &Y3, &t);
  mod_mul(F, &Z3, &p->z, &q->z);
  mod_mul(F, &Z3, &Z3, &H);
  r->x = X3; r->y = Y3; r->z = Z3; r->inf = 0;
}

void jp_mul(const curve_t* C, jpoint* r, const jpoint* p, const u256* k) {
  jpoint acc;
  memset(&acc, 0, sizeof acc);
  acc.inf = 1;
  for (int i = 255; i >= 0; i--) {
    jp_dbl(C, &acc, &acc);
    if (u256_bit(k, i)) jp_add(C, &acc, &acc, p);
  }
  *r = acc;
}

void jp_mul2(const curve_t* C, jpoint* r, const jpoint* p, const u256* u1, const u256* u2) {
  jpoint g = {C->gx, C->gy, C->fp->one, 0}, gp, acc;
  jp_add(C, &gp, &g, p);
  memset(&acc, 0, sizeof acc);
  acc.inf = 1;
  for (int i = 255; i >= 0; i--) {
    jp_dbl(C, &acc, &acc);
    int b1 = u256_bit(u1, i), b2 = u256_bit(u2, i);
    if (b1 && b2) jp_add(C, &acc, &acc, &gp);
    else if (b1) jp_add(C, &acc, &acc, &g);
    else if (b2) jp_add(C, &acc, &acc, p);
  }
  *r = acc;
}

int jp_to_affine(const curve_t* C, u256* x, u256* y, const jpoint* p) {
  const modctx* F = C->fp;
  if (p->inf || u256_is_zero(&p->z)) return 0;
  u256 zi, zi2, zi3, xm, ym;
  mod_inv(F, &zi, &p->z);
  mod_mul(F, &zi2, &zi, &zi);
  mod_mul(F, &zi3, &zi2, &zi);
  mod_mul(F, &xm, &p->x, &zi2);
  mod_mul(F, &ym, &p->y, &zi3);
  mod_from(F, x, &xm);
  mod_from(F, y, &ym);
  return 1;
}

/* y^2 = x^3 + a x + b (Montgomery domain in, returns rhs) */
static void curve_rhs(const curve_t* C, u256* rhs, const u256* xm) {
  const modctx* F = C->fp;
  u256 t, ax;
  mod_mul(F, &t, xm, xm);
  mod_mul(F, &t, &t, xm);
  mod_mul(F, &ax, &C->a, xm);
  mod_add(F, &t, &t, &ax);
  mod_add(F, rhs, &t, &C->b);
}

static int load_affine(const curve_t* C, jpoint* P, const uint8_t in[64]) {
  u256 x, y, xm, ym, rhs, yy;
  u256_from_be(&x, in);
  u256_from_be(&y, in + 32);
  if (u256_cmp(&x, &C->fp->m) >= 0 || u256_cmp(&y, &C->fp->m) >= 0) return 0;
  mod_to(C->fp, &xm, &x);
  mod_to(C->fp, &ym, &y);
  curve_rhs(C, &rhs, &xm);
  mod_mul(C->fp, &yy, &ym, &ym);
  if (!mod_eq(&yy, &rhs)) return 0;
  P->x = xm; P->y = ym; P->z = C->fp->one; P->inf = 0;
  return 1;
}

static void store_affine(const curve_t* C, uint8_t out[64], const jpoint* P) {
  u256 x, y;
  if (!jp_to_affine(C, &x, &y, P)) { memset(out, 0, 64); return; }
  u256_to_be(out, &x);
  u256_to_be(out + 32, &y);
}

int or_secp_recover(int curve, const uint8_t sig[64], const uint8_t digest[32], uint8_t out_pub[64]) {
  const curve_t* C = or_curve(curve);
  const modctx* Fp = C->fp;
  const modctx* Fn = C->fn;
  memset(out_pub, 0, 64);
  uint8_t sb[32];
  memcpy(sb, sig + 32, 32);
  int y_odd = (sb[0] & 0x80) != 0;
  sb[0] &= 0x7f;
  u256 r, s, z, zr;
  u256_from_be(&r, sig);
  u256_from_be(&s, sb);
  u256_from_be(&z, digest);
  if (u256_is_zero(&r) || u256_cmp(&r, &Fn->m) >= 0) return OR_ST_BAD_R;
  if (u256_is_zero(&s) || u256_cmp(&s, &Fn->m) >= 0) return OR_ST_BAD_S;
  mod_reduce256(Fn, &zr, &z);
  /* lift x = r (r < n < p) */
  u256 xm, rhs, ym, chk, y;
  mod_to(Fp, &xm, &r);
  curve_rhs(C, &rhs, &xm);
  mod_pow(Fp, &ym, &rhs, &C->sqrt_exp);
  mod_mul(Fp, &chk, &ym, &ym);
  if (!mod_eq(&chk, &rhs)) return OR_ST_NO_POINT;
  mod_from(Fp, &y, &ym);
  if ((int)(y.v[0] & 1) != y_odd) mod_neg(Fp, &ym, &ym);
  jpoint R = {xm, ym, Fp->one, 0};
  /* u1 = -z/r, u2 = s/r mod n (computed in the Montgomery domain of n) */
  u256 rm, rinv, zm, sm, u1m, u2m, u1, u2;
  mod_to(Fn, &rm, &r);
  mod_inv(Fn, &rinv, &rm);
  mod_to(Fn, &zm, &zr);
  mod_to(Fn, &sm, &s);
  mod_mul(Fn, &u1m, &zm, &rinv);
  mod_neg(Fn, &u1m, &u1m);
  mod_mul(Fn, &u2m, &sm, &rinv);
  mod_from(Fn, &u1, &u1m);
  mod_from(Fn, &u2, &u2m);
  jpoint Q;
  jp_mul2(C, &Q, &R, &u1, &u2);
  if (Q.inf) return OR_ST_INFINITY;
  store_affine(C, out_pub, &Q);
  return OR_ST_OK;
}

/*
 * ECDSA verification with a known public key, fuel `Signature::verify(&pk, &message)`
 * (the Fuel predicate next to recovery). secp256k1 follows fuel-crypto 0.49 over libsecp256k1
 * (secp256k1_ecdsa_signature_parse_compact, secp256k1_ec_pubkey_parse of 0x04 || X || Y,
 * secp256k1_ecdsa_verify); secp256r1 follows p256 0.13 / ecdsa 0.16 VerifyingKey::verify_prehash:
 *   r, s in [1, n-1] after clearing the recid bit of s (which verification ignores);
 *   secp256k1 only: s <= (n-1)/2 (libsecp256k1 accepts lower-S signatures only);
 *   Q = X || Y with X, Y < p on the curve;
 *   z = digest mod n, u1 = z/s, u2 = r/s, P = u1*G + u2*Q != infinity, accept iff x(P) mod n == r.
 */
int or_secp_verify(int curve, const uint8_t sig[64], const uint8_t digest[32], const uint8_t pub[64]) {
  const curve_t* C = or_curve(curve);
  const modctx* Fn = C->fn;
  uint8_t sb[32];
  memcpy(sb, sig + 32, 32);
  sb[0] &= 0x7f;
  u256 r, s, z, zr;
  u256_from_be(&r, sig);
  u256_from_be(&s, sb);
  u256_from_be(&z, digest);
  if (u256_is_zero(&r) || u256_cmp(&r, &Fn->m) >= 0) return OR_ST_BAD_R;
  if (u256_is_zero(&s) || u256_cmp(&s, &Fn->m) >= 0) return OR_ST_BAD_S;
  if (curve == OR_CURVE_K1) {
    u256 half = Fn->m;
    for (int i = 0; i < 4; i++) half.v[i] = (half.v[i] >> 1) | (i < 3 ? half.v[i + 1] << 63 : 0);
    if (u256_cmp(&s, &half) > 0) return OR_ST_HIGH_S;
  }
  jpoint Q;
  if (!load_affine(C, &Q, pub)) return OR_ST_BAD_KEY;
  mod_reduce256(Fn, &zr, &z);
  u256 sm, sinv, zm, rm, u1m, u2m, u1, u2;
  mod_to(Fn, &sm, &s);
  mod_inv(Fn, &sinv, &sm);
  mod_to(Fn, &zm, &zr);
  mod_to(Fn, &rm, &r);
  mod_mul(Fn, &u1m, &zm, &sinv);
  mod_mul(Fn, &u2m, &rm, &sinv);
  mod_from(Fn, &u1, &u1m);
  mod_from(Fn, &u2, &u2m);
  jpoint P;
  jp_mul2(C, &P, &Q, &u1, &u2);
  u256 x, y;
  if (!jp_to_affine(C, &x, &y, &P)) return OR_ST_INFINITY;
  if (u256_cmp(&x, &Fn->m) >= 0) u256_sub(&x, &x, &Fn->m); /* x < p < 2n */
  return u256_cmp(&x, &r) == 0 ? OR_ST_OK : OR_ST_MISMATCH;
}

int or_secp_pubkey(int curve, const uint8_t sk[32], uint8_t out_pub[64]) {
  const curve_t* C = or_curve(curve);
  u256 d;
  u256_from_be(&d, sk);
  memset(out_pub, 0, 64);
  if (u256_is_zero(&d) || u256_cmp(&d, &C->fn->m) >= 0) return -1;
  jpoint G = {C->gx, C->gy, C->fp->one, 0}, Q;
  jp_mul(C, &Q, &G, &d);
  store_affine(C, out_pub, &Q);
  return 0;
}

int or_secp_sign(int curve, const uint8_t sk[32], const uint8_t digest[32], const uint8_t nonce[32],
                 uint8_t out_sig[64]) {
  const curve_t* C = or_curve(curve);
  const modctx* Fn = C->fn;
  u256 d, k, z, zr;
  u256_from_be(&d, sk);
  u256_from_be(&k, nonce);
  u256_from_be(&z, digest);
  if (u256_is_zero(&k) || u256_cmp(&k, &Fn->m) >= 0) return -1;
  if (u256_is_zero(&d) || u256_cmp(&d, &Fn->m) >= 0) return -1;
  jpoint G = {C->gx, C->gy, C->fp->one, 0}, Rp;
  jp_mul(C, &Rp, &G, &k);
  u256 rx, ry;
  if (!jp_to_affine(C, &rx, &ry, &Rp)) return -1;
  if (u256_cmp(&rx, &Fn->m) >= 0) return -1; /* x-reduced R: fuel cannot encode it */
  u256 r = rx;
  if (u256_is_zero(&r)) return -1;
  mod_reduce256(Fn, &zr, &z);
  u256 km, kinv, rm, dm, zm, t, sm, s;
  mod_to(Fn, &km, &k);
  mod_inv(Fn, &kinv, &km);
  mod_to(Fn, &rm, &r);
  mod_to(Fn, &dm, &d);
  mod_to(Fn, &zm, &zr);
  mod_mul(Fn, &t, &rm, &dm);
  mod_add(Fn, &t, &t, &zm);
  mod_mul(Fn, &sm, &t, &kinv);
  mod_from(Fn, &s, &sm);
  if (u256_is_zero(&s)) return -1;
  int odd = (int)(ry.v[0] & 1);
  /* low-s normalisation: s > n/2 -> n - s and flip the parity */
  u256 half = Fn->m;
  for (int i = 0; i < 4; i++) half.v[i] = (half.v[i] >> 1) | (i < 3 ? half.v[i + 1] << 63 : 0);
  if (u256_cmp(&s, &half) > 0) {
    u256_sub(&s, &Fn->m, &s);
    odd ^= 1;
  }
  u256_to_be(out_sig, &r);
  u256_to_be(out_sig + 32, &s);
  out_sig[32] |= (uint8_t)(odd << 7);
  return 0;
}

/* ---- primitives for pinning tests ---- */
void or_mod_mul(int mod, const uint8_t a[32], const uint8_t b[32], uint8_t out[32]) {
  const modctx* F = or_modulus(mod);
  u256 x, y, xm, ym, rm, r;
  u256_from_be(&x, a);
  u256_from_be(&y, b);
  mod_to(F, &xm, &x);
  mod_to(F, &ym, &y);
  mod_mul(F, &rm, &xm, &ym);
  mod_from(F, &r, &rm);
  u256_to_be(out, &r);
}
void or_mod_inv(int mod, const uint8_t a[32], uint8_t out[32]) {
  const modctx* F = or_modulus(mod);
  u256 x, xm, rm, r;
  u256_from_be(&x, a);
  mod_to(F, &xm, &x);
  mod_inv(F, &rm, &xm);
  mod_from(F, &r, &rm);
  u256_to_be(out, &r);
}
void or_mod_pow(int mod, const uint8_t a[32], const uint8_t e[32], uint8_t out[32]) {
  const modctx* F = or_modulus(mod);
  u256 x, ee, xm, rm, r;
  u256_from_be(&x, a);
  u256_from_be(&ee, e);
  mod_to(F, &xm, &x);
  mod_pow(F, &rm, &xm, &ee);
  mod_from(F, &r, &rm);
  u256_to_be(out, &r);
}
int or_mod_sqrt(int mod, const uint8_t a[32], uint8_t out[32]) {
  const curve_t* C = or_curve(mod == 0 ? 0 : 1);
  const modctx* F = C->fp;
  u256 x, xm, ym, chk, y;
  u256_from_be(&x, a);
  mod_to(F, &xm, &x);
  mod_pow(F, &ym, &xm, &C->sqrt_exp);
  mod_mul(F, &chk, &ym, &ym);
  mod_from(F, &y, &ym);
  u256_to_be(out, &y);
  return mod_eq(&chk, &xm);
}
int or_secp_mul(int curve, const uint8_t P[64], const uint8_t k[32], uint8_t out[64]) {
  const curve_t* C = or_curve(curve);
  jpoint p, q;
  u256 kk;
  memset(out, 0, 64);
  if (!load_affine(C, &p, P)) return -1;
  u256_from_be(&kk, k);
  jp_mul(C, &q, &p, &kk);
  if (q.inf) return 1;
  store_affine(C, out, &q);
  return 0;
}
int or_secp_mul2(int curve, const uint8_t P[64], const uint8_t u1[32], const uint8_t u2[32], uint8_t out[64]) {
  const curve_t* C = or_curve(curve);
  jpoint p, q;
  u256 a, b;
  memset(out, 0, 64);
  if (!load_affine(C, &p, P)) return -1;
  u256_from_be(&a, u1);
  u256_from_be(&b, u2);
  jp_mul2(C, &q, &p, &a, &b);
  if (q.inf) return 1;
  store_affine(C, out, &q);
  return 0;
}
