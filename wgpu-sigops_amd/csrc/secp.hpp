// secp.hpp -- secp256k1 / secp256r1 public-key recovery, one signature per lane.
//
// Reference semantics: src/wgsl/secp256k1_ecdsa.wgsl:7-130 and src/wgsl/secp256r1_ecdsa.wgsl:7-130
// (decode, lift R, u1 = -z/r, u2 = s/r, Q = u1*G + u2*R, affine output), with the acceptance rules
// of the reference's CPU path (fuel-crypto over libsecp256k1 / ecdsa 0.16): r, s in [1, n-1],
// z = digest mod n, R must exist, Q != infinity. Errors write 64 zero bytes and a status code.
//
// Algorithm (gfx950-oriented; the reference uses plain double-and-add, src/wgsl/secp256k1_curve.wgsl:277-294,
// and a w=4 fixed-base loop, :388-447, with Strauss-Shamir disabled, :336-383):
//   * one shared doubling chain (Strauss-Shamir) for u1*G + u2*R;
//   * both scalars in the signed all-odd digit form k = sum (2c_i - (2^w-1)) 2^(wi) with
//     C = (k >> 1) | 2^255 for odd k (even k: use n - k and negate the base), so every window
//     performs exactly one addition and no lane ever branches on a zero digit;
//   * u2*R: w = 4, table of odd multiples R..15R in per-lane memory, made affine with one
//     simultaneous inversion so that every addition of the R stream is a mixed addition;
//   * u1*G: w = 8, table of odd multiples G..255G (affine) staged in LDS by the kernel;
//   * Jacobian formulas dbl-2009-l (a = 0) / dbl-2001-b (a = -3), add-2007-bl, madd-2007-bl.
//     These are incomplete: an addition of P and +-P is detected (H == 0) and the lane is marked
//     for the careful path, which redoes the item with complete, branching formulas.
#pragma once
#include "inv.hpp"

namespace sigops {

enum : u32 {
  ST_OK = 0,
  ST_BAD_R = 1,
  ST_BAD_S = 2,
  ST_NO_POINT = 3,
  ST_INFINITY = 4,    // recovery: Q = infinity; verification: u1 G + u2 Q = infinity
  ST_HIGH_S = 5,      // verification, secp256k1: s > (n-1)/2
  ST_BAD_KEY = 6,     // verification: the key is not a curve point with X, Y < p
  ST_MISMATCH = 7,    // verification: x(u1 G + u2 Q) mod n != r
  ST_PENDING = 0xFF,  // exceptional addition met: the careful path finishes the item
};

struct CurveK1 {
  typedef ModPk1 Fp;
  typedef ModNk1 Fn;
  static constexpr bool A_IS_M3 = false;
  static constexpr bool GLV = true;    // phi(x, y) = (beta x, y) = [lambda](x, y)
  static constexpr bool LOW_S = true;  // verification accepts s <= (n-1)/2 only (libsecp256k1)
  DEV static fe b() { return fe_from_const(K1_B); }
  DEV static const u32 (&n32())[8] { return K1_N32; }
  DEV static const u32 (&p32())[8] { return K1_P32; }
  DEV static const u32 (&pmn32())[8] { return K1_PMN32; }
  DEV static const u32 (&nhalf32())[8] { return K1_NHALF32; }
  typedef CH_K1_SQRT SqrtChain;  // x^((p+1)/4)
};
struct CurveR1 {
  typedef ModPr1 Fp;
  typedef ModNr1 Fn;
  static constexpr bool A_IS_M3 = true;
  static constexpr bool GLV = false;
  static constexpr bool LOW_S = false;  // p256 / ecdsa 0.16 verification accepts either s
  DEV static fe b() { return fe_from_const(R1_B); }
  DEV static const u32 (&n32())[8] { return R1_N32; }
  DEV static const u32 (&p32())[8] { return R1_P32; }
  DEV static const u32 (&pmn32())[8] { return R1_PMN32; }
  DEV static const u32 (&nhalf32())[8] { return R1_N32; }  // unused (LOW_S is false)
  typedef CH_R1_SQRT SqrtChain;
};

struct jac {
  fe X, Y, Z;
};

// Point formulas on lazy field arithmetic (field.hpp). Inputs: coordinates normalised with value
// < 2M. Outputs: X3, Y3 from fnorm_red (value < (1+2^-19) M), Z3 from a multiplication (< 2M).
// Intermediate bounds are noted as [limb bound in units of 2^29, value bound in units of M];
// a product needs La*Lb < 7.1 and Va*Vb < 32.

template <class C>
DEV jac jdbl(const jac& P) {
  typedef typename C::Fp F;
  jac R;
  if (!C::A_IS_M3) {
    // a = 0, 3M + 4S: S = 4XY^2, M = 3X^2, X3 = M^2 - 2S, Y3 = M(S - X3) - 8Y^4, Z3 = 2YZ
    const fe XX = mont_sqr<F>(P.X);
    const fe YY = mont_sqr<F>(P.Y);                           // [1, 1.13]
    const fe S = mont_mul<F>(fscale_nr(P.X, 4), YY);          // [4, 8] x [1, 1.13] -> [1, 1.29]
    const fe Mn = fnorm_red<F>(fscale_nr(XX, 3));              // [1, 1+]
    R.X = fnorm_red<F>(fsub_nr<F, 3, 2>(mont_sqr<F>(Mn), fscale_nr(S, 2)));
    const fe Y4x4 = mont_sqr<F>(fscale_nr(YY, 2));            // 4 Y^4: [2, 2.26]^2 -> [1, 1.16]
    const fe MSX = mont_mul<F>(Mn, fsub_nr<F, 2, 1>(S, R.X));  // [1, 1+] x [3, 3.29]
    R.Y = fnorm_red<F>(fsub_nr<F, 3, 2>(MSX, fscale_nr(Y4x4, 2)));
    R.Z = mont_mul<F>(fscale_nr(P.Y, 2), P.Z);                // [2, 4] x [1, 2]
  } else {
    // a = -3, dbl-2001-b with 4beta and 2YZ taken as products:
    //   alpha = 3(X - Z^2)(X + Z^2), X3 = alpha^2 - 8beta, Y3 = alpha(4beta - X3) - 8Y^4, Z3 = 2YZ
    const fe delta = mont_sqr<F>(P.Z);
    const fe gamma = mont_sqr<F>(P.Y);
    const fe beta4 = mont_mul<F>(fscale_nr(P.X, 4), gamma);   // [4, 8] x [1, 1.13]
    const fe t = mont_mul<F>(fsub_nr<F, 2, 1>(P.X, delta), fadd_nr(P.X, delta));  // [3, 4] x [2, 3.13]
    const fe alpha = fnorm_red<F>(fscale_nr(t, 3));
    R.X = fnorm_red<F>(fsub_nr<F, 3, 2>(mont_sqr<F>(alpha), fscale_nr(beta4, 2)));
    const fe g4 = mont_sqr<F>(fscale_nr(gamma, 2));            // 4 gamma^2
    const fe aBX = mont_mul<F>(alpha, fsub_nr<F, 2, 1>(beta4, R.X));
    R.Y = fnorm_red<F>(fsub_nr<F, 3, 2>(aBX, fscale_nr(g4, 2)));
    R.Z = mont_mul<F>(fscale_nr(P.Y, 2), P.Z);
  }
  return R;
}

// add-2007-bl: 11M + 5S. exc is set when H == 0 (P == +-Q). Q.Y may be a lazy negation [2, 3].
template <class C>
DEV jac jadd(const jac& P, const jac& Q, bool& exc) {
  typedef typename C::Fp F;
  const fe Z1Z1 = mont_sqr<F>(P.Z);
  const fe Z2Z2 = mont_sqr<F>(Q.Z);
  const fe U1 = mont_mul<F>(P.X, Z2Z2);
  const fe U2 = mont_mul<F>(Q.X, Z1Z1);
  const fe S1 = mont_mul<F>(mont_mul<F>(P.Y, Q.Z), Z2Z2);
  const fe S2 = mont_mul<F>(mont_mul<F>(Q.Y, P.Z), Z1Z1);
  const fe H = fnorm_red<F>(fsub_nr<F, 2, 1>(U2, U1));
  exc = exc || fis_zero<F>(H);
  const fe I = mont_sqr<F>(fscale_nr(H, 2));
  const fe J = mont_mul<F>(H, I);
  const fe r = fnorm_red<F>(fscale_nr(fsub_nr<F, 2, 1>(S2, S1), 2));  // [6, 6.x] -> [1, 1+]
  const fe V = mont_mul<F>(U1, I);
  jac R;
  R.X = fnorm_red<F>(fsub_nr<F, 4, 3>(mont_sqr<F>(r), fadd_nr(J, fscale_nr(V, 2))));
  const fe rVX = mont_mul<F>(r, fsub_nr<F, 2, 1>(V, R.X));
  R.Y = fnorm_red<F>(fsub_nr<F, 2, 1>(rVX, mont_mul<F>(fscale_nr(S1, 2), J)));
  // ((Z1 + Z2)^2 - Z1Z1 - Z2Z2) H = 2 Z1 Z2 H
  const fe Zs = mont_sqr<F>(fadd_nr(P.Z, Q.Z));             // [2, 4]^2
  R.Z = mont_mul<F>(fsub_nr<F, 3, 2>(Zs, fadd_nr(Z1Z1, Z2Z2)), H);  // [4, 4.x] x [1, 1+]
  return R;
}

// madd-2007-bl (Q affine, y2 possibly a lazy negation [2, 3]): 8M + 3S
template <class C>
DEV jac jmadd(const jac& P, const fe& x2, const fe& y2, bool& exc) {
  typedef typename C::Fp F;
  const fe Z1Z1 = mont_sqr<F>(P.Z);
  const fe U2 = mont_mul<F>(x2, Z1Z1);
  const fe S2 = mont_mul<F>(mont_mul<F>(y2, P.Z), Z1Z1);
  const fe H = fnorm_red<F>(fsub_nr<F, 2, 1>(U2, P.X));
  exc = exc || fis_zero<F>(H);
  const fe I4 = fscale_nr(mont_sqr<F>(H), 4);               // I = 4 H^2: [4, 4.x]
  const fe J = mont_mul<F>(H, I4);
  const fe r = fnorm_red<F>(fscale_nr(fsub_nr<F, 2, 1>(S2, P.Y), 2));
  const fe V = mont_mul<F>(P.X, I4);
  jac R;
  R.X = fnorm_red<F>(fsub_nr<F, 4, 3>(mont_sqr<F>(r), fadd_nr(J, fscale_nr(V, 2))));
  const fe rVX = mont_mul<F>(r, fsub_nr<F, 2, 1>(V, R.X));
  R.Y = fnorm_red<F>(fsub_nr<F, 2, 1>(rVX, mont_mul<F>(fscale_nr(P.Y, 2), J)));
  R.Z = mont_mul<F>(P.Z, fscale_nr(H, 2));                  // (Z1 + H)^2 - Z1Z1 - HH = 2 Z1 H
  return R;
}

// -y as a lazy value [2, 3] when neg, y otherwise (no carries, no v_cndmask). y < 2M can have a
// top limb up to (2M >> 232), one above the top limb of 2M's LB = 1 bias, hence K = 3.
template <class F>
DEV fe fneg_if(bool neg, const fe& y) {
  return fsel(neg, fsub_nr<F, 3, 1>(fe_zero(), y), y);
}

// ---- complete (branching) group law for the careful path ----
struct jacf {
  jac P;
  bool inf;
};

template <class C>
DEV jacf jdbl_c(const jacf& P) {
  typedef typename C::Fp F;
  if (P.inf || fis_zero<F>(P.P.Y)) return jacf{P.P, true};
  return jacf{jdbl<C>(P.P), false};
}

template <class C>
DEV jacf jadd_c(const jacf& P, const jacf& Q) {
  typedef typename C::Fp F;
  if (P.inf) return Q;
  if (Q.inf) return P;
  const fe Z1Z1 = mont_sqr<F>(P.P.Z);
  const fe Z2Z2 = mont_sqr<F>(Q.P.Z);
  const fe U1 = mont_mul<F>(P.P.X, Z2Z2);
  const fe U2 = mont_mul<F>(Q.P.X, Z1Z1);
  const fe S1 = mont_mul<F>(mont_mul<F>(P.P.Y, Q.P.Z), Z2Z2);
  const fe S2 = mont_mul<F>(mont_mul<F>(Q.P.Y, P.P.Z), Z1Z1);
  if (feq<F>(U1, U2)) {
    if (feq<F>(S1, S2)) return jdbl_c<C>(P);
    return jacf{P.P, true};
  }
  bool exc = false;
  return jacf{jadd<C>(P.P, Q.P, exc), false};
}

// ---- integer words (little-endian u32) for the GLV split ----
// r = a * b mod 2^(32 NR); every step a[i]*b[j] + r + carry < 2^64
template <int NA, int NB, int NR>
DEV void wmul(const u32* a, const u32* b, u32* r) {
#pragma unroll
  for (int k = 0; k < NR; k++) r[k] = 0;
#pragma unroll
  for (int i = 0; i < NA; i++) {
    u64 c = 0;
#pragma unroll
    for (int j = 0; j < NB; j++) {
      if (i + j < NR) {
        const u64 t = (u64)a[i] * b[j] + r[i + j] + c;
        r[i + j] = (u32)t;
        c = t >> 32;
      }
    }
    if (i + NB < NR) r[i + NB] = (u32)c;
  }
}

DEV void wadd8(u32 a[8], const u32 b[8]) {  // a += b mod 2^256
  u64 c = 0;
#pragma unroll
  for (int i = 0; i < 8; i++) {
    const u64 t = (u64)a[i] + b[i] + c;
    a[i] = (u32)t;
    c = t >> 32;
  }
}

DEV void wsub8(u32 a[8], const u32 b[8]) {  // a -= b mod 2^256
  u32 t[8];
  words_sub(a, b, t);
#pragma unroll
  for (int i = 0; i < 8; i++) a[i] = t[i];
}

// GLV split for secp256k1 (constants and their derivation: gen_consts.py glv_k1). For k < n:
//   c1 = round(k g1 / 2^384) = round(k b2 / n), c2 = round(k |b1| / n)  (both < 2^128)
//   (k1, k2) = (k, 0) - c1 (a1, b1) - c2 (a2, b2), so k1 + k2 lambda = k (mod n), |k_i| < 2^128
//   then one lattice vector v1 = (a1, b1) (flips both parities), v2 = (a2, b2) (flips k2) or
//   v1 + v2 (flips k1) makes both halves odd: |k_i| < 2^129.
// Output: magnitudes h1, h2 (5 words) and signs, k = (-1)^n1 h1 + (-1)^n2 h2 lambda (mod n).
DEV void glv_split(const u32 k[8], u32 h1[5], u32 h2[5], bool& n1, bool& n2) {
  u32 p[16], c1[4], c2[4];
#pragma unroll
  for (int pass = 0; pass < 2; pass++) {
    wmul<8, 8, 16>(k, pass == 0 ? GLV_G1 : GLV_G2, p);
    u64 cy = p[11] >> 31;  // + 2^383: round to nearest
    u32* c = pass == 0 ? c1 : c2;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      cy += p[12 + i];
      c[i] = (u32)cy;
      cy >>= 32;
    }
  }
  u32 k1[8], k2[8], t[8];
#pragma unroll
  for (int i = 0; i < 8; i++) k1[i] = k[i];
  wmul<4, 8, 8>(c1, GLV_A1, t);
  wsub8(k1, t);
  wmul<4, 8, 8>(c2, GLV_A2, t);
  wsub8(k1, t);
  wmul<4, 8, 8>(c1, GLV_B1N, k2);
  wmul<4, 8, 8>(c2, GLV_B2, t);
  wsub8(k2, t);
  // parity correction: both even -> v1, k2 even -> v2, k1 even -> v1 + v2
  const bool e1 = (k1[0] & 1u) == 0, e2 = (k2[0] & 1u) == 0;
  const int sel = e1 ? (e2 ? 0 : 2) : 1;
  const u32 any = lane_mask(e1 || e2);
  u32 f1[8], f2[8];
#pragma unroll
  for (int i = 0; i < 8; i++) {
    u32 a = GLV_FIX_K1[i], b = GLV_FIX_K2[i];
    if (sel == 1) { a = GLV_FIX_K1[8 + i]; b = GLV_FIX_K2[8 + i]; }
    if (sel == 2) { a = GLV_FIX_K1[16 + i]; b = GLV_FIX_K2[16 + i]; }
    f1[i] = a & any;
    f2[i] = b & any;
  }
  wadd8(k1, f1);
  wadd8(k2, f2);
  n1 = (k1[7] >> 31) != 0;
  n2 = (k2[7] >> 31) != 0;
  u32 z[8] = {0, 0, 0, 0, 0, 0, 0, 0}, m1[8], m2[8];
  words_sub(z, k1, m1);
  words_sub(z, k2, m2);
#pragma unroll
  for (int i = 0; i < 5; i++) {
    h1[i] = n1 ? m1[i] : k1[i];
    h2[i] = n2 ? m2[i] : k2[i];
  }
}

// digit carrier of an odd h < 2^B over 5 words: C = (h >> 1) | 2^(B-1); the w-bit digits c_i of C
// give h = sum (2 c_i - (2^w - 1)) 2^(w i), every digit odd and the top one positive
template <int B>
DEV void glv_carrier(const u32 h[5], u32 c[5]) {
#pragma unroll
  for (int i = 0; i < 5; i++) c[i] = (h[i] >> 1) | (i < 4 ? h[i + 1] << 31 : 0u);
  c[(B - 1) >> 5] |= 1u << ((B - 1) & 31);
}

constexpr int GLV_BR = 132;  // R streams: 33 windows of 4 bits (halves < 2^129)
constexpr int GLV_BG = 136;  // G streams: 17 windows of 8 bits

// ---- per-item recovery ----
// sig_raw: the 64 signature bytes as 16 little-endian u32 loads; dig_raw: 8 words of the digest.
// Prologue results shared by the fast and careful ladders.
struct secp_pro {
  u32 status;
  fe rx, ry;         // R affine, Montgomery form of p; ry = -y when ry_neg
  bool ry_neg;
  u32 c1[8], c2[8];  // digit carriers C = (k' >> 1) | 2^255 for u1 (G) and u2 (R) (no GLV)
  bool g_neg;        // G digits must be negated (u1 replaced by n - u1) (no GLV)
  u32 cg1[5], cg2[5], cr1[5], cr2[5];  // GLV: carriers of the halves of u1 (G) and u2 (R)
  bool sg1, sg2, sr1, sr2;             // GLV: signs of those halves
  u32 u1[8], u2[8];  // plain u1, u2 (for the careful path)
};

// Q = u1 G + u2 P ladder set-up shared by recovery (P = R) and verification (P = the key):
// plain u1, u2 for the careful path, and the digit carriers of the fast ladder (GLV halves for
// secp256k1, odd representatives otherwise). u1m, u2m in Montgomery form of n; P affine (x, y).
template <class C>
DEV void secp_scalars(secp_pro& o, const fe& u1m, const fe& u2m, const fe& x, const fe& y) {
  typedef typename C::Fp Fp;
  typedef typename C::Fn Fn;
  const u32(&n)[8] = C::n32();
  fe_to_words(from_mont<Fn>(u1m), o.u1);
  fe_to_words(from_mont<Fn>(u2m), o.u2);
  o.rx = x;
  if constexpr (C::GLV) {
    u32 h1[5], h2[5];
    glv_split(o.u1, h1, h2, o.sg1, o.sg2);
    glv_carrier<GLV_BG>(h1, o.cg1);
    glv_carrier<GLV_BG>(h2, o.cg2);
    glv_split(o.u2, h1, h2, o.sr1, o.sr2);
    glv_carrier<GLV_BR>(h1, o.cr1);
    glv_carrier<GLV_BR>(h2, o.cr2);
    o.ry = y;
    o.ry_neg = false;
    o.g_neg = false;
    return;
  }
  // odd representatives
  o.g_neg = (o.u1[0] & 1u) == 0;
  const bool r_neg = (o.u2[0] & 1u) == 0;
  u32 k1[8], k2[8];
  words_sub(n, o.u1, k1);
  words_sub(n, o.u2, k2);
#pragma unroll
  for (int i = 0; i < 8; i++) {
    k1[i] = o.g_neg ? k1[i] : o.u1[i];
    k2[i] = r_neg ? k2[i] : o.u2[i];
  }
#pragma unroll
  for (int i = 0; i < 8; i++) {
    o.c1[i] = (k1[i] >> 1) | (i < 7 ? k1[i + 1] << 31 : 0x80000000u);
    o.c2[i] = (k2[i] >> 1) | (i < 7 ? k2[i + 1] << 31 : 0x80000000u);
  }
  o.ry = r_neg ? fneg<Fp>(y) : y;
  o.ry_neg = r_neg;
}

template <class C>
DEV void secp_prologue(const u32 sig_raw[16], const u32 dig_raw[8], secp_pro& o) {
  typedef typename C::Fp Fp;
  typedef typename C::Fn Fn;
  const u32(&n)[8] = C::n32();
  u32 r[8], s[8], z[8], sraw[8];
#pragma unroll
  for (int i = 0; i < 8; i++) sraw[i] = sig_raw[8 + i];
  const bool y_odd = (sraw[0] & 0x80u) != 0;  // first byte of s
  sraw[0] &= ~0x80u;
  words_from_be_bytes(sig_raw, r);
  words_from_be_bytes(sraw, s);
  words_from_be_bytes(dig_raw, z);
  o.status = ST_OK;
  if (words_is_zero(r) || !words_lt(r, n)) o.status = ST_BAD_R;
  else if (words_is_zero(s) || !words_lt(s, n)) o.status = ST_BAD_S;
  if (!words_lt(z, n)) {
    u32 t[8];
    words_sub(z, n, t);
#pragma unroll
    for (int i = 0; i < 8; i++) z[i] = t[i];
  }
  // lift x = r
  const fe x = to_mont<Fp>(fe_from_words(r));
  fe rhs = mont_mul<Fp>(mont_sqr<Fp>(x), x);
  if (C::A_IS_M3) rhs = fsub<Fp>(rhs, fadd<Fp>(fdbl<Fp>(x), x));
  rhs = fadd<Fp>(rhs, C::b());
  fe y = fpow_chain<Fp, typename C::SqrtChain>(rhs);
  if (o.status == ST_OK && !feq<Fp>(mont_sqr<Fp>(y), rhs)) o.status = ST_NO_POINT;
  const fe yc = from_mont<Fp>(y);
  if ((bool)(yc.v[0] & 1u) != y_odd) y = fneg<Fp>(y);
  // u1 = -z/r, u2 = s/r (mod n)
  const fe rn = to_mont<Fn>(fe_from_words(r));
  const fe rinv = finv<Fn>(rn);
  const fe u1m = fneg<Fn>(mont_mul<Fn>(to_mont<Fn>(fe_from_words(z)), rinv));
  const fe u2m = mont_mul<Fn>(to_mont<Fn>(fe_from_words(s)), rinv);
  secp_scalars<C>(o, u1m, u2m, x, y);
}

// ---- verification with a known public key (fuel Signature::verify) ----
// Rules of the reference's CPU stack (fuel-crypto 0.49 over libsecp256k1 for secp256k1, p256 /
// ecdsa 0.16 verify_prehash for secp256r1): r, s in [1, n-1] after clearing the recid bit (which
// verification ignores); secp256k1 also s <= (n-1)/2; the key X || Y with X, Y < p on the curve;
// z = digest mod n, u1 = z/s, u2 = r/s, P = u1 G + u2 Q != infinity, accept iff x(P) mod n == r.
// pub_raw: the 64 key bytes as 16 little-endian u32 loads. The same ladders then run with the key
// in place of R.
template <class C>
DEV void secp_prologue_verify(const u32 sig_raw[16], const u32 dig_raw[8], const u32 pub_raw[16], secp_pro& o) {
  typedef typename C::Fp Fp;
  typedef typename C::Fn Fn;
  const u32(&n)[8] = C::n32();
  u32 r[8], s[8], z[8], sraw[8], xw[8], yw[8];
#pragma unroll
  for (int i = 0; i < 8; i++) sraw[i] = sig_raw[8 + i];
  sraw[0] &= ~0x80u;
  words_from_be_bytes(sig_raw, r);
  words_from_be_bytes(sraw, s);
  words_from_be_bytes(dig_raw, z);
  words_from_be_bytes(pub_raw, xw);
  words_from_be_bytes(pub_raw + 8, yw);
  o.status = ST_OK;
  if (words_is_zero(r) || !words_lt(r, n)) o.status = ST_BAD_R;
  else if (words_is_zero(s) || !words_lt(s, n)) o.status = ST_BAD_S;
  else if (C::LOW_S && words_lt(C::nhalf32(), s)) o.status = ST_HIGH_S;
  else if (!words_lt(xw, C::p32()) || !words_lt(yw, C::p32())) o.status = ST_BAD_KEY;
  if (!words_lt(z, n)) {
    u32 t[8];
    words_sub(z, n, t);
#pragma unroll
    for (int i = 0; i < 8; i++) z[i] = t[i];
  }
  const fe x = to_mont<Fp>(fe_from_words(xw));
  const fe y = to_mont<Fp>(fe_from_words(yw));
  fe rhs = mont_mul<Fp>(mont_sqr<Fp>(x), x);
  if (C::A_IS_M3) rhs = fsub<Fp>(rhs, fadd<Fp>(fdbl<Fp>(x), x));
  rhs = fadd<Fp>(rhs, C::b());
  if (o.status == ST_OK && !feq<Fp>(mont_sqr<Fp>(y), rhs)) o.status = ST_BAD_KEY;
  // u1 = z/s, u2 = r/s (mod n)
  const fe sinv = finv<Fn>(to_mont<Fn>(fe_from_words(s)));
  const fe u1m = mont_mul<Fn>(to_mont<Fn>(fe_from_words(z)), sinv);
  const fe u2m = mont_mul<Fn>(to_mont<Fn>(fe_from_words(r)), sinv);
  secp_scalars<C>(o, u1m, u2m, x, y);
}

// the verifier's final check on P = (X : Y : Z) (Jacobian) without an inversion: x(P) < p < 2n,
// so x(P) mod n == r means X == r Z^2, or X == (r + n) Z^2 when r + n < p. sig_raw: the signature
// as loaded (r is decoded again here rather than kept live through the ladder).
template <class C>
DEV u32 secp_verify_x(const jac& P, const u32 sig_raw[16]) {
  typedef typename C::Fp F;
  u32 r[8], rn[8];
  words_from_be_bytes(sig_raw, r);
  words_add(r, C::n32(), rn);  // meaningful only when r < p - n
  const fe z2 = mont_sqr<F>(P.Z);
  const bool eq1 = feq<F>(P.X, mont_mul<F>(to_mont<F>(fe_from_words(r)), z2));
  const bool eq2 = feq<F>(P.X, mont_mul<F>(to_mont<F>(fe_from_words(rn)), z2));
  if (fis_zero<F>(P.Z)) return ST_INFINITY;
  return (eq1 || (eq2 && words_lt(r, C::pmn32()))) ? ST_OK : ST_MISMATCH;
}

DEV u32 nibble(const u32 c[8], int i) { return (c[i >> 3] >> ((i & 7) * 4)) & 15u; }
DEV u32 byte8(const u32 c[8], int j) { return (c[j >> 2] >> ((j & 3) * 8)) & 255u; }
// 9-bit digit j of a carrier (bits 9j .. 9j + 8, which may straddle two words)
DEV u32 digit9(const u32* c, int j) {
  const int b = 9 * j, w = b >> 5, sh = b & 31;
  u32 v = c[w] >> sh;
  if (sh > 23) v |= c[w + 1] << (32 - sh);
  return v & 511u;
}

// fixed-base windows of 9 bits: 256 odd multiples of G (LDS 1.75x larger, 4 fewer G additions per
// secp256k1 item and 3 per P-256 item). The carriers come from the 8-bit ones: a carrier is
// (h >> 1) | 2^(B-1) for the odd h it spells, so moving the marker bit re-spells h over B' bits.
constexpr int GLV_BG9 = 135;  // G streams of secp256k1: 15 windows of 9 bits
constexpr int R1_BG9 = 261;   // the G stream of P-256: 29 windows of 9 bits

// acc + the odd multiple of the G table's point that 9-bit digit d spells (negated per digit and
// stream sign); PHI: the entry read as phi(entry) = (beta x, y)
template <class C, bool PHI, class GTab>
DEV jac g_add9(const jac& acc, const GTab& gt, u32 d, bool neg_stream, bool& exc) {
  typedef typename C::Fp F;
  fe gx, gy;
  const u32 idx = ((d < 256) ? ~d : d) & 255u;
  if constexpr (PHI)
    gt.load_phi(idx, gx, gy);
  else
    gt.load(idx, gx, gy);
  return jmadd<C>(acc, gx, fneg_if<F>((d < 256) != neg_stream, gy), exc);
}

// GLV ladder with 9-bit G windows (see secp_ladder_glv): one doubling per iteration over bits
// 127..0; the R digits enter at multiples of 4, the G digits at multiples of 9
template <class C, class Tab, class GTab>
DEV jac secp_ladder_glv9(const secp_pro& o, Tab& tab, const GTab& gt, bool& exc) {
  typedef typename C::Fp F;
  u32 cg1[5], cg2[5];
#pragma unroll
  for (int i = 0; i < 5; i++) {
    cg1[i] = o.cg1[i];
    cg2[i] = o.cg2[i];
  }
  // marker bit GLV_BG - 1 = 135 -> GLV_BG9 - 1 = 134 (the halves are below 2^130)
  cg1[4] = (cg1[4] & ~(1u << 7)) | (1u << 6);
  cg2[4] = (cg2[4] & ~(1u << 7)) | (1u << 6);
  jac acc;
  {
    fe x, y;
    tab.load_affine(nibble(o.cr1, 32) - 8, x, y);
    acc.X = fnorm_red<F>(x);
    acc.Y = fnorm_red<F>(fneg_if<F>(o.sr1, y));
    acc.Z = fe_from_const(F::ONE);
  }
  {
    fe bx, ey;
    tab.load_affine_phi(nibble(o.cr2, 32) - 8, bx, ey);
    acc = jmadd<C>(acc, bx, fneg_if<F>(o.sr2, ey), exc);
  }
  fe x1, y1;
  u32 c1 = 0;
#pragma unroll 1
  for (int p = 127; p >= 0; p--) {
    if ((p & 3) == 3) {  // the window's first table entry, loaded ahead of its four doublings
      c1 = nibble(o.cr1, p >> 2);
      tab.load_affine((c1 < 8 ? ~c1 : c1) & 7u, x1, y1);
    }
    acc = jdbl<C>(acc);
    if ((p & 3) == 0) {
      const u32 c2 = nibble(o.cr2, p >> 2);
      fe x2, y2;
      tab.load_affine_phi((c2 < 8 ? ~c2 : c2) & 7u, x2, y2);  // x2 = beta x
      acc = jmadd<C>(acc, x1, fneg_if<F>((c1 < 8) != o.sr1, y1), exc);
      acc = jmadd<C>(acc, x2, fneg_if<F>((c2 < 8) != o.sr2, y2), exc);
    }
    if (p % 9 == 0) {
      acc = g_add9<C, false>(acc, gt, digit9(cg1, p / 9), o.sg1, exc);
      acc = g_add9<C, true>(acc, gt, digit9(cg2, p / 9), o.sg2, exc);
    }
  }
  return acc;
}

// P-256 ladder with 9-bit G windows (see secp_ladder): one doubling per iteration over bits
// 251..0; the R digits enter at multiples of 4, the G digits at multiples of 9 (the top one, at
// bit 252, with the top R digit)
template <class C, class Tab, class GTab>
DEV jac secp_ladder9(const secp_pro& o, Tab& tab, const GTab& gt, bool& exc) {
  typedef typename C::Fp F;
  u32 cg[9];
#pragma unroll
  for (int i = 0; i < 8; i++) cg[i] = o.c1[i];
  cg[7] &= 0x7fffffffu;  // marker bit 255 -> R1_BG9 - 1 = 260 (u1 < n < 2^256)
  cg[8] = 1u << (R1_BG9 - 1 - 256);
  jac acc;
  {
    fe x, y;
    tab.load_affine(nibble(o.c2, 63) - 8, x, y);
    acc.X = fnorm_red<F>(x);
    acc.Y = fnorm_red<F>(y);
    acc.Z = fe_from_const(F::ONE);
  }
  acc = g_add9<C, false>(acc, gt, digit9(cg, 28), o.g_neg, exc);
  fe ex, ey;
  u32 c = 0;
#pragma unroll 1
  for (int p = 251; p >= 0; p--) {
    if ((p & 3) == 3) {
      c = nibble(o.c2, p >> 2);
      tab.load_affine((c < 8 ? ~c : c) & 7u, ex, ey);
    }
    acc = jdbl<C>(acc);
    if ((p & 3) == 0) acc = jmadd<C>(acc, ex, fneg_if<F>(c < 8, ey), exc);
    if (p % 9 == 0) acc = g_add9<C, false>(acc, gt, digit9(cg, p / 9), o.g_neg, exc);
  }
  return acc;
}

// GLV ladder (secp256k1): Q = h1 G + h2 phi(G) + h3 R + h4 phi(R) on one doubling chain of 132
// bits, with phi(x, y) = (beta x, y): both tables store beta x beside x (Tab::load_affine_phi,
// GTab::load_phi read (beta x, y)). R streams use 33
// windows of 4 bits over the per-lane table of odd multiples of R; G streams 17 windows of 8 bits
// over the table of odd multiples of G, added at every second R window.
template <class C, class Tab, class GTab>
DEV jac secp_ladder_glv(const secp_pro& o, Tab& tab, const GTab& gt, bool& exc) {
  typedef typename C::Fp F;
  // top windows: every stream's top digit is positive; only the stream signs apply
  // the accumulator enters jmadd as a subtrahend, so its coordinates are folded to < (1 + 2^-19) M
  jac acc;
  {
    fe x, y;
    tab.load_affine(nibble(o.cr1, 32) - 8, x, y);
    acc.X = fnorm_red<F>(x);
    acc.Y = fnorm_red<F>(fneg_if<F>(o.sr1, y));
    acc.Z = fe_from_const(F::ONE);
  }
  {
    fe bx, ey;
    tab.load_affine_phi(nibble(o.cr2, 32) - 8, bx, ey);
    acc = jmadd<C>(acc, bx, fneg_if<F>(o.sr2, ey), exc);
  }
#pragma unroll 1
  for (int i = 32; i >= 0; i--) {
    if (i < 32) {
      // the first table load is issued before the doublings and the second before the first
      // addition; each is hidden behind that work without keeping both entries live
      const u32 c1 = nibble(o.cr1, i), c2 = nibble(o.cr2, i);
      fe x1, y1, x2, y2;
      tab.load_affine((c1 < 8 ? ~c1 : c1) & 7u, x1, y1);
#pragma unroll 1
      for (int d = 0; d < 4; d++) acc = jdbl<C>(acc);
      tab.load_affine_phi((c2 < 8 ? ~c2 : c2) & 7u, x2, y2);  // x2 = beta x
      acc = jmadd<C>(acc, x1, fneg_if<F>((c1 < 8) != o.sr1, y1), exc);
      acc = jmadd<C>(acc, x2, fneg_if<F>((c2 < 8) != o.sr2, y2), exc);
    }
    if ((i & 1) == 0) {
      fe gx, gy;
      const u32 b1 = byte8(o.cg1, i >> 1), b2 = byte8(o.cg2, i >> 1);
      gt.load(((b1 < 128) ? ~b1 : b1) & 127u, gx, gy);
      gy = fneg_if<F>((b1 < 128) != o.sg1, gy);
      acc = jmadd<C>(acc, gx, gy, exc);
      gt.load_phi(((b2 < 128) ? ~b2 : b2) & 127u, gx, gy);  // gx = beta x
      gy = fneg_if<F>((b2 < 128) != o.sg2, gy);
      acc = jmadd<C>(acc, gx, gy, exc);
    }
  }
  return acc;
}

// Per-lane table of the odd multiples (2e + 1) P, e = 0..7, made affine with one inversion
// (Montgomery's simultaneous inversion), so that the ladder adds its entries with the mixed formula
// (8M + 3S instead of 11M + 5S per addition). The entries are built in Jacobian form with the
// running product of the Z coordinates before them stored alongside; one inversion of the full
// product then yields every 1/Z on the way back down, and each entry is rewritten in place as
// (x, y). Tab: store_full(e, jac, before) / load_full(e, jac&, before&) and store_affine(e, x, y)
// / load_affine(e, x&, y&).
// The entries are built on an isomorphic curve where 2P is affine, so the seven additions are
// mixed ones too: with 2P = (X2, Y2, Z2), the map (x, y) -> (Z2^2 x, Z2^3 y) takes the curve to
// y^2 = x^3 + a Z2^4 x + b Z2^6, 2P to the affine (X2, Y2) and P to (Z2^2 x, Z2^3 y). The
// addition formulas do not involve a or b. A Jacobian (X, Y, Z) there is (X, Y, Z Z2) on the
// curve, so Z2 joins the product that is inverted.
template <class C, class Tab>
DEV void secp_build_table(const secp_pro& o, Tab& tab, bool& exc) {
  typedef typename C::Fp F;
  jac P;
  P.X = o.rx;
  P.Y = o.ry;
  P.Z = fe_from_const(F::ONE);
  const jac P2 = jdbl<C>(P);
  // secp256k1 (GLV): beta x beside each entry, for the phi(R) stream
  const fe beta = C::GLV ? fe_from_const(K1_BETA) : fe_zero();
  if constexpr (C::GLV)
    tab.store_affine_phi(0, P.X, P.Y, mont_mul<F>(P.X, beta));
  else
    tab.store_affine(0, P.X, P.Y);
  // P on the isomorphic curve; the accumulator enters jmadd as a subtrahend, so it is folded
  const fe zz = mont_sqr<F>(P2.Z);
  jac T;
  T.X = fnorm_red<F>(mont_mul<F>(P.X, zz));
  T.Y = fnorm_red<F>(mont_mul<F>(P.Y, mont_mul<F>(zz, P2.Z)));
  T.Z = P.Z;
  fe before = P.Z;  // Z_1 * ... * Z_(e-1), isomorphic-curve Z coordinates
#pragma unroll 1
  for (int e = 1; e < 8; e++) {
    T = jmadd<C>(T, P2.X, P2.Y, exc);
    tab.store_full(e, T, before);
    before = mont_mul<F>(before, T.Z);
  }
  fe inv = finv<F>(mont_mul<F>(before, P2.Z));  // 1 / (Z2 * Z_1 * ... * Z_7)
#pragma unroll 1
  for (int e = 7; e >= 1; e--) {
    jac E;
    fe b;
    tab.load_full(e, E, b);
    const fe zi = mont_mul<F>(inv, b);  // 1 / (Z2 * Z_e), the curve's 1 / Z
    inv = mont_mul<F>(inv, E.Z);        // 1 / (Z2 * Z_1 * ... * Z_(e-1))
    const fe zi2 = mont_sqr<F>(zi);
    const fe ax = mont_mul<F>(E.X, zi2), ay = mont_mul<F>(E.Y, mont_mul<F>(zi2, zi));
    if constexpr (C::GLV)
      tab.store_affine_phi(e, ax, ay, mont_mul<F>(ax, beta));
    else
      tab.store_affine(e, ax, ay);
  }
}

// R-table stored per lane (secp_build_table). GTab provides load(idx, fe& x, fe& y) for the odd
// multiples (2 idx + 1) G: 128 of them for fixed-base windows of WGB = 8 bits, 256 for 9.
template <class C, int WGB = 8, class Tab, class GTab>
DEV jac secp_ladder(const secp_pro& o, Tab& tab, const GTab& gt, bool& exc) {
  typedef typename C::Fp F;
  static_assert(WGB == 8 || WGB == 9, "fixed-base windows of 8 or 9 bits");
  secp_build_table<C>(o, tab, exc);
  if constexpr (WGB == 9) {
    if constexpr (C::GLV)
      return secp_ladder_glv9<C>(o, tab, gt, exc);
    else
      return secp_ladder9<C>(o, tab, gt, exc);
  }
  if constexpr (C::GLV) return secp_ladder_glv<C>(o, tab, gt, exc);
  // top R digit is always positive (C's bit 255 is set); the accumulator enters jmadd as a
  // subtrahend, so its coordinates are folded to < (1 + 2^-19) M
  jac acc;
  {
    fe x, y;
    tab.load_affine(nibble(o.c2, 63) - 8, x, y);
    acc.X = fnorm_red<F>(x);
    acc.Y = fnorm_red<F>(y);
    acc.Z = fe_from_const(F::ONE);
  }
#pragma unroll 1
  for (int i = 62; i >= 0; i--) {
    // issue the HBM load of this window's table entry before the doublings, which hide its latency
    const u32 c = nibble(o.c2, i);
    const bool neg = c < 8;
    fe ex, ey;
    tab.load_affine((neg ? ~c : c) & 7u, ex, ey);
#pragma unroll 1
    for (int d = 0; d < 4; d++) acc = jdbl<C>(acc);
    acc = jmadd<C>(acc, ex, fneg_if<F>(neg, ey), exc);
    if ((i & 1) == 0) {
      const u32 cg = byte8(o.c1, i >> 1);
      const bool gneg = (cg < 128) != o.g_neg;
      fe gx, gy;
      gt.load(((cg < 128) ? ~cg : cg) & 127u, gx, gy);
      gy = fneg_if<F>(gneg, gy);
      acc = jmadd<C>(acc, gx, gy, exc);
    }
  }
  return acc;
}

// careful path: Q = u1*G + u2*R by double-and-add with the complete law (slow, branching)
template <class C, class GTab>
DEV jacf secp_ladder_careful(const secp_pro& o, const GTab& gt) {
  typedef typename C::Fp F;
  jacf G, R, GR, acc;
  fe gx, gy;
  gt.load(0, gx, gy);
  G.P.X = gx; G.P.Y = gy; G.P.Z = fe_from_const(F::ONE); G.inf = false;
  // o.u2 is the true u2; undo the negation of o.ry that came with the fast ladder's n - u2
  R.P.X = o.rx; R.P.Y = o.ry_neg ? fneg<F>(o.ry) : o.ry; R.P.Z = fe_from_const(F::ONE); R.inf = false;
  GR = jadd_c<C>(G, R);
  acc.inf = true;
  acc.P = G.P;
  for (int i = 255; i >= 0; i--) {
    acc = jdbl_c<C>(acc);
    const bool b1 = (o.u1[i >> 5] >> (i & 31)) & 1u;
    const bool b2 = (o.u2[i >> 5] >> (i & 31)) & 1u;
    if (b1 && b2) acc = jadd_c<C>(acc, GR);
    else if (b1) acc = jadd_c<C>(acc, G);
    else if (b2) acc = jadd_c<C>(acc, R);
  }
  return acc;
}

// Jacobian -> affine big-endian X||Y (16 words as stored to memory)
template <class C>
DEV void secp_output(const jac& Q, u32 out_raw[16]) {
  typedef typename C::Fp F;
  const fe zi = finv<F>(Q.Z);
  const fe zi2 = mont_sqr<F>(zi);
  const fe x = from_mont<F>(mont_mul<F>(Q.X, zi2));
  const fe y = from_mont<F>(mont_mul<F>(Q.Y, mont_mul<F>(zi2, zi)));
  u32 w[8];
  fe_to_words(x, w);
  words_to_be_bytes(w, out_raw);
  fe_to_words(y, w);
  words_to_be_bytes(w, out_raw + 8);
}

}  // namespace sigops
