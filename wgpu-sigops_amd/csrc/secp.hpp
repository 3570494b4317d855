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
//   * u2*R: w = 4, table of odd multiples R..15R (Jacobian) in per-lane memory;
//   * u1*G: w = 8, table of odd multiples G..255G (affine) staged in LDS by the kernel;
//   * Jacobian formulas dbl-2009-l (a = 0) / dbl-2001-b (a = -3), add-2007-bl, madd-2007-bl.
//     These are incomplete: an addition of P and +-P is detected (H == 0) and the lane is marked
//     for the careful path, which redoes the item with complete, branching formulas.
#pragma once
#include "field.hpp"

namespace sigops {

enum : u32 {
  ST_OK = 0,
  ST_BAD_R = 1,
  ST_BAD_S = 2,
  ST_NO_POINT = 3,
  ST_INFINITY = 4,
  ST_PENDING = 0xFF,  // exceptional addition met: the careful path finishes the item
};

struct CurveK1 {
  typedef ModPk1 Fp;
  typedef ModNk1 Fn;
  static constexpr bool A_IS_M3 = false;
  DEV static fe b() { return fe_from_const(K1_B); }
  DEV static const u32 (&n32())[8] { return K1_N32; }
  DEV static const u32 (&sqrt_exp())[8] { return K1_SQRT_EXP; }
};
struct CurveR1 {
  typedef ModPr1 Fp;
  typedef ModNr1 Fn;
  static constexpr bool A_IS_M3 = true;
  DEV static fe b() { return fe_from_const(R1_B); }
  DEV static const u32 (&n32())[8] { return R1_N32; }
  DEV static const u32 (&sqrt_exp())[8] { return R1_SQRT_EXP; }
};

struct jac {
  fe X, Y, Z;
};

template <class C>
DEV jac jdbl(const jac& P) {
  typedef typename C::Fp F;
  jac R;
  if (!C::A_IS_M3) {
    // dbl-2009-l: 2M + 5S
    const fe A = mont_sqr<F>(P.X);
    const fe B = mont_sqr<F>(P.Y);
    const fe Cc = mont_sqr<F>(B);
    fe D = fsub<F>(fsub<F>(mont_sqr<F>(fadd<F>(P.X, B)), A), Cc);
    D = fdbl<F>(D);
    const fe E = fadd<F>(fdbl<F>(A), A);
    const fe Fv = mont_sqr<F>(E);
    R.X = fsub<F>(Fv, fdbl<F>(D));
    const fe C8 = fdbl<F>(fdbl<F>(fdbl<F>(Cc)));
    R.Y = fsub<F>(mont_mul<F>(E, fsub<F>(D, R.X)), C8);
    R.Z = fdbl<F>(mont_mul<F>(P.Y, P.Z));
  } else {
    // dbl-2001-b (a = -3): 3M + 5S
    const fe delta = mont_sqr<F>(P.Z);
    const fe gamma = mont_sqr<F>(P.Y);
    const fe beta = mont_mul<F>(P.X, gamma);
    const fe t = mont_mul<F>(fsub<F>(P.X, delta), fadd<F>(P.X, delta));
    const fe alpha = fadd<F>(fdbl<F>(t), t);
    const fe beta4 = fdbl<F>(fdbl<F>(beta));
    R.X = fsub<F>(mont_sqr<F>(alpha), fdbl<F>(beta4));
    R.Z = fsub<F>(fsub<F>(mont_sqr<F>(fadd<F>(P.Y, P.Z)), gamma), delta);
    const fe g2 = mont_sqr<F>(gamma);
    const fe g8 = fdbl<F>(fdbl<F>(fdbl<F>(g2)));
    R.Y = fsub<F>(mont_mul<F>(alpha, fsub<F>(beta4, R.X)), g8);
  }
  return R;
}

// add-2007-bl: 11M + 5S. exc is set when H == 0 (P == +-Q).
template <class C>
DEV jac jadd(const jac& P, const jac& Q, bool& exc) {
  typedef typename C::Fp F;
  const fe Z1Z1 = mont_sqr<F>(P.Z);
  const fe Z2Z2 = mont_sqr<F>(Q.Z);
  const fe U1 = mont_mul<F>(P.X, Z2Z2);
  const fe U2 = mont_mul<F>(Q.X, Z1Z1);
  const fe S1 = mont_mul<F>(mont_mul<F>(P.Y, Q.Z), Z2Z2);
  const fe S2 = mont_mul<F>(mont_mul<F>(Q.Y, P.Z), Z1Z1);
  const fe H = fsub<F>(U2, U1);
  exc = exc || fis_zero<F>(H);
  const fe I = mont_sqr<F>(fdbl<F>(H));
  const fe J = mont_mul<F>(H, I);
  const fe r = fdbl<F>(fsub<F>(S2, S1));
  const fe V = mont_mul<F>(U1, I);
  jac R;
  R.X = fsub<F>(fsub<F>(mont_sqr<F>(r), J), fdbl<F>(V));
  R.Y = fsub<F>(mont_mul<F>(r, fsub<F>(V, R.X)), fdbl<F>(mont_mul<F>(S1, J)));
  R.Z = mont_mul<F>(fsub<F>(fsub<F>(mont_sqr<F>(fadd<F>(P.Z, Q.Z)), Z1Z1), Z2Z2), H);
  return R;
}

// madd-2007-bl (Q affine): 7M + 4S
template <class C>
DEV jac jmadd(const jac& P, const fe& x2, const fe& y2, bool& exc) {
  typedef typename C::Fp F;
  const fe Z1Z1 = mont_sqr<F>(P.Z);
  const fe U2 = mont_mul<F>(x2, Z1Z1);
  const fe S2 = mont_mul<F>(mont_mul<F>(y2, P.Z), Z1Z1);
  const fe H = fsub<F>(U2, P.X);
  exc = exc || fis_zero<F>(H);
  const fe HH = mont_sqr<F>(H);
  const fe I = fdbl<F>(fdbl<F>(HH));
  const fe J = mont_mul<F>(H, I);
  const fe r = fdbl<F>(fsub<F>(S2, P.Y));
  const fe V = mont_mul<F>(P.X, I);
  jac R;
  R.X = fsub<F>(fsub<F>(mont_sqr<F>(r), J), fdbl<F>(V));
  R.Y = fsub<F>(mont_mul<F>(r, fsub<F>(V, R.X)), fdbl<F>(mont_mul<F>(P.Y, J)));
  R.Z = fsub<F>(fsub<F>(mont_sqr<F>(fadd<F>(P.Z, H)), Z1Z1), HH);
  return R;
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

// ---- per-item recovery ----
// sig_raw: the 64 signature bytes as 16 little-endian u32 loads; dig_raw: 8 words of the digest.
// Prologue results shared by the fast and careful ladders.
struct secp_pro {
  u32 status;
  fe rx, ry;       // R affine, Montgomery form of p (ry already negated when u2 was negated)
  u32 c1[8], c2[8];  // digit carriers C = (k' >> 1) | 2^255 for u1 (G) and u2 (R)
  bool g_neg;      // G digits must be negated (u1 replaced by n - u1)
  u32 u1[8], u2[8];  // plain u1, u2 (for the careful path)
};

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
  fe y = fpow_const<Fp>(rhs, C::sqrt_exp());
  if (o.status == ST_OK && !feq<Fp>(mont_sqr<Fp>(y), rhs)) o.status = ST_NO_POINT;
  const fe yc = from_mont<Fp>(y);
  if ((bool)(yc.v[0] & 1u) != y_odd) y = fneg<Fp>(y);
  // u1 = -z/r, u2 = s/r (mod n)
  const fe rn = to_mont<Fn>(fe_from_words(r));
  const fe rinv = finv<Fn>(rn);
  const fe u1m = fneg<Fn>(mont_mul<Fn>(to_mont<Fn>(fe_from_words(z)), rinv));
  const fe u2m = mont_mul<Fn>(to_mont<Fn>(fe_from_words(s)), rinv);
  fe_to_words(from_mont<Fn>(u1m), o.u1);
  fe_to_words(from_mont<Fn>(u2m), o.u2);
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
  o.rx = x;
  o.ry = r_neg ? fneg<Fp>(y) : y;
}

DEV u32 nibble(const u32 c[8], int i) { return (c[i >> 3] >> ((i & 7) * 4)) & 15u; }
DEV u32 byte8(const u32 c[8], int j) { return (c[j >> 2] >> ((j & 3) * 8)) & 255u; }

// R-table stored per lane; Tab provides store(e, jac) / load(e) -> jac.
// GTab provides load(idx, fe& x, fe& y) for the odd multiples (2 idx + 1) G.
template <class C, class Tab, class GTab>
DEV jac secp_ladder(const secp_pro& o, Tab& tab, const GTab& gt, bool& exc) {
  typedef typename C::Fp F;
  jac P;
  P.X = o.rx;
  P.Y = o.ry;
  P.Z = fe_from_const(F::ONE);
  const jac P2 = jdbl<C>(P);
  tab.store(0, P);
  jac T = P;
#pragma unroll 1
  for (int e = 1; e < 8; e++) {
    T = jadd<C>(T, P2, exc);
    tab.store(e, T);
  }
  // top R digit is always positive (C's bit 255 is set)
  jac acc = tab.load(nibble(o.c2, 63) - 8);
#pragma unroll 1
  for (int i = 62; i >= 0; i--) {
    // issue the HBM load of this window's table entry before the doublings, which hide its latency
    const u32 c = nibble(o.c2, i);
    const bool neg = c < 8;
    jac E = tab.load((neg ? ~c : c) & 7u);
#pragma unroll 1
    for (int d = 0; d < 4; d++) acc = jdbl<C>(acc);
    if (neg) E.Y = fneg<F>(E.Y);
    acc = jadd<C>(acc, E, exc);
    if ((i & 1) == 0) {
      const u32 cg = byte8(o.c1, i >> 1);
      const bool gneg = (cg < 128) != o.g_neg;
      fe gx, gy;
      gt.load(((cg < 128) ? ~cg : cg) & 127u, gx, gy);
      if (gneg) gy = fneg<F>(gy);
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
  // o.ry may have been negated together with u2 -> use the original sign consistently:
  // o.u2 is the true u2 and o.ry = r_neg ? -y : y; undo the negation here.
  const bool r_neg = (o.u2[0] & 1u) == 0;
  R.P.X = o.rx; R.P.Y = r_neg ? fneg<F>(o.ry) : o.ry; R.P.Z = fe_from_const(F::ONE); R.inf = false;
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
