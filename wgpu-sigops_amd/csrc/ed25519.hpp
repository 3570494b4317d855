// ed25519.hpp -- Ed25519 verification with ed25519-dalek `VerifyingKey::verify` semantics, one
// signature per lane.
//
// Reference flow: src/wgsl/ed25519_eddsa.wgsl:6-63 (-A, [s]B + [k](-A)),
// src/wgsl/ed25519_utils.wgsl:27-118 (pow_p58, sqrt_ratio_i, decompression),
// src/wgsl/main/ed25519_eddsa_main_0.wgsl:23-86 (decoding, A.y mod p, preimage) and _5.wgsl:20-46
// (compress and byte-compare with R). Acceptance rules follow dalek (the reference test oracle,
// src/tests/ed25519_eddsa.rs:26) rather than the WGSL, which skips the s < l check (src/wgsl/main/ed25519_eddsa_main_0.wgsl:51-54):
//   s >= l -> reject; A must decompress (y taken mod p, sign bit applied also to x = 0);
//   k = SHA-512(R || A || M) mod l over the given bytes; accept iff compress([s]B - [k]A) == R.
//
// Group law: extended twisted Edwards (a = -1) with the unified formulas of Hisil-Wong-Carter-
// Dawson 2008 in the completed-point form used by curve25519-dalek; they are complete on the
// whole curve (d is a non-square), so torsion components of A need no special handling and there
// is no exceptional path. Scalar multiplication shares one doubling chain (Strauss-Shamir):
//   [s]B: w = 8 signed odd digits over an LDS-staged table of (2j+1)B in affine Niels form;
//   [k](-A): w = 4 signed odd digits over a per-lane table of odd multiples in cached form.
// The kernels that verify only Ed25519 run the lattice-reduced form of the final check instead
// (ed_check_lattice: a 136-bit chain, same results); the mixed-scheme kernels keep the full one.
#pragma once
#include "secp.hpp"  // nibble / byte8 helpers and status codes
#include "sha512.hpp"

namespace sigops {

typedef ModPed Fe;
typedef ModL Fl;

struct ext {
  fe X, Y, Z, T;
};
struct cached {  // (Y+X, Y-X, Z, 2d*T)
  fe ypx, ymx, z, t2d;
};

DEV ext ext_identity() {
  ext r;
  r.X = fe_zero();
  r.Y = fe_from_const(Fe::ONE);
  r.Z = r.Y;
  r.T = fe_zero();
  return r;
}

// Group law on lazy field arithmetic (field.hpp). Inputs: coordinates normalised with value < 2M;
// cached ypx / ymx normalised, z and t2d products (t2d possibly a lazy negation [2, 2]).
// Bounds of intermediates as [limb bound in units of 2^29, value bound in units of M]; a product
// needs La*Lb < 7.1 and Va*Vb < 64 (M = 2^255 - 19, R = 2^261).
DEV cached to_cached(const ext& P) {
  cached c;
  c.ypx = fnorm_red<Fe>(fadd_nr(P.Y, P.X));
  c.ymx = fnorm_red<Fe>(fsub_nr<Fe, 3, 1>(P.Y, P.X));  // P.X may be any value < 2M
  c.z = P.Z;
  c.t2d = mont_mul<Fe>(P.T, fe_from_const(ED_D2));
  return c;
}

// -Q when neg: swap Y+X and Y-X and negate 2dT (lazily)
DEV cached neg_cached_if(bool neg, const cached& c) {
  const u32 m = lane_mask(neg);
  cached r;
  r.ypx = fsel_mask(m, c.ymx, c.ypx);
  r.ymx = fsel_mask(m, c.ypx, c.ymx);
  r.z = c.z;
  r.t2d = fsel_mask(m, fsub_nr<Fe, 2, 1>(fe_zero(), c.t2d), c.t2d);
  return r;
}

// completed point (E:G:H:Tc) with x = E/Tc, y = G/H -> extended; each pair product within bounds
DEV ext completed_to_ext(const fe& E, const fe& G, const fe& H, const fe& Tc, bool want_t) {
  ext r;
  r.X = mont_mul<Fe>(E, Tc);
  r.Y = mont_mul<Fe>(G, H);
  r.Z = mont_mul<Fe>(H, Tc);
  r.T = want_t ? mont_mul<Fe>(E, G) : fe_zero();
  return r;
}

// doubling (a = -1); T of the input is not used
DEV ext edbl(const ext& P, bool want_t) {
  const fe XX = mont_sqr<Fe>(P.X);
  const fe YY = mont_sqr<Fe>(P.Y);
  const fe ZZ = mont_sqr<Fe>(P.Z);
  const fe XpY2 = mont_sqr<Fe>(fadd_nr(P.X, P.Y));           // [2, 4]^2
  const fe G = fadd_nr(YY, XX);                               // YY + XX: [2, 2.2]
  const fe H = fsub_nr<Fe, 2, 1>(YY, XX);                     // YY - XX: [3, 3.1]
  const fe E = fnorm_red<Fe>(fsub_nr<Fe, 3, 2>(XpY2, G));     // 2XY
  const fe Tc = fnorm_red<Fe>(fsub_nr<Fe, 2, 1>(fadd_nr(fscale_nr(ZZ, 2), XX), YY));  // 2ZZ - (YY - XX)
  return completed_to_ext(E, G, H, Tc, want_t);               // [1]x[1], [2]x[3], [3]x[1], [1]x[2]
}

DEV ext eadd_cached(const ext& P, const cached& Q, bool want_t) {
  const fe PP = mont_mul<Fe>(fadd_nr(P.Y, P.X), Q.ypx);
  const fe MM = mont_mul<Fe>(fsub_nr<Fe, 3, 1>(P.Y, P.X), Q.ymx);  // P.X: any value < 2M
  const fe TT2d = mont_mul<Fe>(P.T, Q.t2d);
  const fe ZZ2 = mont_mul<Fe>(fscale_nr(P.Z, 2), Q.z);
  const fe E = fnorm_red<Fe>(fsub_nr<Fe, 2, 1>(PP, MM));
  const fe G = fadd_nr(PP, MM);                               // [2, 2.x]
  const fe H = fadd_nr(ZZ2, TT2d);                            // [2, 2.x]
  const fe Tc = fsub_nr<Fe, 2, 1>(ZZ2, TT2d);                 // [3, 3.x]
  return completed_to_ext(E, G, H, Tc, want_t);               // [1]x[3], [2]x[2], [2]x[3], [1]x[2]
}

// affine Niels (y+x, y-x, 2d*x*y); t2d possibly a lazy negation
DEV ext eadd_niels(const ext& P, const fe& ypx, const fe& ymx, const fe& t2d, bool want_t) {
  const fe PP = mont_mul<Fe>(fadd_nr(P.Y, P.X), ypx);
  const fe MM = mont_mul<Fe>(fsub_nr<Fe, 3, 1>(P.Y, P.X), ymx);
  const fe TT2d = mont_mul<Fe>(P.T, t2d);
  const fe ZZ2 = fscale_nr(P.Z, 2);                           // [2, 4]
  const fe E = fnorm_red<Fe>(fsub_nr<Fe, 2, 1>(PP, MM));
  const fe G = fadd_nr(PP, MM);                               // [2, 2.x]
  const fe H = fadd_nr(ZZ2, TT2d);                            // [3, 5.x]
  const fe Tc = fnorm_red<Fe>(fsub_nr<Fe, 2, 1>(ZZ2, TT2d));
  return completed_to_ext(E, G, H, Tc, want_t);               // [1]x[1], [2]x[3], [3]x[1], [1]x[2]
}

// dalek FieldElement::sqrt_ratio_i (field.rs): returns was_nonzero_square, r = nonnegative root
DEV bool sqrt_ratio_i(const fe& u, const fe& v, fe& r_out) {
  const fe v3 = mont_mul<Fe>(mont_sqr<Fe>(v), v);
  const fe v7 = mont_mul<Fe>(mont_sqr<Fe>(v3), v);
  fe r = mont_mul<Fe>(mont_mul<Fe>(u, v3), fpow_chain<Fe, CH_ED_P58>(mont_mul<Fe>(u, v7)));
  const fe check = mont_mul<Fe>(v, mont_sqr<Fe>(r));
  const fe i = fe_from_const(ED_SQRTM1);
  const fe neg_u = fneg<Fe>(u);
  const bool correct = feq<Fe>(check, u);
  const bool flipped = feq<Fe>(check, neg_u);
  const bool flipped_i = feq<Fe>(check, mont_mul<Fe>(neg_u, i));
  const fe r_prime = mont_mul<Fe>(i, r);
  r = fsel(flipped || flipped_i, r_prime, r);
  const bool r_neg = (from_mont<Fe>(r).v[0] & 1u) != 0;
  r = fsel(r_neg, fneg<Fe>(r), r);
  r_out = r;
  return correct || flipped;
}

// ED_BAD_R and ED_SMALL_ORDER come from verify_strict only
enum : u32 { ED_OK = 0, ED_BAD_S = 1, ED_BAD_KEY = 2, ED_MISMATCH = 3, ED_BAD_R = 4, ED_SMALL_ORDER = 5 };

// dalek CompressedEdwardsY::decompress: y = the 255 low bits taken mod p (non-canonical y is
// accepted), x = sqrt_ratio_i(y^2 - 1, d y^2 + 1), negated when the sign bit is set (also for
// x = 0). enc: 8 raw little-endian words. Returns false if there is no such x.
DEV bool ed_decompress(const u32 enc[8], fe& x, fe& y) {
  u32 yw[8];
#pragma unroll
  for (int i = 0; i < 8; i++) yw[i] = enc[i];
  const bool x_sign = (yw[7] >> 31) != 0;
  yw[7] &= 0x7fffffffu;
  if (!words_lt(yw, ED_P32)) {
    u32 t[8];
    words_sub(yw, ED_P32, t);
#pragma unroll
    for (int i = 0; i < 8; i++) yw[i] = t[i];
  }
  y = to_mont<Fe>(fe_from_words(yw));
  const fe one = fe_from_const(Fe::ONE);
  const fe yy = mont_sqr<Fe>(y);
  const fe u = fsub<Fe>(yy, one);
  const fe v = fadd<Fe>(mont_mul<Fe>(yy, fe_from_const(ED_D)), one);
  const bool ok = sqrt_ratio_i(u, v, x);
  if (x_sign) x = fneg<Fe>(x);
  return ok;
}

// dalek EdwardsPoint::is_small_order: [8]P is the identity (X = 0 and Y = Z after three doublings)
DEV bool ed_small_order(const fe& x, const fe& y) {
  ext P;
  P.X = x;
  P.Y = y;
  P.Z = fe_from_const(Fe::ONE);
  P.T = fe_zero();  // doubling does not read T
#pragma unroll 1
  for (int i = 0; i < 3; i++) P = edbl(P, false);
  return fis_zero<Fe>(P.X) && feq<Fe>(P.Y, P.Z);
}

// k = the 512-bit little-endian digest h (16 words) mod l, as 8 little-endian words
DEV void ed25519_k_from_digest(const u32 h[16], u32 k[8]) {
  const fe hi = mont_mul<Fl>(fe_from_words(h + 8), fe_from_const(L_C256));
  const fe lo = mont_mul<Fl>(fe_from_words(h), fe_from_const(L_C0));
  fe_to_words(fcanon<Fl>(fadd<Fl>(hi, lo)), k);
}

// sig_raw: 16 raw words (R || s); pk_raw: 8 raw words; message in mem at msg_off, len bytes.
// k = SHA-512(R || A || M) mod l as 8 little-endian words
template <class Mem>
DEV void ed25519_challenge(const u32 sig_raw[16], const u32 pk_raw[8], const Mem& mem, u64 msg_off, u32 len,
                           u32 k[8]) {
  u32 hdr[16], h[16];
#pragma unroll
  for (int i = 0; i < 8; i++) {
    hdr[i] = sig_raw[i];
    hdr[8 + i] = pk_raw[i];
  }
  sha512_ed(hdr, mem, msg_off, len, h);
  ed25519_k_from_digest(h, k);
}

// verify_strict's extra rules (dalek VerifyingKey::verify_strict): R must decompress and neither R
// nor A may have small order. The rules on R are evaluated in a pass of their own between the hash
// and the verifier (registers to spare): this returns ED_OK, ED_BAD_R or ED_SMALL_ORDER, which
// travels to the verifier in bits 29..31 of the challenge's top word (k < l < 2^253 leaves those
// bits zero). The rule on A runs in the verifier, which decompresses A anyway.
DEV u32 ed25519_strict_r_checks(const u32 sig_raw[16]) {
  fe rx, ry;
  if (!ed_decompress(sig_raw, rx, ry)) return ED_BAD_R;
  return ed_small_order(rx, ry) ? ED_SMALL_ORDER : ED_OK;
}
constexpr u32 ED_K_STRICT_SHIFT = 29;

// ---- lattice-reduced check (T. Pornin, "Optimized Lattice Basis Reduction In Dimension 2, and
// Fast Schnorr and EdDSA Signature Verification", IACR ePrint 2020/454) ----
// dalek's test compress([s]B - [k]A) == R bytes runs a 253-bit doubling chain. With a pair
// (c0, c1) such that c0 = c1 k (mod 8l), c1 odd and |c0|, c1 < 2^136, the check becomes
//     [c1 s mod l] B - [c0] A - [c1] R == identity
// over a 136-bit chain, with R decoded from its bytes. The left side is [c1]([s]B - [k]A - R):
// [c1 k] A = [c0] A since the order of A divides 8l, and [c1 s] B = [c1 s mod l] B since B has
// order l. c1 is odd and below l, so it is prime to every order in the group, and the point is the
// identity exactly when [s]B - [k]A == R. R must decode canonically (y < p, and no sign bit on
// x = 0): compress never produces another encoding, so dalek rejects those bytes, and so does this
// check. The high half of c1 s mod l runs on its own table of odd multiples of [2^136] B.
constexpr int ED_LAT_BITS = SIGOPS_ED_LAT_BITS;
static_assert(ED_LAT_BITS == 136, "carriers, windows and the split of c1 s assume 136 bits");
constexpr int ED_LAT_MAX_STEPS = 320;  // Euclid steps from 2^255 to 2^128: at most 183, plus corrections
constexpr u32 ED_BTAB_HI = 1u << (SIGOPS_WG - 1);  // index of (2j+1) [2^136] B is ED_BTAB_HI + j

#if defined(SIGOPS_HOST_CHECK) && SIGOPS_HOST_CHECK && !defined(__HIP_DEVICE_COMPILE__)
// host check build only: send every item through the full-length check
inline bool& ed_force_full() {
  static bool f = false;
  return f;
}
#endif

// R decoded as compress() could have produced it: y < p and no sign bit on x = 0
DEV bool ed_decode_canonical(const u32 enc[8], fe& x, fe& y) {
  u32 yw[8];
#pragma unroll
  for (int i = 0; i < 8; i++) yw[i] = enc[i];
  yw[7] &= 0x7fffffffu;
  const bool canon = words_lt(yw, ED_P32);
  const bool ok = ed_decompress(enc, x, y);
  return ok && canon && !((enc[7] >> 31) != 0 && fis_zero<Fe>(x));
}

template <int NW>
DEV double words_f64(const u32* w) {  // nearest-ish double of a multi-word value (7 roundings at most)
  double d = (double)w[NW - 1];
#pragma unroll
  for (int i = NW - 2; i >= 0; i--) d = d * 4294967296.0 + (double)w[i];
  return d;
}

// r = a - q b over N words (a, b unsigned or two's complement); returns the high part that the
// product and the borrow left above the top word (0 when a - q b is in [0, 2^(32N)))
template <int N>
DEV u64 words_submul(const u32* a, const u32* b, u32 q, u32* r) {
  u32 hi = 0, br = 0;
#pragma unroll
  for (int i = 0; i < N; i++) {
    const u64 p = (u64)q * b[i] + hi;
    hi = (u32)(p >> 32);
    const u64 d = (u64)a[i] - (u32)p - br;
    r[i] = (u32)d;
    br = (u32)(d >> 63);
  }
  return (u64)hi + br;
}

template <int N>
DEV void words_neg_if(bool c, u32* w) {  // two's complement negation when c
  const u32 m = lane_mask(c);
  u32 cy = c ? 1u : 0u;
#pragma unroll
  for (int i = 0; i < N; i++) {
    const u64 x = (u64)(w[i] ^ m) + cy;
    w[i] = (u32)x;
    cy = (u32)(x >> 32);
  }
}

// (c0, c1) from Euclid's remainder sequence on (8l, k). Every remainder is r_j = t_j k (mod 8l); the
// first one below 2^128 has |t_j| <= 8l / r_(j-1) < 2^127. When t_j is even, t_(j-1) is odd (two
// consecutive t are coprime) and the pair is (r_(j-1) - m r_j, t_(j-1) - m t_j) with the m that
// balances the two coordinates. Quotients come from double-precision estimates (relative error
// below 2^-49, so off by at most one below 2^32), corrected exactly: one too large is undone, one
// too small leaves a remainder >= b for the next step. Every update is exact integer arithmetic,
// so c0 = c1 k (mod 8l) holds whatever the estimates, and the bounds are checked at the end.
// Returns false when no pair within 2^136 was found (about 3 in 10^6 random k) or a quotient
// reached 2^32 (about 2 in 10^8); the caller then runs the full-length check. c0 >= 0 comes out
// with its sign apart (c0_neg), c1 > 0.
DEV bool ed_lattice(const u32 k[8], u32 c0[5], u32 c1[5], bool& c0_neg) {
  u32 a[8], b[8], ta[5], tb[5];
#pragma unroll
  for (int i = 0; i < 8; i++) {
    a[i] = (ED_L32[i] << 3) | (i > 0 ? ED_L32[i - 1] >> 29 : 0u);  // 8l < 2^256
    b[i] = k[i];
  }
#pragma unroll
  for (int i = 0; i < 5; i++) {
    ta[i] = 0;
    tb[i] = i == 0 ? 1u : 0u;
  }
  bool ok = true;
  int steps = 0;
#pragma unroll 1
  while ((b[4] | b[5] | b[6] | b[7]) != 0) {  // a >= b >= 2^128 here
    if (++steps > ED_LAT_MAX_STEPS) {
      ok = false;
      break;
    }
    const double qd = __builtin_floor(words_f64<8>(a) / words_f64<8>(b));
    if (!(qd < 4294967295.0)) {  // a quotient of 2^32 or more (about 2^-32 per step for random k)
      ok = false;
      break;
    }
    u32 q = qd > 1.0 ? (u32)qd : 1u;
    u32 r[8];
    const u64 over = words_submul<8>(a, b, q, r);
    if (over != 0) {  // q was one too large: add b back, which must bring the value to [0, 2^256)
      const u32 cy = words_add(r, b, r);
      q -= 1;
      if ((u64)cy != over) {
        ok = false;
        break;
      }
    }
    u32 tr[5];
    words_submul<5>(ta, tb, q, tr);  // |t| < 2^128 throughout: no overflow of 160 bits
    const bool lt = words_lt(r, b);  // a full step (else a partial one: a <- r)
#pragma unroll
    for (int i = 0; i < 8; i++) {
      const u32 nb = lt ? r[i] : b[i];
      a[i] = lt ? b[i] : r[i];
      b[i] = nb;
    }
#pragma unroll
    for (int i = 0; i < 5; i++) {
      const u32 nt = lt ? tr[i] : tb[i];
      ta[i] = lt ? tb[i] : tr[i];
      tb[i] = nt;
    }
  }
#if defined(SIGOPS_HOST_CHECK) && SIGOPS_HOST_CHECK && !defined(__HIP_DEVICE_COMPILE__)
  if (ed_force_full()) ok = false;
#endif
  u32 w0[8];
  if (tb[0] & 1u) {
#pragma unroll
    for (int i = 0; i < 8; i++) w0[i] = b[i];
#pragma unroll
    for (int i = 0; i < 5; i++) c1[i] = tb[i];
  } else {
    u32 ua[5], ub[5];
#pragma unroll
    for (int i = 0; i < 5; i++) {
      ua[i] = ta[i];
      ub[i] = tb[i];
    }
    words_neg_if<5>((ta[4] >> 31) != 0, ua);
    words_neg_if<5>((tb[4] >> 31) != 0, ub);
    const double md = __builtin_floor((words_f64<8>(a) - words_f64<5>(ua)) / (words_f64<8>(b) + words_f64<5>(ub)));
    if (!(md < 4294967296.0)) ok = false;
    const u32 m = (ok && md > 0.0) ? (u32)md : 0u;
    if (words_submul<8>(a, b, m, w0) != 0) ok = false;  // m above the quotient a / b
    words_submul<5>(ta, tb, m, c1);  // |c1| < 2^160; a wrapped value fails the bound below
  }
  c0_neg = (c1[4] >> 31) != 0;
  words_neg_if<5>(c0_neg, c1);
#pragma unroll
  for (int i = 0; i < 5; i++) c0[i] = w0[i];
  const u32 hi0 = w0[5] | w0[6] | w0[7] | (w0[4] >> (ED_LAT_BITS - 128));
  return ok && hi0 == 0 && (c1[4] >> (ED_LAT_BITS - 128)) == 0;
}

// table of the odd multiples (2e+1) P, e = 0..7, of P = (x, y) in cached form at entries e0..e0+7
template <class Tab>
DEV void ed_table_odd(Tab& tab, int e0, const fe& x, const fe& y) {
  ext P;
  P.X = x;
  P.Y = y;
  P.Z = fe_from_const(Fe::ONE);
  P.T = mont_mul<Fe>(x, y);
  const cached P2 = to_cached(edbl(P, true));
  ext T = P;
  tab.store(e0, to_cached(P));
#pragma unroll 1
  for (int e = 1; e < 8; e++) {
    T = eadd_cached(T, P2, true);
    tab.store(e0 + e, to_cached(T));
  }
}

// acc + (digit's odd multiple of the table's point, negated per digit and stream sign); the table
// holds (2j+1) B at entries base + j in affine Niels form
template <class BTab>
DEV ext ed_add_b(const ext& acc, const BTab& bt, u32 base, u32 digit, bool neg_stream, bool want_t) {
  const bool neg = (digit < 128) != neg_stream;
  fe ypx, ymx, t2d;
  bt.load(base + (((digit < 128) ? ~digit : digit) & 127u), ypx, ymx, t2d);
  const u32 m = lane_mask(neg);
  return eadd_niels(acc, fsel_mask(m, ymx, ypx), fsel_mask(m, ypx, ymx),
                    fsel_mask(m, fsub_nr<Fe, 2, 1>(fe_zero(), t2d), t2d), want_t);
}

// the cached identity (Y+X, Y-X, Z, 2dT) = (1, 1, 1, 0), or entry e of the table when use
template <class Tab>
DEV cached ed_entry_or_identity(const Tab& tab, u32 e, bool neg, bool use) {
  cached c = neg_cached_if(neg, tab.load(e));
  if (!use) {
    const fe one = fe_from_const(Fe::ONE);
    c.ypx = one;
    c.ymx = one;
    c.z = one;
    c.t2d = fe_zero();
  }
  return c;
}

// the 136-bit check: [v0](+-B) + [v1](+-[2^136]B) + [a](+-(-A)) + [c1](-R) (- the +-(-A) that made
// an even |c0| odd) == identity. Carriers of 5 words; tab entries 0..7 hold -A's odd multiples,
// 8..15 -R's. Per 4-bit window: four doublings and the A and R digits; per 8-bit window also the
// two B digits.
template <class Tab, class BTab>
DEV bool ed_check_lattice(const u32 ca[5], bool a_neg, bool a_fix, const u32 cr[5], const u32 cb0[5], bool b0_neg,
                          const u32 cb1[5], bool b1_neg, const Tab& tab, const BTab& bt) {
  constexpr int W = ED_LAT_BITS / 4;
  // top windows: the digits are positive; the A stream's sign applies
  ext acc = eadd_cached(ext_identity(), neg_cached_if(a_neg, tab.load(nibble(ca, W - 1) - 8)), true);
  acc = eadd_cached(acc, tab.load(8 + nibble(cr, W - 1) - 8), false);
#pragma unroll 1
  for (int i = W - 2; i >= 0; i--) {
    const u32 da = nibble(ca, i), dr = nibble(cr, i);
    cached ea = tab.load((da < 8 ? ~da : da) & 7u);
    acc = edbl(acc, false);
    acc = edbl(acc, false);
    acc = edbl(acc, false);
    acc = edbl(acc, true);
    cached er = tab.load(8 + ((dr < 8 ? ~dr : dr) & 7u));
    acc = eadd_cached(acc, neg_cached_if((da < 8) != a_neg, ea), true);
    acc = eadd_cached(acc, neg_cached_if(dr < 8, er), (i & 1) == 0);
    if ((i & 1) == 0) {
      acc = ed_add_b(acc, bt, 0, byte8(cb0, i >> 1), b0_neg, true);
      acc = ed_add_b(acc, bt, ED_BTAB_HI, byte8(cb1, i >> 1), b1_neg, i == 0);
    }
  }
  // the A stream ran on |c0| + 1 when |c0| was even: take its point off once
  acc = eadd_cached(acc, ed_entry_or_identity(tab, 0, !a_neg, a_fix), false);
  return fis_zero<Fe>(acc.X) && feq<Fe>(acc.Y, acc.Z);
}

// the full-length check: compress([s]B + [k](-A)) == R bytes, tab entries 0..7 = -A's odd multiples
template <class Tab, class BTab>
DEV bool ed_check_full(const u32 s[8], const u32 k[8], const u32 sig_raw[16], const Tab& tab, const BTab& bt) {
  // odd representatives: s even -> l - s with B negated; k even -> k + 1 and a final +A
  const bool b_neg = (s[0] & 1u) == 0;
  const bool k_even = (k[0] & 1u) == 0;
  u32 s1[8], k1[8];
  {
    u32 t[8];
    words_sub(ED_L32, s, t);  // only used when s < l (else the result is discarded)
    u32 c = k_even ? 1u : 0u;
#pragma unroll
    for (int i = 0; i < 8; i++) {
      s1[i] = b_neg ? t[i] : s[i];
      const u64 kk = (u64)k[i] + c;
      k1[i] = (u32)kk;
      c = (u32)(kk >> 32);
    }
  }
  u32 c1[8], c2[8];
#pragma unroll
  for (int i = 0; i < 8; i++) {
    c1[i] = (s1[i] >> 1) | (i < 7 ? s1[i + 1] << 31 : 0x80000000u);
    c2[i] = (k1[i] >> 1) | (i < 7 ? k1[i + 1] << 31 : 0x80000000u);
  }
  ext acc = eadd_cached(ext_identity(), tab.load(nibble(c2, 63) - 8), true);
#pragma unroll 1
  for (int i = 62; i >= 0; i--) {
    // issue the HBM load of this window's table entry before the doublings, which hide its latency
    const u32 c = nibble(c2, i);
    const bool neg = c < 8;
    cached E = tab.load((neg ? ~c : c) & 7u);
    acc = edbl(acc, false);
    acc = edbl(acc, false);
    acc = edbl(acc, false);
    acc = edbl(acc, true);
    E = neg_cached_if(neg, E);
    // T is needed only by a following addition: doublings do not read it (uniform branches)
    acc = eadd_cached(acc, E, (i & 1) == 0);
    if ((i & 1) == 0) acc = ed_add_b(acc, bt, 0, byte8(c1, i >> 1), b_neg, i == 0);
  }
  // + A when k was made odd by adding one, else + identity (complete formulas)
  acc = eadd_cached(acc, ed_entry_or_identity(tab, 0, true, k_even), false);
  // compress and compare with R
  const fe zi = finv<Fe>(acc.Z);
  const fe xa = from_mont<Fe>(mont_mul<Fe>(acc.X, zi));
  const fe ya = from_mont<Fe>(mont_mul<Fe>(acc.Y, zi));
  u32 enc[8];
  fe_to_words(ya, enc);
  enc[7] ^= (xa.v[0] & 1u) << 31;
  u32 diff = 0;
#pragma unroll
  for (int i = 0; i < 8; i++) diff |= enc[i] ^ sig_raw[i];
  return diff == 0;
}

// verification given the challenge k (from ed25519_challenge).
// LATTICE: the lattice-reduced check (and the full-length one for challenges without a short
// pair); otherwise the full-length check only. Measured on the MI355X: 3.5% faster in the kernels
// that verify only Ed25519, 2.5% slower in the mixed-scheme kernel, whose secp paths share its
// registers and code.
// BTab: load(idx, ypx, ymx, t2d) for (2 idx + 1) B at idx < ED_BTAB_HI and, with LATTICE,
// (2 j + 1) [2^136] B at idx = ED_BTAB_HI + j. Tab: per-lane store(e, cached) / load(e), 8 entries
// (16 with LATTICE).
// STRICT: dalek VerifyingKey::verify_strict; the status of its rules on R arrives in the top bits
// of k (ed25519_strict_r_checks) and A's small order is checked here.
template <bool STRICT = false, bool LATTICE = false, class Tab, class BTab>
DEV u32 ed25519_verify_core(const u32 sig_raw[16], const u32 pk_raw[8], const u32 k_in[8], Tab& tab,
                            const BTab& bt) {
  u32 k[8];
#pragma unroll
  for (int i = 0; i < 8; i++) k[i] = k_in[i];
  u32 strict_status = ED_OK;
  if constexpr (STRICT) {
    strict_status = k[7] >> ED_K_STRICT_SHIFT;
    k[7] &= (1u << ED_K_STRICT_SHIFT) - 1u;
  }
  u32 status = ED_OK;
  // s (little-endian bytes = little-endian words)
  u32 s[8];
#pragma unroll
  for (int i = 0; i < 8; i++) s[i] = sig_raw[8 + i];
  if (!words_lt(s, ED_L32)) status = ED_BAD_S;
  fe x, y;
  const bool ok = ed_decompress(pk_raw, x, y);
  if (!ok && status == ED_OK) status = ED_BAD_KEY;
  if (STRICT && status == ED_OK) status = strict_status;
  // verify_strict: A of small order (checked after R's rules, as dalek checks R first)
  if (STRICT && status == ED_OK && ed_small_order(x, y)) status = ED_SMALL_ORDER;
  ed_table_odd(tab, 0, fneg<Fe>(x), y);
  u32 c0[5], c1[5];
  bool c0_neg = false;
  bool equal;
  if (LATTICE && ed_lattice(k, c0, c1, c0_neg)) {
    fe rx, ry;
    const bool r_ok = ed_decode_canonical(sig_raw, rx, ry);
    ed_table_odd(tab, 8, fneg<Fe>(rx), ry);
    // v = c1 s mod l, made odd (l - v with B negated), split at bit 136; an even high half takes
    // one more: v = (v_lo - 2^136) + 2^136 (v_hi + 1), the low stream then negated
    u32 v[8];
    {
      u32 cw[8];
#pragma unroll
      for (int i = 0; i < 8; i++) cw[i] = i < 5 ? c1[i] : 0u;
      fe_to_words(fcanon<Fl>(to_mont<Fl>(mont_mul<Fl>(fe_from_words(cw), fe_from_words(s)))), v);
    }
    const bool b_neg = (v[0] & 1u) == 0;
    {
      u32 t[8];
      words_sub(ED_L32, v, t);
#pragma unroll
      for (int i = 0; i < 8; i++) v[i] = b_neg ? t[i] : v[i];
    }
    u32 lo[5] = {v[0], v[1], v[2], v[3], v[4] & 0xffu};
    u32 hi[5] = {(v[4] >> 8) | (v[5] << 24), (v[5] >> 8) | (v[6] << 24), (v[6] >> 8) | (v[7] << 24), v[7] >> 8, 0u};
    const bool hi_even = (hi[0] & 1u) == 0;
    hi[0] |= 1u;
    {
      const u32 top[5] = {0, 0, 0, 0, 1u << (ED_LAT_BITS - 128)};
      u32 t[5];
      u32 br = 0;
#pragma unroll
      for (int i = 0; i < 5; i++) {
        const u64 d = (u64)top[i] - lo[i] - br;
        t[i] = (u32)d;
        br = (u32)(d >> 63);
      }
#pragma unroll
      for (int i = 0; i < 5; i++) lo[i] = hi_even ? t[i] : lo[i];
    }
    const bool a_fix = (c0[0] & 1u) == 0;
    c0[0] |= 1u;
    u32 ca[5], cr[5], cb0[5], cb1[5];
    glv_carrier<ED_LAT_BITS>(c0, ca);
    glv_carrier<ED_LAT_BITS>(c1, cr);
    glv_carrier<ED_LAT_BITS>(lo, cb0);
    glv_carrier<ED_LAT_BITS>(hi, cb1);
    equal = ed_check_lattice(ca, c0_neg, a_fix, cr, cb0, b_neg != hi_even, cb1, b_neg, tab, bt) && r_ok;
  } else {
    equal = ed_check_full(s, k, sig_raw, tab, bt);
  }
  if (status == ED_OK && !equal) status = ED_MISMATCH;
  return status;
}

template <bool STRICT = false, bool LATTICE = false, class Tab, class BTab, class Mem>
DEV u32 ed25519_verify_item(const u32 sig_raw[16], const u32 pk_raw[8], const Mem& mem, u64 msg_off, u32 len,
                            Tab& tab, const BTab& bt) {
  u32 k[8];
  ed25519_challenge(sig_raw, pk_raw, mem, msg_off, len, k);
  if constexpr (STRICT) k[7] |= ed25519_strict_r_checks(sig_raw) << ED_K_STRICT_SHIFT;
  return ed25519_verify_core<STRICT, LATTICE>(sig_raw, pk_raw, k, tab, bt);
}

}  // namespace sigops
