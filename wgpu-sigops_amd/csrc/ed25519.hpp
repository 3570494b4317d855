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

// verification given the challenge k (from ed25519_challenge).
// BTab: load(idx, ypx, ymx, t2d) for (2 idx + 1) B. Tab: per-lane store(e, cached) / load(e).
// STRICT: dalek VerifyingKey::verify_strict; the status of its rules on R arrives in the top bits
// of k (ed25519_strict_r_checks) and A's small order is checked here. The final comparison stays
// on the compressed bytes of R.
template <bool STRICT = false, class Tab, class BTab>
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
  const fe one = fe_from_const(Fe::ONE);
  if (STRICT && status == ED_OK) status = strict_status;
  // verify_strict: A of small order (checked after R's rules, as dalek checks R first)
  if (STRICT && status == ED_OK && ed_small_order(x, y)) status = ED_SMALL_ORDER;
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
  // table of odd multiples of -A
  ext nA;
  nA.X = fneg<Fe>(x);
  nA.Y = y;
  nA.Z = one;
  nA.T = mont_mul<Fe>(nA.X, y);
  const cached nA2 = to_cached(edbl(nA, true));
  ext T = nA;
  tab.store(0, to_cached(nA));
#pragma unroll 1
  for (int e = 1; e < 8; e++) {
    T = eadd_cached(T, nA2, true);
    tab.store(e, to_cached(T));
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
    if ((i & 1) == 0) {
      const u32 cb = byte8(c1, i >> 1);
      const bool bneg = (cb < 128) != b_neg;
      fe ypx, ymx, t2d;
      bt.load(((cb < 128) ? ~cb : cb) & 127u, ypx, ymx, t2d);
      const u32 m = lane_mask(bneg);
      acc = eadd_niels(acc, fsel_mask(m, ymx, ypx), fsel_mask(m, ypx, ymx),
                       fsel_mask(m, fsub_nr<Fe, 2, 1>(fe_zero(), t2d), t2d), i == 0);
    }
  }
  {
    // + A when k was made odd by adding one, else + identity (complete formulas)
    cached corr = neg_cached_if(true, tab.load(0));
    if (!k_even) {
      corr.ypx = one;
      corr.ymx = one;
      corr.z = one;
      corr.t2d = fe_zero();
    }
    acc = eadd_cached(acc, corr, false);
  }
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
  if (status == ED_OK && diff != 0) status = ED_MISMATCH;
  return status;
}

template <bool STRICT = false, class Tab, class BTab, class Mem>
DEV u32 ed25519_verify_item(const u32 sig_raw[16], const u32 pk_raw[8], const Mem& mem, u64 msg_off, u32 len,
                            Tab& tab, const BTab& bt) {
  u32 k[8];
  ed25519_challenge(sig_raw, pk_raw, mem, msg_off, len, k);
  if constexpr (STRICT) k[7] |= ed25519_strict_r_checks(sig_raw) << ED_K_STRICT_SHIFT;
  return ed25519_verify_core<STRICT>(sig_raw, pk_raw, k, tab, bt);
}

}  // namespace sigops
