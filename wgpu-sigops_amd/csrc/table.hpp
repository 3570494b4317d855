// table.hpp -- caller-supplied generator tables in the reference's layout, per-lane steps.
//
// The reference passes its fixed-base table with every call (`table_limbs`, built by
// src/precompute.rs:12-69): 16 affine points [P, 2P, ..., 16P], each coordinate x * 2^(nl*l) mod p
// stored as nl = 256 / l + 1 little-endian limbs of l = 11..15 bits (src/tests/mod.rs:94-149;
// secp entries hold (x, y), Ed25519 entries (x, y, x*y)). Its stage-1 shaders multiply u1 (or s)
// by that table with unsigned 4-bit windows (src/wgsl/secp256k1_curve.wgsl:388-447), so for a
// table of multiples of a point P the result is u1 * P. These steps turn such a table into this
// engine's odd-multiple table of P (128 entries, w = 8, the resident tables' format), checking on
// the way that the table is what its layout promises: every limb below 2^l, every value below p,
// every entry on the curve and entry j equal to (j + 1) * P; for Ed25519 also t = x*y and P in the
// prime-order subgroup, which the signed-digit recoding of s relies on.
// A workgroup of 128 lanes runs check_entry on lanes 0..15 (entry j), shares entry 0, then
// build_entry on every lane (odd multiple 2j + 1). Failures OR bits into `bad`:
// 1 = a limb, value or curve equation, 2 = entry j is not (j + 1) P, 4 = an odd multiple is the
// point at infinity, 8 = [l] P is not the identity (Ed25519).
#pragma once
#include "ed25519.hpp"

namespace sigops {

constexpr u32 TAB_BAD_FORMAT = 1, TAB_BAD_MULTIPLE = 2, TAB_BAD_INFINITY = 4, TAB_BAD_ORDER = 8;

// C_l for limb width lbits (11..15): the product with it turns x * 2^(nl*l) into x * 2^261
template <class F>
DEV const u32* refconv_table();
template <>
DEV const u32* refconv_table<ModPk1>() { return REFCONV_K1; }
template <>
DEV const u32* refconv_table<ModPr1>() { return REFCONV_R1; }
template <>
DEV const u32* refconv_table<ModPed>() { return REFCONV_ED; }

template <class F>
DEV fe refconv(u32 lbits) {
  const u32* c = refconv_table<F>();
  fe r;
#pragma unroll
  for (int k = 0; k < NL; k++) r.v[k] = c[9 * (lbits - 11) + k];
  BOUNDS_SET(r, LMASK, (1u << 24) - 1u, p2(256));
  return r;
}

// one coordinate of nl limbs of lbits bits -> Montgomery form; false if a limb has bits at or
// above 2^lbits or the value is not below p
template <class F>
DEV bool ref_coord(const u32* limbs, u32 nl, u32 lbits, const u32 (&p32)[8], fe& out) {
  u32 w[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  bool ok = true;
  for (u32 i = 0; i < nl; i++) {
    const u32 limb = limbs[i];
    if (limb >> lbits) ok = false;
    const u32 b = lbits * i;  // < 256 for every lbits in 11..15
    const u64 v = (u64)(limb & ((1u << lbits) - 1u)) << (b & 31);
    w[b >> 5] |= (u32)v;
    w[(b >> 5) + 1] |= (u32)(v >> 32);
  }
  if (w[8] != 0 || !words_lt(w, p32)) ok = false;
  out = mont_mul<F>(fe_from_words(w), refconv<F>(lbits));
  return ok;
}

// ---- secp256k1 / secp256r1: entries (x, y) ----
template <class C>
DEV void tab_secp_check_entry(u32 j, const u32* limbs, u32 nl, u32 lbits, fe& x, fe& y, u32& bad) {
  typedef typename C::Fp F;
  const bool okx = ref_coord<F>(limbs + (2 * j) * nl, nl, lbits, C::p32(), x);
  const bool oky = ref_coord<F>(limbs + (2 * j + 1) * nl, nl, lbits, C::p32(), y);
  fe rhs = mont_mul<F>(mont_sqr<F>(x), x);
  if (C::A_IS_M3) rhs = fsub<F>(rhs, fadd<F>(fdbl<F>(x), x));
  rhs = fadd<F>(rhs, C::b());
  if (!okx || !oky || !feq<F>(mont_sqr<F>(y), rhs)) bad |= TAB_BAD_FORMAT;
}

// k * P0 with the complete law (k < 256)
template <class C>
DEV jacf tab_secp_mul_small(const jacf& P0, u32 k) {
  jacf acc;
  acc.P = P0.P;
  acc.inf = true;
#pragma unroll 1
  for (int b = 7; b >= 0; b--) {
    acc = jdbl_c<C>(acc);
    if ((k >> b) & 1u) acc = jadd_c<C>(acc, P0);
  }
  return acc;
}

// p0: canonical x, y of entry 0. Lanes j < 16 pass their own decoded entry (x, y) to compare with
// (j + 1) P0; every lane writes the 20 words of odd multiple (2j + 1) P0 to out.
template <class C>
DEV void tab_secp_build_entry(u32 j, const u32 p0[2 * NL], const fe& x, const fe& y, u32& bad, u32 out[20]) {
  typedef typename C::Fp F;
  jacf P0;
#pragma unroll
  for (int k = 0; k < NL; k++) {
    P0.P.X.v[k] = p0[k];
    P0.P.Y.v[k] = p0[NL + k];
  }
  BOUNDS_SET(P0.P.X, LMASK, (1u << 24) - 1u, mod_ld<F>());
  BOUNDS_SET(P0.P.Y, LMASK, (1u << 24) - 1u, mod_ld<F>());
  P0.P.Z = fe_from_const(F::ONE);
  P0.inf = false;
  if (j < 16) {
    const jacf Q = tab_secp_mul_small<C>(P0, j + 1);
    const fe z2 = mont_sqr<F>(Q.P.Z);
    const bool eq = !Q.inf && feq<F>(Q.P.X, mont_mul<F>(x, z2)) &&
                    feq<F>(Q.P.Y, mont_mul<F>(y, mont_mul<F>(z2, Q.P.Z)));
    if (!eq) bad |= TAB_BAD_MULTIPLE;
  }
  const jacf Q = tab_secp_mul_small<C>(P0, 2 * j + 1);
  if (Q.inf) bad |= TAB_BAD_INFINITY;
  const fe zi = finv<F>(Q.P.Z);
  const fe zi2 = mont_sqr<F>(zi);
  const fe ax = fcanon<F>(mont_mul<F>(Q.P.X, zi2));
  const fe ay = fcanon<F>(mont_mul<F>(Q.P.Y, mont_mul<F>(zi2, zi)));
#pragma unroll
  for (int k = 0; k < NL; k++) {
    out[k] = ax.v[k];
    out[NL + k] = ay.v[k];
  }
  out[18] = out[19] = 0;
}

// ---- Ed25519: entries (x, y, t = x*y); table out in affine Niels form (y+x, y-x, 2d*x*y) ----
DEV void tab_ed_check_entry(u32 j, const u32* limbs, u32 nl, u32 lbits, fe& x, fe& y, u32& bad) {
  typedef ModPed F;
  fe t;
  const bool okx = ref_coord<F>(limbs + (3 * j) * nl, nl, lbits, ED_P32, x);
  const bool oky = ref_coord<F>(limbs + (3 * j + 1) * nl, nl, lbits, ED_P32, y);
  const bool okt = ref_coord<F>(limbs + (3 * j + 2) * nl, nl, lbits, ED_P32, t);
  const fe one = fe_from_const(F::ONE);
  const fe xx = mont_sqr<F>(x), yy = mont_sqr<F>(y);
  const bool on = feq<F>(fsub<F>(yy, xx), fadd<F>(one, mont_mul<F>(fe_from_const(ED_D), mont_mul<F>(xx, yy))));
  if (!okx || !oky || !okt || !on || !feq<F>(t, mont_mul<F>(x, y))) bad |= TAB_BAD_FORMAT;
}

// a point in extended coordinates as a table entry: affine Niels form (y+x, y-x, 2d*x*y), canonical
DEV void tab_ed_niels(const ext& Q, u32 out[28]) {
  typedef ModPed F;
  const fe zi = finv<F>(Q.Z);
  const fe ax = mont_mul<F>(Q.X, zi), ay = mont_mul<F>(Q.Y, zi);
  const fe ypx = fcanon<F>(fadd<F>(ay, ax));
  const fe ymx = fcanon<F>(fsub<F>(ay, ax));
  const fe t2d = fcanon<F>(mont_mul<F>(mont_mul<F>(ax, ay), fe_from_const(ED_D2)));
#pragma unroll
  for (int k = 0; k < NL; k++) {
    out[k] = ypx.v[k];
    out[NL + k] = ymx.v[k];
    out[2 * NL + k] = t2d.v[k];
  }
  out[27] = 0;
}

DEV ext tab_ed_mul_small(const cached& P0c, u32 k, int top) {
  ext acc = ext_identity();
#pragma unroll 1
  for (int b = top; b >= 0; b--) {
    acc = edbl(acc, true);
    if ((k >> b) & 1u) acc = eadd_cached(acc, P0c, true);
  }
  return acc;
}

DEV void tab_ed_build_entry(u32 j, const u32 p0[2 * NL], const fe& x, const fe& y, u32& bad, u32 out[28]) {
  typedef ModPed F;
  const fe one = fe_from_const(F::ONE);
  ext P0;
#pragma unroll
  for (int k = 0; k < NL; k++) {
    P0.X.v[k] = p0[k];
    P0.Y.v[k] = p0[NL + k];
  }
  BOUNDS_SET(P0.X, LMASK, (1u << 24) - 1u, mod_ld<F>());
  BOUNDS_SET(P0.Y, LMASK, (1u << 24) - 1u, mod_ld<F>());
  P0.Z = one;
  P0.T = mont_mul<F>(P0.X, P0.Y);
  const cached P0c = to_cached(P0);
  if (j < 16) {
    const ext Q = tab_ed_mul_small(P0c, j + 1, 4);
    if (!(feq<F>(Q.X, mont_mul<F>(x, Q.Z)) && feq<F>(Q.Y, mont_mul<F>(y, Q.Z)))) bad |= TAB_BAD_MULTIPLE;
  }
  if (j == 0) {  // [l] P0 must be the identity
    ext acc = ext_identity();
#pragma unroll 1
    for (int b = 252; b >= 0; b--) {
      acc = edbl(acc, true);
      if ((ED_L32[b >> 5] >> (b & 31)) & 1u) acc = eadd_cached(acc, P0c, true);
    }
    if (!(fis_zero<F>(acc.X) && feq<F>(acc.Y, acc.Z))) bad |= TAB_BAD_ORDER;
  }
  tab_ed_niels(tab_ed_mul_small(P0c, 2 * j + 1, 7), out);
}

// entry j of the high-half table: (2j+1) [2^136] P0 (P0 = (p0 x, p0 y), affine Montgomery form)
DEV void tab_ed_build_hi_entry(u32 j, const u32 p0[2 * NL], u32 out[28]) {
  typedef ModPed F;
  ext P;
#pragma unroll
  for (int k = 0; k < NL; k++) {
    P.X.v[k] = p0[k];
    P.Y.v[k] = p0[NL + k];
  }
  BOUNDS_SET(P.X, LMASK, (1u << 24) - 1u, mod_ld<F>());
  BOUNDS_SET(P.Y, LMASK, (1u << 24) - 1u, mod_ld<F>());
  P.Z = fe_from_const(F::ONE);
  P.T = fe_zero();  // doubling does not read T
#pragma unroll 1
  for (int b = 0; b < ED_LAT_BITS; b++) P = edbl(P, b == ED_LAT_BITS - 1);
  tab_ed_niels(tab_ed_mul_small(to_cached(P), 2 * j + 1, 7), out);
}

}  // namespace sigops
