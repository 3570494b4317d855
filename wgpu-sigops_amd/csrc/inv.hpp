// inv.hpp -- modular inversion by divsteps (binary GCD), one element per lane.
//
// The reference inverts with a Fermat exponentiation (src/wgsl/ff.wgsl:60-120, "fails if x == 0"
// :57-58). Here the inputs are public signature data, so the inversion need not be constant time,
// but SIMD lanes want one fixed instruction stream; the Bernstein-Yang divstep recurrence ("Fast
// constant-time gcd computation and modular inversion", 2019) gives both:
//
//   divstep(delta, f, g): if delta > 0 and g odd: (1 - delta, g, (g - f)/2)
//                         elif g odd:             (1 + delta, f, (g + f)/2)
//                         else:                   (1 + delta, f, g/2)
//
// from (delta, f, g) = (1/2, M, x) reaches g = 0, f = +-1 within 590 steps for any 256-bit odd M and
// 0 < x < M (half-delta variant; the bound is the computed worst case of P. Wuille's analysis of
// that variant). Each step is a linear map of (f, g) with a factor 1/2, so after N steps
// 2^N (f, g) = T (M, x) for an integer matrix T = [[u, v], [q, r]], and v x = +-2^N (mod M).
// The same maps applied to (d, e) = (0, e0) mod M give d = f * x^-1 * e0 at the end.
//
// GPU shape: 20 batches of 30 steps. A batch runs its 30 decisions on the low 32 bits of f and g
// (enough for 30 steps) with 32-bit ops and no branches, producing T with |u|+|v|, |q|+|r| <= 2^30;
// then applies T to the full f, g (exact division by 2^30) and to d, e (division by 2^30 mod M, by
// adding the multiple of M that clears the low 30 bits, taken in [-2^29, 2^29) so |d|, |e| grow by
// at most M/2 per batch). Full-width values are 9 signed limbs of 30 bits.
#pragma once
#include "field.hpp"

namespace sigops {

constexpr int INV_BATCHES = 20;  // 600 divsteps >= 590
// Random inputs take 517 divsteps on average and at most 533 in a sample of 19,200 (for both
// secp256k1 moduli); every 64-lane wave of that sample was done within 18 batches. Once g = 0 on
// every lane of the wave, further batches leave f and d unchanged (u = 2^30, v = 0 and no multiple
// of M added), so from this batch on the loop ends as soon as the whole wave has converged.
constexpr int INV_CHECK_FROM = 18;
constexpr i32 L30 = (1 << 30) - 1;

// true when c holds on every active lane of the wavefront (on the host: c)
DEV bool wave_all(bool c) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __all(c);
#else
  return c;
#endif
}

// 9 limbs of 29 bits (value < 2^261) -> 9 limbs of 30 bits
DEV void limbs29_to_30(const fe& a, i32 o[9]) {
#pragma unroll
  for (int k = 0; k < 9; k++) {
    const int bit = 30 * k;
    const int li = bit / 29, sh = bit % 29;
    u64 x = (u64)a.v[li] >> sh;
    if (li + 1 < NL) x |= (u64)a.v[li + 1] << (29 - sh);
    if (li + 2 < NL) x |= (u64)a.v[li + 2] << (58 - sh);
    o[k] = (i32)((u32)x & (u32)L30);
  }
}

// 9 non-negative normalised limbs of 30 bits (value < 2^261) -> 9 limbs of 29 bits
DEV fe limbs30_to_29(const i32 a[9]) {
  fe r;
#pragma unroll
  for (int k = 0; k < NL; k++) {
    const int bit = 29 * k;
    const int li = bit / 30, sh = bit % 30;
    u64 x = (u64)(u32)a[li] >> sh;
    if (li + 1 < 9) x |= (u64)(u32)a[li + 1] << (30 - sh);
    r.v[k] = (u32)x & LMASK;
  }
  return r;
}

// 30 divsteps on the low words of f and g: eta = 2 delta, returns T = [[u, v], [q, r]]
DEV void divsteps30(i32& eta, u32 f, u32 g, i32& u, i32& v, i32& q, i32& r) {
  i32 uu = 1, vv = 0, qq = 0, rr = 1;
#pragma unroll
  for (int j = 0; j < 30; j++) {
    // swap when delta > 0 and g odd: (delta, f, g, u, v, q, r) <- (-delta, g, -f, q, r, -u, -v)
    const u32 m1 = (u32)((-eta) >> 31) & (0u - (g & 1u));
    const u32 nf = (g & m1) | (f & ~m1);
    const u32 ng = ((0u - f) & m1) | (g & ~m1);
    const i32 nu = (qq & (i32)m1) | (uu & ~(i32)m1);
    const i32 nv = (rr & (i32)m1) | (vv & ~(i32)m1);
    const i32 nq = ((-uu) & (i32)m1) | (qq & ~(i32)m1);
    const i32 nr = ((-vv) & (i32)m1) | (rr & ~(i32)m1);
    eta = ((-eta) & (i32)m1) | (eta & ~(i32)m1);
    f = nf;
    uu = nu;
    vv = nv;
    // g odd: g += f, (q, r) += (u, v); then g /= 2 (exact), (u, v) *= 2, delta += 1
    const u32 m2 = 0u - (ng & 1u);
    g = (ng + (f & m2)) >> 1;
    qq = nq + (uu & (i32)m2);
    rr = nr + (vv & (i32)m2);
    // doubled as unsigned words: a left shift of a negative signed value is undefined in C++17
    uu = (i32)((u32)uu << 1);
    vv = (i32)((u32)vv << 1);
    eta += 2;
  }
  u = uu;
  v = vv;
  q = qq;
  r = rr;
}

// (f, g) <- (u f + v g, q f + r g) / 2^30, exact
DEV void apply_fg(i32 f[9], i32 g[9], i32 u, i32 v, i32 q, i32 r) {
  i64 cf = (i64)u * f[0] + (i64)v * g[0];
  i64 cg = (i64)q * f[0] + (i64)r * g[0];
  cf >>= 30;
  cg >>= 30;
#pragma unroll
  for (int i = 1; i < 9; i++) {
    cf += (i64)u * f[i] + (i64)v * g[i];
    cg += (i64)q * f[i] + (i64)r * g[i];
    f[i - 1] = (i32)cf & L30;
    g[i - 1] = (i32)cg & L30;
    cf >>= 30;
    cg >>= 30;
  }
  f[8] = (i32)cf;
  g[8] = (i32)cg;
}

// (d, e) <- (u d + v e + md M, q d + r e + me M) / 2^30 with md, me in [-2^29, 2^29) chosen so the
// divisions are exact
template <class F>
DEV void apply_de(i32 d[9], i32 e[9], i32 u, i32 v, i32 q, i32 r) {
  const u32 td = (u32)u * (u32)d[0] + (u32)v * (u32)e[0];
  const u32 te = (u32)q * (u32)d[0] + (u32)r * (u32)e[0];
  // sign-extend the low 30 bits of -t * M^-1
  const i32 md = (i32)((0u - td * F::MINV30) << 2) >> 2;
  const i32 me = (i32)((0u - te * F::MINV30) << 2) >> 2;
  i64 cd = (i64)u * d[0] + (i64)v * e[0] + (i64)md * F::M30[0];
  i64 ce = (i64)q * d[0] + (i64)r * e[0] + (i64)me * F::M30[0];
  cd >>= 30;
  ce >>= 30;
#pragma unroll
  for (int i = 1; i < 9; i++) {
    cd += (i64)u * d[i] + (i64)v * e[i] + (i64)md * F::M30[i];
    ce += (i64)q * d[i] + (i64)r * e[i] + (i64)me * F::M30[i];
    d[i - 1] = (i32)cd & L30;
    e[i - 1] = (i32)ce & L30;
    cd >>= 30;
    ce >>= 30;
  }
  d[8] = (i32)cd;
  e[8] = (i32)ce;
}

// Montgomery-form inverse of a Montgomery-form x (x != 0 mod M): returns x^-1 * R (mod M),
// normalised, value < 2M.
template <class F>
DEV fe finv_gcd(const fe& x) {
  const fe xc = fcanon<F>(x);  // canonical: 0 < xc < M
  i32 f[9], g[9], d[9], e[9];
  limbs29_to_30(xc, g);
  // seed e0 such that the result is already in Montgomery form:
  //   f d = xc^-1 e0 = (x R)^-1 e0  ->  e0 = R^2 gives x^-1 R directly (sparse moduli, followed by
  //   a fold), e0 = R gives x^-1, turned into x^-1 R by one product with R^2 (dense moduli)
#pragma unroll
  for (int i = 0; i < 9; i++) {
    f[i] = F::M30[i];
    d[i] = 0;
    e[i] = F::SPARSE ? F::R2MOD30[i] : F::RMOD30[i];
  }
  i32 eta = 1;
#pragma unroll 1
  for (int b = 0; b < INV_BATCHES; b++) {
    if (b >= INV_CHECK_FROM) {
      u32 gz = 0;
#pragma unroll
      for (int i = 0; i < 9; i++) gz |= (u32)g[i];
      if (wave_all(gz == 0)) break;
    }
    i32 u, v, q, r;
    divsteps30(eta, (u32)f[0] | ((u32)f[1] << 30), (u32)g[0] | ((u32)g[1] << 30), u, v, q, r);
    apply_fg(f, g, u, v, q, r);
    apply_de<F>(d, e, u, v, q, r);
  }
  // now g = 0 and f = +-1: the inverse times e0 is f d; take f d + 16 M >= 0 (|d| < 11 M)
  const u32 neg = (u32)(f[8] >> 31);  // 0 or ~0
  i64 c = 0;
  i32 t[9];
#pragma unroll
  for (int i = 0; i < 9; i++) {
    const i32 di = (i32)(((u32)d[i] ^ neg) - neg);  // -d limbwise when f = -1
    c += (i64)di + 16 * (i64)F::M30[i];
    t[i] = (i32)c & L30;
    c >>= 30;
  }
  fe rr = limbs30_to_29(t);  // 0 <= f d + 16 M < 27 M < 2^261
#if SIGOPS_BOUNDS
  {
    // x = 0 (an invalid signature's r, whose result is discarded) has no inverse; any other input
    // must have converged
    u32 gz = 0, xz = 0;
    for (int i = 0; i < 9; i++) gz |= (u32)g[i];
    for (int k = 0; k < NL; k++) xz |= xc.v[k];
    BCHECK(xz == 0 || (gz == 0 && (f[8] == 0 || f[8] == -1)), "divsteps did not reach g = 0 within the batches");
    BOUNDS_SET(rr, LMASK, LMASK, 27.0L * mod_ld<F>());
  }
#endif
  if constexpr (F::SPARSE) return fnorm_red<F>(rr);
  else return mont_mul<F>(rr, fe_from_const(F::R2));
}

template <class F>
DEV fe finv(const fe& x) {  // x != 0 mod M
  return finv_gcd<F>(x);
}

}  // namespace sigops
