// field.hpp -- 256-bit modular arithmetic for gfx950 (CDNA4), one element per lane.
//
// Replaces the reference's WGSL limb code: bigint.wgsl (add/sub/compare, :25-141), ff.wgsl
// (Barrett ff_mul :169-198, ff_inverse :60-120, ff_add/ff_sub :6-48), mont.wgsl (mont_mul :5-70,
// modpow :80-99). The reference packs 20 x 13-bit limbs because WGSL has no 64-bit multiply; here
// the limb width is chosen for the hardware instead:
//
//   * 9 limbs of 29 bits (261 bits), Montgomery form with R = 2^261 for every modulus.
//   * A product a_i*b_j < 2^58 goes into a 64-bit column accumulator with ONE v_mad_u64_u32 whose
//     64-bit addend and result are the same even-aligned VGPR pair, so the product and the
//     interleaved Montgomery reduction issue no carry or register-move instructions; a column
//     holds up to 9 a*b plus 9 m*M products (< 2^62.2) without overflow. (With 8 x 32-bit limbs
//     every product needs a carry add plus moves to rebuild aligned pairs: ~2x the VALU work.)
//   * Values are kept "almost reduced" in [0, 2M) with normalised limbs (< 2^29); only
//     comparisons and outputs canonicalise to [0, M).
#pragma once
#include <cstdint>

#if defined(__HIPCC__) || defined(__HIP__)
#include <hip/hip_runtime.h>
#define DEV __device__ __forceinline__
// The modular multiplication is inlined: an out-of-line copy (SIGOPS_OUTLINE_MUL=1) bounds the live
// registers around each call but pays for call frames in scratch memory at every call site.
#if defined(SIGOPS_OUTLINE_MUL) && SIGOPS_OUTLINE_MUL
#define DEV_MUL __device__ __noinline__
#else
#define DEV_MUL __device__ __forceinline__
#endif
#else
// Host build of the same arithmetic (tests/devcheck compiles these headers with g++ to check the
// device algorithms against the oracle without a GPU).
#define DEV inline
#define DEV_MUL inline
#endif

#include "consts.hpp"

namespace sigops {

typedef uint32_t u32;
typedef uint64_t u64;
typedef int32_t i32;
typedef int64_t i64;

constexpr int NL = SIGOPS_NLIMBS;
constexpr u32 LMASK = (1u << SIGOPS_LIMB_BITS) - 1u;

struct fe {
  u32 v[NL];
};

DEV fe fe_from_const(const u32 (&c)[NL]) {
  fe r;
#pragma unroll
  for (int k = 0; k < NL; k++) r.v[k] = c[k];
  return r;
}

DEV fe fe_zero() {
  fe r;
#pragma unroll
  for (int k = 0; k < NL; k++) r.v[k] = 0;
  return r;
}

// Montgomery product a*b*2^-261 mod M for a, b < 2M (output < 2M, limbs normalised).
// Operand-scanning with interleaved reduction: after row i is accumulated, column i is final
// (it only lacks the carry of column i-1), so m_i = col_i * N0 mod 2^29 and adding m_i*M zeroes
// its low 29 bits.
template <class F>
DEV_MUL fe mont_mul(const fe& a, const fe& b) {
  u64 acc[2 * NL - 1];
#pragma unroll
  for (int k = 0; k < 2 * NL - 1; k++) acc[k] = 0;
#pragma unroll
  for (int i = 0; i < NL; i++) {
#pragma unroll
    for (int j = 0; j < NL; j++) acc[i + j] += (u64)a.v[i] * b.v[j];
    if (i > 0) acc[i] += acc[i - 1] >> SIGOPS_LIMB_BITS;
    const u32 m = ((u32)acc[i] * F::N0) & LMASK;
#pragma unroll
    for (int j = 0; j < NL; j++) acc[i + j] += (u64)m * F::M[j];
  }
  fe r;
  u64 c = acc[NL - 1] >> SIGOPS_LIMB_BITS;
#pragma unroll
  for (int k = 0; k < NL - 1; k++) {
    const u64 t = acc[NL + k] + c;
    r.v[k] = (u32)t & LMASK;
    c = t >> SIGOPS_LIMB_BITS;
  }
  r.v[NL - 1] = (u32)c;
  return r;
}

// Montgomery square: 45 instead of 81 partial products (cross terms use a pre-doubled operand,
// < 2^30, so a cross product is < 2^59 and a column still stays below 2^63).
template <class F>
DEV_MUL fe mont_sqr(const fe& a) {
  u32 a2[NL];
#pragma unroll
  for (int k = 0; k < NL; k++) a2[k] = a.v[k] << 1;
  u64 acc[2 * NL - 1];
#pragma unroll
  for (int k = 0; k < 2 * NL - 1; k++) acc[k] = 0;
#pragma unroll
  for (int i = 0; i < NL; i++) {
    acc[2 * i] += (u64)a.v[i] * a.v[i];
#pragma unroll
    for (int j = i + 1; j < NL; j++) acc[i + j] += (u64)a.v[i] * a2[j];
    if (i > 0) acc[i] += acc[i - 1] >> SIGOPS_LIMB_BITS;
    const u32 m = ((u32)acc[i] * F::N0) & LMASK;
#pragma unroll
    for (int j = 0; j < NL; j++) acc[i + j] += (u64)m * F::M[j];
  }
  fe r;
  u64 c = acc[NL - 1] >> SIGOPS_LIMB_BITS;
#pragma unroll
  for (int k = 0; k < NL - 1; k++) {
    const u64 t = acc[NL + k] + c;
    r.v[k] = (u32)t & LMASK;
    c = t >> SIGOPS_LIMB_BITS;
  }
  r.v[NL - 1] = (u32)c;
  return r;
}

// s in [0, 4M) with normalised limbs -> [0, 2M)
template <class F>
DEV fe reduce_2m(const fe& s) {
  fe t;
  i32 br = 0;
#pragma unroll
  for (int k = 0; k < NL; k++) {
    const i32 x = (i32)s.v[k] - (i32)F::M2[k] + br;
    t.v[k] = (u32)x & LMASK;
    br = x >> SIGOPS_LIMB_BITS;  // arithmetic shift: 0 or -1
  }
  const bool keep = br < 0;
  fe r;
#pragma unroll
  for (int k = 0; k < NL; k++) r.v[k] = keep ? s.v[k] : t.v[k];
  return r;
}

template <class F>
DEV fe fadd(const fe& a, const fe& b) {
  fe s;
  u32 c = 0;
#pragma unroll
  for (int k = 0; k < NL; k++) {
    const u32 x = a.v[k] + b.v[k] + c;
    s.v[k] = x & LMASK;
    c = x >> SIGOPS_LIMB_BITS;
  }
  return reduce_2m<F>(s);
}

// a - b + 2M in (0, 4M), then to [0, 2M)
template <class F>
DEV fe fsub(const fe& a, const fe& b) {
  fe s;
  i32 c = 0;
#pragma unroll
  for (int k = 0; k < NL; k++) {
    const i32 x = (i32)a.v[k] - (i32)b.v[k] + (i32)F::M2[k] + c;
    s.v[k] = (u32)x & LMASK;
    c = x >> SIGOPS_LIMB_BITS;
  }
  return reduce_2m<F>(s);
}

template <class F>
DEV fe fneg(const fe& a) {
  return fsub<F>(fe_zero(), a);
}

template <class F>
DEV fe fdbl(const fe& a) {
  return fadd<F>(a, a);
}

// [0, 2M) -> [0, M)
template <class F>
DEV fe fcanon(const fe& s) {
  fe t;
  i32 br = 0;
#pragma unroll
  for (int k = 0; k < NL; k++) {
    const i32 x = (i32)s.v[k] - (i32)F::M[k] + br;
    t.v[k] = (u32)x & LMASK;
    br = x >> SIGOPS_LIMB_BITS;
  }
  const bool keep = br < 0;
  fe r;
#pragma unroll
  for (int k = 0; k < NL; k++) r.v[k] = keep ? s.v[k] : t.v[k];
  return r;
}

// x == 0 mod M for x in [0, 2M)
template <class F>
DEV bool fis_zero(const fe& x) {
  u32 z = 0, zm = 0;
#pragma unroll
  for (int k = 0; k < NL; k++) {
    z |= x.v[k];
    zm |= x.v[k] ^ F::M[k];
  }
  return z == 0 || zm == 0;
}

template <class F>
DEV bool feq(const fe& a, const fe& b) {
  return fis_zero<F>(fsub<F>(a, b));
}

DEV fe fsel(bool c, const fe& a, const fe& b) {
  fe r;
#pragma unroll
  for (int k = 0; k < NL; k++) r.v[k] = c ? a.v[k] : b.v[k];
  return r;
}

template <class F>
DEV fe to_mont(const fe& x) {  // x any value < 2^261 with normalised limbs, x * R2 < M * 2^261
  return mont_mul<F>(x, fe_from_const(F::R2));
}

template <class F>
DEV fe from_mont(const fe& x) {  // canonical plain value
  fe one = fe_zero();
  one.v[0] = 1;
  return fcanon<F>(mont_mul<F>(x, one));
}

// ---- conversions between 8 x 32-bit words (little-endian) and 9 x 29-bit limbs ----
DEV fe fe_from_words(const u32 w[8]) {
  fe r;
#pragma unroll
  for (int k = 0; k < NL; k++) {
    const int bit = SIGOPS_LIMB_BITS * k;
    const int wi = bit >> 5, sh = bit & 31;
    u64 x = (u64)w[wi] >> sh;
    if (wi + 1 < 8) x |= (u64)w[wi + 1] << (32 - sh);
    r.v[k] = (u32)x & LMASK;
  }
  return r;
}

DEV void fe_to_words(const fe& a, u32 w[8]) {  // a < 2^256
#pragma unroll
  for (int i = 0; i < 8; i++) {
    const int bit = 32 * i;
    const int k = bit / SIGOPS_LIMB_BITS, sh = bit % SIGOPS_LIMB_BITS;
    u64 x = (u64)a.v[k] >> sh;
    if (k + 1 < NL) x |= (u64)a.v[k + 1] << (SIGOPS_LIMB_BITS - sh);
    if (k + 2 < NL) x |= (u64)a.v[k + 2] << (2 * SIGOPS_LIMB_BITS - sh);
    w[i] = (u32)x;
  }
}

// big-endian bytes (as 8 u32 loaded little-endian from memory) <-> words
DEV u32 bswap32(u32 x) { return __builtin_bswap32(x); }

DEV void words_from_be_bytes(const u32 raw[8], u32 w[8]) {
#pragma unroll
  for (int i = 0; i < 8; i++) w[i] = bswap32(raw[7 - i]);
}

DEV void words_to_be_bytes(const u32 w[8], u32 raw[8]) {
#pragma unroll
  for (int i = 0; i < 8; i++) raw[i] = bswap32(w[7 - i]);
}

// word compare a < b (little-endian 8 x u32)
DEV bool words_lt(const u32 a[8], const u32 b[8]) {
  bool lt = false, done = false;
#pragma unroll
  for (int i = 7; i >= 0; i--) {
    if (!done && a[i] != b[i]) {
      lt = a[i] < b[i];
      done = true;
    }
  }
  return lt;
}

DEV bool words_is_zero(const u32 a[8]) {
  u32 z = 0;
#pragma unroll
  for (int i = 0; i < 8; i++) z |= a[i];
  return z == 0;
}

// r = a - b (a >= b assumed), returns borrow
DEV u32 words_sub(const u32 a[8], const u32 b[8], u32 r[8]) {
  u32 br = 0;
#pragma unroll
  for (int i = 0; i < 8; i++) {
    const u64 x = (u64)a[i] - b[i] - br;
    r[i] = (u32)x;
    br = (u32)(x >> 63);
  }
  return br;
}

// ---- exponentiation by a fixed (kernel-constant) 256-bit exponent, fixed window of 3 bits ----
// The window value is uniform across the wavefront, so selecting the table entry is a scalar
// branch, not divergence.
template <class F>
DEV fe fpow_const(const fe& x, const u32 (&e)[8]) {
  fe t[7];  // t[i] = x^(i+1)
  t[0] = x;
  t[1] = mont_sqr<F>(x);
#pragma unroll
  for (int i = 2; i < 7; i++) t[i] = mont_mul<F>(t[i - 1], x);
  fe r = fe_from_const(F::ONE);
  // 256 bits = 85 windows of 3 + 1 leading bit (bit 255)
  {
    const u32 top = (e[7] >> 31) & 1u;
    if (top) r = x;
  }
  for (int w = 84; w >= 0; w--) {
    r = mont_sqr<F>(r);
    r = mont_sqr<F>(r);
    r = mont_sqr<F>(r);
    const int bit = 3 * w;
    const int wi = bit >> 5, sh = bit & 31;
    u32 d = e[wi] >> sh;
    if (sh > 29 && wi + 1 < 8) d |= e[wi + 1] << (32 - sh);
    d &= 7u;
    if (d) {
      fe m = t[0];
#pragma unroll
      for (int i = 1; i < 7; i++)
        if (d == (u32)(i + 1)) m = t[i];
      r = mont_mul<F>(r, m);
    }
  }
  return r;
}

template <class F>
DEV fe finv(const fe& x) {  // x != 0; Fermat
  return fpow_const<F>(x, F::EXP_INV);
}

}  // namespace sigops
