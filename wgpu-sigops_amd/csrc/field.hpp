// field.hpp -- 256-bit modular arithmetic for gfx950 (CDNA4), one element per lane.
//
// Replaces the reference's WGSL limb code: bigint.wgsl (add/sub/compare, :25-141), ff.wgsl
// (Barrett ff_mul :169-198, ff_inverse :60-120, ff_add/ff_sub :6-48), mont.wgsl (mont_mul :5-70,
// modpow :80-99). The reference packs 20 x 13-bit limbs because WGSL has no 64-bit multiply; here
// the limb width is chosen for the hardware instead:
//
//   * 9 limbs of 29 bits (261 bits), Montgomery form with R = 2^261 for every modulus.
//   * A product a_i*b_j goes into a 64-bit column accumulator with ONE v_mad_u64_u32 whose 64-bit
//     addend and result are the same even-aligned VGPR pair, so products issue no carry or move
//     instructions. (With 8 x 32-bit limbs every product needs a carry add plus moves to rebuild
//     aligned pairs: ~2x the VALU work.) The 3 bits of headroom per limb also let sums of a few
//     values feed a multiplication unreduced.
//   * Two representations:
//       - normalised ("tight"): limbs 0..7 < 2^29, value < 2M. Every multiplication returns this
//         form; points stored in memory use it.
//       - lazy: limbwise sums, differences (against a biased multiple of M) and small multiples,
//         with no carries (one VOP2 instruction per limb). fnorm_red() brings a lazy value back
//         to normalised limbs and a value just above M with one fold of the top bits and one
//         carry chain. On gfx950 a normalised modular add costs ~14x a lazy one.
//     Which lazy values may feed a multiplication is fixed by two bounds (column sums below 2^64,
//     a*b below M*2^261); the point formulas state the bounds of each intermediate. Host builds
//     with SIGOPS_HOST_CHECK=1 carry the worst-case bounds through every operation and check each
//     precondition, so one run of a formula verifies its bounds for all inputs.
#pragma once
#include <cstdint>
#include <type_traits>

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

#if defined(SIGOPS_HOST_CHECK) && SIGOPS_HOST_CHECK && !defined(__HIP_DEVICE_COMPILE__)
#define SIGOPS_BOUNDS 1
#include <execinfo.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#else
#define SIGOPS_BOUNDS 0
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
#if SIGOPS_BOUNDS
  // worst-case bounds (host check builds): limbs 0..7 <= lmax, limb 8 <= tmax, value < vmax.
  // The defaults describe a normalised value below 2^257 (>= 2M for every modulus here).
  u64 lmax = LMASK;
  u64 tmax = 1u << 25;
  const char* origin = "default";
  long double vmax = 231584178474632390847141970017375815706539969331281128078915168015826259279872.0L;
#endif
};

#if SIGOPS_BOUNDS
[[noreturn]] inline void bound_fail(const char* what) {
  fprintf(stderr, "sigops: bound check failed: %s\n", what);
  void* frames[32];
  backtrace_symbols_fd(frames, backtrace(frames, 32), 2);
  abort();
}
// switched off only by tests that feed values outside every invariant on purpose
inline bool& bounds_enabled() {
  static bool on = true;
  return on;
}
#define BCHECK(cond, what)                            \
  do {                                                \
    if (bounds_enabled() && !(cond)) bound_fail(what); \
  } while (0)
template <class F>
inline long double mod_ld() {
  long double x = 0;
  for (int k = NL - 1; k >= 0; k--) x = ldexpl(x, 29) + F::M[k];
  return x;
}
inline long double fe_ld(const fe& a) {
  long double x = 0;
  for (int k = NL - 1; k >= 0; k--) x = ldexpl(x, 29) + a.v[k];
  return x;
}
inline long double p2(int e) { return ldexpl(1.0L, e); }
// the actual limbs and value respect the recorded bounds
inline void bcheck_actual(const fe& a, const char* what) {
  for (int k = 0; k < NL - 1; k++) BCHECK(a.v[k] <= a.lmax, what);
  BCHECK(a.v[NL - 1] <= a.tmax, what);
  BCHECK(fe_ld(a) < a.vmax * (1.0L + 1e-15L), what);
}
inline u64 fe_lim(const fe& a) { return a.lmax > a.tmax ? a.lmax : a.tmax; }
// limb 8 of a value below vmax is below vmax / 2^232 (the lower limbs are non-negative)
inline u64 top_from_value(long double vmax) { return (u64)(vmax / p2(232)); }
#define BOUNDS_SET(x, l, t, vv)                                    \
  do {                                                             \
    (x).lmax = (l);                                                \
    (x).vmax = (vv);                                               \
    const u64 tv_ = top_from_value((x).vmax);                      \
    (x).tmax = (u64)(t) < tv_ ? (u64)(t) : tv_;                    \
    (x).origin = __func__;                                         \
  } while (0)
#else
#define BCHECK(cond, what) \
  do {                     \
  } while (0)
#define BOUNDS_SET(x, l, t, vv) \
  do {                          \
  } while (0)
#endif

DEV fe fe_from_const(const u32 (&c)[NL]) {
  fe r;
#pragma unroll
  for (int k = 0; k < NL; k++) r.v[k] = c[k];
#if SIGOPS_BOUNDS
  u64 lm = 0;
  for (int k = 0; k < NL - 1; k++) lm = c[k] > lm ? c[k] : lm;
  BOUNDS_SET(r, lm, c[NL - 1], fe_ld(r) + 1.0L);
#endif
  return r;
}

DEV fe fe_zero() {
  fe r;
#pragma unroll
  for (int k = 0; k < NL; k++) r.v[k] = 0;
  BOUNDS_SET(r, 0, 0, 1.0L);
  return r;
}

// ---- Montgomery multiplication ----
// Operand scanning with interleaved reduction: after row i is accumulated, column i is final (it
// only lacks the carry of column i-1), so m_i = col_i * N0 mod 2^29 and adding m_i*M zeroes its
// low 29 bits. Two forms of the reduction step:
//   * dense moduli (the group orders): m_i * M_j for the nine limbs of M, all unsigned; a column
//     holds at most 9 a*b + 9 m*M products.
//   * sparse moduli (the three field primes, F::SPARSE): M = sum_t SP_COEF[t] * 2^(29 SP_OFF[t])
//     with 2 (2^255-19), 3 (secp256k1 p) or 5 (P-256 p) signed terms, so a step is 1..4
//     multiply-adds instead of 9. Only SP_COEF[0] sits at column i, so M = SP_COEF[0] (mod 2^29)
//     and N0 = -SP_COEF[0]^-1 (mod 2^29). A coefficient -1 at column i costs nothing: m is then
//     col_i mod 2^29 and subtracting it only clears the bits the carry shift drops.
//     Negative terms (-977 m, -8 m for secp256k1 p; -19 m for 2^255-19; magnitude < 2^40) are
//     absorbed by column offsets: column 0 starts at COL_E, columns 1..15 at COL_E - COL_E/2^29 and
//     the top column at -COL_E/2^29, so with the +COL_E/2^29 that each incoming carry brings,
//     every column below the top is held as its true value + COL_E >= 0 while it is shifted (a
//     negative term or a negative carry from below never exceeds COL_E), and the offsets cancel.
//     The top column needs none: the result is non-negative, so its final value is too. Columns
//     stay unsigned and carries are plain shifts. P-256 p has no negative term and no offsets.
// Column bound: 9 * A * B + (positive reduction terms + offset + carry) < 2^64 for operand limbs
// <= A, B; the second part is below 2^54 for secp256k1 p and 2^255-19 and below 2^58.1 for P-256.
// Output: (a*b + m*M) / 2^261 < a*b/2^261 + M, so it is < 2M whenever a*b < M * 2^261.
constexpr u64 COL_E = 1ull << 40;

template <class F>
DEV constexpr bool has_neg_terms() {  // a reduction term other than a column-0 coefficient -1 is negative
  for (int t = 0; t < F::SP_N; t++)
    if (F::SP_COEF[t] < 0 && !(t == 0 && F::SP_COEF[t] == -1)) return true;
  return false;
}

template <class F>
DEV constexpr u64 col_init(int k) {
  if constexpr (!F::SPARSE) {
    return 0;
  } else {
    if (!has_neg_terms<F>()) return 0;
    if (k == 2 * NL - 2) return (u64)0 - (COL_E >> SIGOPS_LIMB_BITS);
    return k == 0 ? COL_E : COL_E - (COL_E >> SIGOPS_LIMB_BITS);
  }
}

#if defined(__HIP_DEVICE_COMPILE__)
// Column offsets read from constant memory: the compiler does not fold the load of an externally
// visible object, so each offset stays a 64-bit scalar operand of its column's first multiply-add
// instead of becoming a separate 64-bit add after it (16 fewer instructions per secp256k1 /
// 2^255-19 product). With internal linkage (`static`) the compiler proves the array constant and
// folds it back into immediates, so it stays extern; each code object holds its own copy.
__constant__ u64 g_col_offsets[3] = {COL_E, COL_E - (COL_E >> SIGOPS_LIMB_BITS),
                                     (u64)0 - (COL_E >> SIGOPS_LIMB_BITS)};
#endif

// The sparse coefficients of the three field primes in one table (secp256k1 p's, then P-256 p's,
// then 2^255-19's). On the device every reduction term reads its coefficient from the extern
// constant copy, for the same reason as the column offsets: folded into an immediate, a
// power-of-two coefficient becomes a 64-bit shift plus add (or a 2-instruction subtract) and
// 2^24-1 a shift and subtract; read from memory, each term stays one multiply-add with a 32-bit
// scalar operand, and the values cost no registers between products.
template <class F>
DEV constexpr int sp_base() {
  if constexpr (std::is_same<F, ModPk1>::value) return 0;
  if constexpr (std::is_same<F, ModPr1>::value) return ModPk1::SP_N;
  if constexpr (std::is_same<F, ModPed>::value) return ModPk1::SP_N + ModPr1::SP_N;
  return -1;
}
constexpr int SP_TABLE_N = ModPk1::SP_N + ModPr1::SP_N + ModPed::SP_N;
struct SpTable {
  i32 v[SP_TABLE_N];
};
constexpr SpTable sp_table() {
  SpTable t{};
  for (int k = 0; k < ModPk1::SP_N; k++) t.v[sp_base<ModPk1>() + k] = ModPk1::SP_COEF[k];
  for (int k = 0; k < ModPr1::SP_N; k++) t.v[sp_base<ModPr1>() + k] = ModPr1::SP_COEF[k];
  for (int k = 0; k < ModPed::SP_N; k++) t.v[sp_base<ModPed>() + k] = ModPed::SP_COEF[k];
  return t;
}
#if defined(__HIP_DEVICE_COMPILE__)
__constant__ SpTable g_sp_coef = sp_table();
#endif

template <class F>
DEV void mont_acc_init(u64 (&acc)[2 * NL - 1]) {
#if defined(__HIP_DEVICE_COMPILE__)
  if constexpr (F::SPARSE && has_neg_terms<F>()) {
    const u64 e0 = g_col_offsets[0], e1 = g_col_offsets[1], et = g_col_offsets[2];
#pragma unroll
    for (int k = 0; k < 2 * NL - 1; k++) acc[k] = k == 0 ? e0 : (k == 2 * NL - 2 ? et : e1);
    return;
  }
#endif
#pragma unroll
  for (int k = 0; k < 2 * NL - 1; k++) acc[k] = col_init<F>(k);
}

template <class F>
DEV void mont_step(u64 (&acc)[2 * NL - 1], int i) {
  if (i > 0) acc[i] += acc[i - 1] >> SIGOPS_LIMB_BITS;
  const u32 m = ((u32)acc[i] * F::N0) & LMASK;
  if constexpr (F::SPARSE) {
#pragma unroll
    for (int t = 0; t < F::SP_N; t++) {
      i32 coef = F::SP_COEF[t];
      // coefficient -1 at column i: subtracting m = col_i mod 2^29 only clears the low bits that
      // the carry shift discards anyway
      if (t == 0 && coef == -1) continue;
#if defined(__HIP_DEVICE_COMPILE__)
      static_assert(sp_base<F>() >= 0, "a sparse modulus outside the coefficient table");
      coef = g_sp_coef.v[sp_base<F>() + t];  // the same value, kept out of the compiler's folding
#endif
      if (F::SP_COEF[t] < 0)
        acc[i + F::SP_OFF[t]] += (u64)((i64)coef * (i64)(i32)m);
      else
        acc[i + F::SP_OFF[t]] += (u64)(u32)coef * m;
    }
  } else {
#pragma unroll
    for (int j = 0; j < NL; j++) acc[i + j] += (u64)m * F::M[j];
  }
}

template <class F>
DEV fe mont_final(const u64 (&acc)[2 * NL - 1]) {
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

#if SIGOPS_BOUNDS
template <class F>
inline void mont_bounds(const fe& a, const fe& b, long double col, fe& r) {
  bcheck_actual(a, "mont input a");
  bcheck_actual(b, "mont input b");
  // column sums: products plus reduction terms, offsets and carries
  long double red = p2(36) + (F::SPARSE ? (long double)COL_E : 9.0L * (long double)LMASK * LMASK);
  if constexpr (F::SPARSE)
    for (int t = 0; t < F::SP_N; t++)
      if (F::SP_COEF[t] > 0) red += (long double)F::SP_COEF[t] * LMASK;
  BCHECK(col + red < p2(64), "mont column bound");
  const long double M = mod_ld<F>();
  BCHECK(a.vmax * b.vmax < M * p2(261), "mont value bound");
  const long double vo = a.vmax * b.vmax / p2(261) + M;
  if (getenv("SIGOPS_BDEBUG") && vo > 1.5L * M)
    fprintf(stderr, "mont: a %.4Lf M (lmax %llu) x b %.4Lf M (lmax %llu) -> %.4Lf M\n", a.vmax / M,
            (unsigned long long)a.lmax, b.vmax / M, (unsigned long long)b.lmax, vo / M);
  r.lmax = LMASK;
  r.tmax = (u64)(vo / p2(232)) + 1;
  r.vmax = vo;
  r.origin = "mont";
  bcheck_actual(r, "mont output");
}
#endif

// a*b*2^-261 mod M; inputs within the column and value bounds above; output normalised, < 2M
template <class F>
DEV_MUL fe mont_mul(const fe& a, const fe& b) {
  u64 acc[2 * NL - 1];
  mont_acc_init<F>(acc);
#pragma unroll
  for (int i = 0; i < NL; i++) {
#pragma unroll
    for (int j = 0; j < NL; j++) acc[i + j] += (u64)a.v[i] * b.v[j];
    mont_step<F>(acc, i);
  }
  fe r = mont_final<F>(acc);
#if SIGOPS_BOUNDS
  mont_bounds<F>(a, b, 9.0L * fe_lim(a) * fe_lim(b), r);
#endif
  return r;
}

// Montgomery square: 45 instead of 81 partial products; the cross terms use a pre-doubled
// operand, so column 8 holds 4 cross products a_i*2a_j and one square: 9 * A^2 in all.
template <class F>
DEV_MUL fe mont_sqr(const fe& a) {
  u32 a2[NL];
#pragma unroll
  for (int k = 0; k < NL; k++) a2[k] = a.v[k] << 1;
  u64 acc[2 * NL - 1];
  mont_acc_init<F>(acc);
#pragma unroll
  for (int i = 0; i < NL; i++) {
    acc[2 * i] += (u64)a.v[i] * a.v[i];
#pragma unroll
    for (int j = i + 1; j < NL; j++) acc[i + j] += (u64)a.v[i] * a2[j];
    mont_step<F>(acc, i);
  }
  fe r = mont_final<F>(acc);
#if SIGOPS_BOUNDS
  BCHECK(fe_lim(a) < (1ull << 31), "sqr doubled limb");
  mont_bounds<F>(a, a, 9.0L * fe_lim(a) * fe_lim(a), r);
#endif
  return r;
}

// ---- selects without v_cndmask: a lane mask (0 or ~0) and bitfield inserts ----
DEV u32 lane_mask(bool c) { return 0u - (u32)c; }

DEV fe fsel_mask(u32 mask, const fe& a, const fe& b) {  // mask ? a : b, limbwise
  fe r;
#pragma unroll
  for (int k = 0; k < NL; k++) r.v[k] = (a.v[k] & mask) | (b.v[k] & ~mask);
#if SIGOPS_BOUNDS
  BOUNDS_SET(r, a.lmax > b.lmax ? a.lmax : b.lmax, a.tmax > b.tmax ? a.tmax : b.tmax, a.vmax > b.vmax ? a.vmax : b.vmax);
#endif
  return r;
}

DEV fe fsel(bool c, const fe& a, const fe& b) { return fsel_mask(lane_mask(c), a, b); }

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
  BCHECK(s.lmax <= LMASK && s.vmax <= 4.0L * mod_ld<F>(), "reduce_2m input must be normalised < 4M");
  fe r = fsel_mask((u32)br, s, t);  // borrow out: s < 2M, keep it
  BOUNDS_SET(r, LMASK, LMASK, 2.0L * mod_ld<F>());
  return r;
}

template <class F>
DEV fe fadd(const fe& a, const fe& b) {
  BCHECK(a.lmax <= LMASK && b.lmax <= LMASK && a.vmax <= 2.0L * mod_ld<F>() && b.vmax <= 2.0L * mod_ld<F>(),
         "fadd inputs must be normalised < 2M");
  fe s;
  u32 c = 0;
#pragma unroll
  for (int k = 0; k < NL; k++) {
    const u32 x = a.v[k] + b.v[k] + c;
    s.v[k] = x & LMASK;
    c = x >> SIGOPS_LIMB_BITS;
  }
  BOUNDS_SET(s, LMASK, LMASK, a.vmax + b.vmax);
  return reduce_2m<F>(s);
}

// a - b + 2M in (0, 4M), then to [0, 2M)
template <class F>
DEV fe fsub(const fe& a, const fe& b) {
  BCHECK(a.lmax <= LMASK && b.lmax <= LMASK && a.vmax <= 2.0L * mod_ld<F>() && b.vmax <= 2.0L * mod_ld<F>(),
         "fsub inputs must be normalised < 2M");
  fe s;
  i32 c = 0;
#pragma unroll
  for (int k = 0; k < NL; k++) {
    const i32 x = (i32)a.v[k] - (i32)b.v[k] + (i32)F::M2[k] + c;
    s.v[k] = (u32)x & LMASK;
    c = x >> SIGOPS_LIMB_BITS;
  }
  BOUNDS_SET(s, LMASK, LMASK, a.vmax + 2.0L * mod_ld<F>());
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
  BCHECK(s.lmax <= LMASK && s.vmax <= 2.0L * mod_ld<F>(), "fcanon input must be normalised < 2M");
  fe t;
  i32 br = 0;
#pragma unroll
  for (int k = 0; k < NL; k++) {
    const i32 x = (i32)s.v[k] - (i32)F::M[k] + br;
    t.v[k] = (u32)x & LMASK;
    br = x >> SIGOPS_LIMB_BITS;
  }
  fe r = fsel_mask((u32)br, s, t);
  BOUNDS_SET(r, LMASK, LMASK, mod_ld<F>());
  return r;
}

// ---- lazy arithmetic (no carries, no reduction) ----
// a + b limbwise
DEV fe fadd_nr(const fe& a, const fe& b) {
  fe r;
#pragma unroll
  for (int k = 0; k < NL; k++) r.v[k] = a.v[k] + b.v[k];
#if SIGOPS_BOUNDS
  BOUNDS_SET(r, a.lmax + b.lmax, a.tmax + b.tmax, a.vmax + b.vmax);
  BCHECK(fe_lim(r) < (1ull << 32), "fadd_nr limb overflow");
#endif
  return r;
}

// c * a limbwise, small c
DEV fe fscale_nr(const fe& a, u32 c) {
  fe r;
#pragma unroll
  for (int k = 0; k < NL; k++) r.v[k] = a.v[k] * c;
#if SIGOPS_BOUNDS
  BOUNDS_SET(r, a.lmax * c, a.tmax * c, a.vmax * c);
  BCHECK(fe_lim(r) < (1ull << 32), "fscale_nr limb overflow");
#endif
  return r;
}

// K*M written with limbs 0..7 in [LB*2^29 - LB, (LB+1)*2^29): subtracting any b whose limbs are
// <= LB*(2^29-1) (a sum of LB normalised values) limbwise from it leaves every limb >= 0.
struct limbs9 {
  u32 v[NL];
};
template <class F, u32 K, u32 LB>
DEV constexpr limbs9 sub_bias() {
  limbs9 r{};
  u64 c = 0;
  for (int k = 0; k < NL; k++) {
    const u64 x = (u64)F::M[k] * K + c;
    r.v[k] = (u32)x & LMASK;
    c = x >> SIGOPS_LIMB_BITS;
  }
  r.v[NL - 1] += (u32)(c << SIGOPS_LIMB_BITS);
  for (int k = 0; k < NL - 1; k++) {
    r.v[k] += LB << SIGOPS_LIMB_BITS;
    r.v[k + 1] -= LB;
  }
  return r;
}

// a - b + K*M limbwise. b's limbs must be <= LB*(2^29-1) (K*M's top limb - LB for limb 8)
template <class F, u32 K, u32 LB>
DEV fe fsub_nr(const fe& a, const fe& b) {
  constexpr limbs9 B = sub_bias<F, K, LB>();
  fe r;
#pragma unroll
  for (int k = 0; k < NL; k++) r.v[k] = a.v[k] + (B.v[k] - b.v[k]);
#if SIGOPS_BOUNDS
  u64 bmin = ~0ull, bmax = 0;
  for (int k = 0; k < NL - 1; k++) {
    bmin = B.v[k] < bmin ? B.v[k] : bmin;
    bmax = B.v[k] > bmax ? B.v[k] : bmax;
  }
  if (bounds_enabled() && !(b.lmax <= bmin && b.tmax <= B.v[NL - 1])) {
    fprintf(stderr, "fsub_nr<K=%u, LB=%u>: b.lmax %llu (bias min %llu), b.tmax %llu (bias top %u), b.vmax %.6Lf M\n",
            K, LB, (unsigned long long)b.lmax, (unsigned long long)bmin, (unsigned long long)b.tmax, B.v[NL - 1],
            b.vmax / mod_ld<F>());
    fprintf(stderr, "  b from %s\n", b.origin);
    bound_fail("fsub_nr bias too small for b");
  }
  BOUNDS_SET(r, a.lmax + bmax, a.tmax + B.v[NL - 1], a.vmax + K * mod_ld<F>());
  BCHECK(fe_lim(r) < (1ull << 32), "fsub_nr limb overflow");
#endif
  return r;
}

// Lazy value (limbs < 2^32 - 2^29 - 2^5) -> normalised limbs, value < (1 + 2^-19) M, same residue.
// Fold: q = top bits at 2^TOP (TOP = 256, or 255 for 2^255-19) read from limb 8 alone, which
// underestimates floor(v / 2^TOP), so v - q*M >= 0; v - q*M < 2^TOP + 2^235.1 + q*(2^TOP mod M).
//   secp256k1 p: 2^256 = 2^32 + 977 (mod p)  -> limb0 += 977 q, limb1 += 8 q
//   2^255 - 19:  2^255 = 19 (mod p)          -> limb0 += 19 q
//   P-256 p:     2^256 = 2^224 - 2^192 - 2^96 + 1 (mod p); the two negative terms borrow
//                2^29 at limbs 3..6 from one unit of limb 7 (only when q > 0, when limb 7 also
//                receives q * 2^21), so no limb goes negative.
// Then one carry chain.
template <class F>
DEV fe fnorm_red(const fe& a) {
  static_assert(F::SPARSE, "fnorm_red is defined for the field primes");
  u32 l[NL];
#pragma unroll
  for (int k = 0; k < NL; k++) l[k] = a.v[k];
  if constexpr (F::SP_N == 2) {  // 2^255 - 19
    const u32 q = l[8] >> 23;
    l[8] &= (1u << 23) - 1u;
    l[0] += 19u * q;
  } else if constexpr (F::SP_N == 3) {  // secp256k1 p
    const u32 q = l[8] >> 24;
    l[8] &= (1u << 24) - 1u;
    l[0] += 977u * q;
    l[1] += 8u * q;
  } else {  // P-256 p
    const u32 q = l[8] >> 24;
    const u32 s = q < 1u ? q : 1u;
    const u32 s29 = s << 29;
    l[8] &= (1u << 24) - 1u;
    l[0] += q;
    l[3] += s29 - (q << 9);
    l[4] += s29 - s;
    l[5] += s29 - s;
    l[6] += s29 - s - (q << 18);
    l[7] += (q << 21) - s;
  }
  fe r;
  u32 c = 0;
#pragma unroll
  for (int k = 0; k < NL - 1; k++) {
    const u32 x = l[k] + c;
    r.v[k] = x & LMASK;
    c = x >> SIGOPS_LIMB_BITS;
  }
  r.v[NL - 1] = l[NL - 1] + c;
#if SIGOPS_BOUNDS
  bcheck_actual(a, "fnorm_red input");
  BCHECK(fe_lim(a) <= (1ull << 32) - (1ull << 29) - 64, "fnorm_red input limbs");
  BOUNDS_SET(r, LMASK, (F::SP_N == 2 ? (1ull << 23) : (1ull << 24)) + 64, mod_ld<F>() * (1.0L + 1.0L / 524288.0L));
  bcheck_actual(r, "fnorm_red output");
#endif
  return r;
}

// x == 0 mod M for x in [0, 2M)
template <class F>
DEV bool fis_zero(const fe& x) {
  BCHECK(x.lmax <= LMASK && x.vmax <= 2.0L * mod_ld<F>(), "fis_zero input must be normalised < 2M");
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

template <class F>
DEV fe to_mont(const fe& x) {  // x any value < 2^261 with normalised limbs, x * R2 < M * 2^261
  return mont_mul<F>(x, fe_from_const(F::R2));
}

template <class F>
DEV fe from_mont(const fe& x) {  // canonical plain value
  fe one = fe_zero();
  one.v[0] = 1;
  BOUNDS_SET(one, 1, 0, 2.0L);
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
  BOUNDS_SET(r, LMASK, (1u << 24) - 1u, p2(256));
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

// the low 32 bits of (hi:lo) >> sh, sh in {0, 8, 16, 24} (v_alignbit_b32)
DEV u32 funnel32(u32 hi, u32 lo, u32 sh) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_alignbit(hi, lo, sh);
#else
  return (u32)((((u64)hi << 32) | lo) >> sh);
#endif
}

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

// r = a + b, returns the carry
DEV u32 words_add(const u32 a[8], const u32 b[8], u32 r[8]) {
  u32 c = 0;
#pragma unroll
  for (int i = 0; i < 8; i++) {
    const u64 x = (u64)a[i] + b[i] + c;
    r[i] = (u32)x;
    c = (u32)(x >> 32);
  }
  return c;
}

// ---- exponentiation by a fixed exponent along an addition chain ----
// CH (generated into consts.hpp) lists steps t[DST] = t[SRC]^(2^NSQ) * t[MUL] (MUL < 0: no
// multiplication) with t[0] = x on entry. The steps are expanded at compile time, so every slot
// index is a constant and the slots stay in registers; only the squaring runs loop.
template <class F, class CH, int S>
DEV void chain_steps(fe* t) {
  if constexpr (S < CH::N) {
    fe v = t[CH::SRC[S]];
#pragma unroll 1
    for (int k = 0; k < CH::NSQ[S]; k++) v = mont_sqr<F>(v);
    if constexpr (CH::MUL[S] >= 0) v = mont_mul<F>(v, t[CH::MUL[S]]);
    t[CH::DST[S]] = v;
    chain_steps<F, CH, S + 1>(t);
  }
}

template <class F, class CH>
DEV fe fpow_chain(const fe& x) {
  fe t[CH::SLOTS];
  t[0] = x;
  chain_steps<F, CH, 0>(t);
  return t[CH::RESULT];
}

// x^(M-2) (Fermat inverse) for each modulus
template <class F>
struct InvChain;
template <>
struct InvChain<ModPk1> {
  typedef CH_K1_INV type;
};
template <>
struct InvChain<ModNk1> {
  typedef CH_NK1_INV type;
};
template <>
struct InvChain<ModPr1> {
  typedef CH_R1_INV type;
};
template <>
struct InvChain<ModNr1> {
  typedef CH_NR1_INV type;
};
template <>
struct InvChain<ModPed> {
  typedef CH_ED_INV type;
};
template <>
struct InvChain<ModL> {
  typedef CH_L_INV type;
};

// Fermat inverse x^(M-2); the kernels invert with divsteps (inv.hpp), this form checks it
template <class F>
DEV fe finv_fermat(const fe& x) {  // x != 0
  return fpow_chain<F, typename InvChain<F>::type>(x);
}

}  // namespace sigops
