// devcheck.cpp -- host (g++) build of the device headers in wgpu-sigops_amd/csrc, exposing the
// per-item algorithms through C functions so that CPU tests can compare them with the oracle
// before anything runs on a GPU. Test infrastructure only; the product runs the HIP kernels.
#include <cstring>

#include "ed25519.hpp"
#include "gen.hpp"
#include "secp.hpp"
#include "sha256.hpp"
#include "table.hpp"

using namespace sigops;

namespace {
// the generator tables hold canonical values (< p); the bound checker is told so
template <class F>
void canonical_bounds(fe& x) {
#if SIGOPS_BOUNDS
  BOUNDS_SET(x, LMASK, F::M[NL - 1], mod_ld<F>());
  bcheck_actual(x, "table entry");
#else
  (void)x;
#endif
}
template <class F>
struct GTabHost {
  const u32* base;
  void load(u32 idx, fe& x, fe& y) const {
    const u32* e = base + idx * 20;
    for (int k = 0; k < NL; k++) {
      x.v[k] = e[k];
      y.v[k] = e[9 + k];
    }
    canonical_bounds<F>(x);
    canonical_bounds<F>(y);
  }
  // the LDS staging's product (stage_gtab): beta times the canonical x
  void load_phi(u32 idx, fe& bx, fe& y) const {
    fe x;
    load(idx, x, y);
    bx = mont_mul<F>(x, fe_from_const(K1_BETA));
  }
};
struct BTabHost {
  void load(u32 idx, fe& ypx, fe& ymx, fe& t2d) const {
    const u32* e = ED_BTABLE + idx * 28;
    for (int k = 0; k < NL; k++) {
      ypx.v[k] = e[k];
      ymx.v[k] = e[9 + k];
      t2d.v[k] = e[18 + k];
    }
    canonical_bounds<ModPed>(ypx);
    canonical_bounds<ModPed>(ymx);
    canonical_bounds<ModPed>(t2d);
  }
};
struct STabHost {
  jac t[8];
  fe before[8];
  fe ax[8], ay[8], abx[8];
  void store_full(int e, const jac& p, const fe& b) { t[e] = p; before[e] = b; }
  void load_full(u32 e, jac& p, fe& b) const { p = t[e]; b = before[e]; }
  void store_affine(int e, const fe& x, const fe& y) { ax[e] = x; ay[e] = y; }
  void load_affine(u32 e, fe& x, fe& y) const { x = ax[e]; y = ay[e]; }
  void store_affine_phi(int e, const fe& x, const fe& y, const fe& bx) { ax[e] = x; ay[e] = y; abx[e] = bx; }
  void load_affine_phi(u32 e, fe& bx, fe& y) const { bx = abx[e]; y = ay[e]; }
};
struct CTabHost {
  cached t[16];
  void store(int e, const cached& p) { t[e] = p; }
  cached load(u32 e) const { return t[e]; }
};
struct MemHost {
  const u32* w;
  u32 word(u64 i) const { return w[i]; }
};

template <class C>
int recover_one(const unsigned char* sig, const unsigned char* dig, unsigned char* out, const u32* gtab,
                int force_careful) {
  u32 sr[16], dr[8], o[16];
  memcpy(sr, sig, 64);
  memcpy(dr, dig, 32);
  secp_pro pro;
  secp_prologue<C>(sr, dr, pro);
  GTabHost<typename C::Fp> gt{gtab};
  memset(out, 0, 64);
  if (pro.status != ST_OK) return (int)pro.status;
  bool exc = false;
  STabHost tab;
  jac Q = (force_careful & 2) ? secp_ladder<C, 9>(pro, tab, gt, exc) : secp_ladder<C>(pro, tab, gt, exc);
  int careful = exc || (force_careful & 1);
  if (!careful) {
    secp_output<C>(Q, o);
    memcpy(out, o, 64);
    return 0;
  }
  jacf Qc = secp_ladder_careful<C>(pro, gt);
  if (Qc.inf) return (int)ST_INFINITY;
  secp_output<C>(Qc.P, o);
  memcpy(out, o, 64);
  return exc ? 0x100 : 0;  // flag that the careful path ran (status OK)
}
template <class C>
int verify_one(const unsigned char* sig, const unsigned char* dig, const unsigned char* pub, const u32* gtab,
               int force_careful) {
  u32 sr[16], dr[8], pr[16];
  memcpy(sr, sig, 64);
  memcpy(dr, dig, 32);
  memcpy(pr, pub, 64);
  secp_pro pro;
  secp_prologue_verify<C>(sr, dr, pr, pro);
  GTabHost<typename C::Fp> gt{gtab};
  if (pro.status != ST_OK) return (int)pro.status;
  bool exc = false;
  STabHost tab;
  jac P = (force_careful & 2) ? secp_ladder<C, 9>(pro, tab, gt, exc) : secp_ladder<C>(pro, tab, gt, exc);
  if (!exc && !(force_careful & 1)) return (int)secp_verify_x<C>(P, sr);
  jacf Pc = secp_ladder_careful<C>(pro, gt);
  const int st = Pc.inf ? (int)ST_INFINITY : (int)secp_verify_x<C>(Pc.P, sr);
  return st | (exc ? 0x100 : 0);
}
// the 128-lane workgroup of kern_table.hip, lane by lane: check entries 0..15, then build
template <class C>
u32 table_secp_host(const u32* limbs, u32 lbits, u32* out) {
  typedef typename C::Fp F;
  const u32 nl = 256 / lbits + 1;
  fe xs[16], ys[16];
  u32 bad = 0;
  for (u32 j = 0; j < 16; j++) tab_secp_check_entry<C>(j, limbs, nl, lbits, xs[j], ys[j], bad);
  u32 p0[2 * NL];
  const fe xc = fcanon<F>(xs[0]), yc = fcanon<F>(ys[0]);
  for (int k = 0; k < NL; k++) {
    p0[k] = xc.v[k];
    p0[NL + k] = yc.v[k];
  }
  for (u32 j = 0; j < 128; j++)
    tab_secp_build_entry<C>(j, p0, j < 16 ? xs[j] : fe_zero(), j < 16 ? ys[j] : fe_zero(), bad, out + 20 * j);
  return bad;
}

u32 table_ed_host(const u32* limbs, u32 lbits, u32* out) {
  const u32 nl = 256 / lbits + 1;
  fe xs[16], ys[16];
  u32 bad = 0;
  for (u32 j = 0; j < 16; j++) tab_ed_check_entry(j, limbs, nl, lbits, xs[j], ys[j], bad);
  u32 p0[2 * NL];
  const fe xc = fcanon<ModPed>(xs[0]), yc = fcanon<ModPed>(ys[0]);
  for (int k = 0; k < NL; k++) {
    p0[k] = xc.v[k];
    p0[NL + k] = yc.v[k];
  }
  for (u32 j = 0; j < 128; j++)
    tab_ed_build_entry(j, p0, j < 16 ? xs[j] : fe_zero(), j < 16 ? ys[j] : fe_zero(), bad, out + 28 * j);
  for (u32 j = 0; j < 128; j++) tab_ed_build_hi_entry(j, p0, out + 28 * (128 + j));
  return bad;
}
}  // namespace

extern "C" {
// caller-table conversion (host build of kern_table.hip): returns the TAB_BAD_* bits, writes the
// 128-entry table (20 words per secp entry), or for Ed25519 (28 words per entry) the 128 odd
// multiples of P0 followed by the 128 of [2^136] P0; which: 0 k1, 1 r1, 2 ed25519
unsigned dc_table(int which, const u32* limbs, unsigned lbits, u32* out) {
  if (which == 0) return table_secp_host<CurveK1>(limbs, lbits, out);
  if (which == 1) return table_secp_host<CurveR1>(limbs, lbits, out);
  return table_ed_host(limbs, lbits, out);
}

// the resident generator tables as uploaded to HBM (same layout as dc_table's output; the secp
// tables hold 256 entries, of which 8-bit windows use the first 128)
void dc_resident_table(int which, u32* out) {
  if (which == 0) memcpy(out, K1_GTABLE, sizeof K1_GTABLE);
  else if (which == 1) memcpy(out, R1_GTABLE, sizeof R1_GTABLE);
  else memcpy(out, ED_BTABLE, sizeof ED_BTABLE);
}

// verification with a known key: status; bit 8 set when the fast ladder met an exceptional addition.
// flags (also of dc_secp_recover): bit 0 forces the careful path, bit 1 selects 9-bit G windows
int dc_secp_verify(int curve, const unsigned char* sig, const unsigned char* dig, const unsigned char* pub,
                   int flags) {
  if (curve == 0) return verify_one<CurveK1>(sig, dig, pub, K1_GTABLE, flags);
  return verify_one<CurveR1>(sig, dig, pub, R1_GTABLE, flags);
}

// returns status; bit 8 set when the fast ladder met an exceptional addition
int dc_secp_recover(int curve, const unsigned char* sig, const unsigned char* dig, unsigned char* out,
                    int flags) {
  if (curve == 0) return recover_one<CurveK1>(sig, dig, out, K1_GTABLE, flags);
  return recover_one<CurveR1>(sig, dig, out, R1_GTABLE, flags);
}

// the lattice reduction of the Ed25519 check: returns 1 when a pair within the bounds was found;
// c0 (5 words, magnitude), c1 (5 words), *c0_neg
int dc_ed_lattice(const u32* k, u32* c0, u32* c1, int* c0_neg) {
  bool neg = false;
  const bool ok = ed_lattice(k, c0, c1, neg);
  *c0_neg = neg;
  return ok;
}

// 1: every Ed25519 item takes the full-length check (the lattice reduction reports failure)
void dc_set_ed_force_full(int on) { ed_force_full() = on != 0; }

// flags: bit 0 verify_strict rules, bit 1 the full-length check only (the mixed-scheme kernels'
// form); otherwise the lattice-reduced check of the Ed25519-only kernels
int dc_ed25519_verify(const unsigned char* sig, const unsigned char* pk, const unsigned char* msg, unsigned len,
                      int flags) {
  u32 sr[16], pr[8];
  memcpy(sr, sig, 64);
  memcpy(pr, pk, 32);
  // message in a word-aligned buffer with 16 bytes of slack
  u32* buf = new u32[(len + 16) / 4 + 4]();
  memcpy(buf, msg, len);
  MemHost mem{buf};
  CTabHost tab;
  BTabHost bt;
  u32 st;
  if (flags & 2)
    st = (flags & 1) ? ed25519_verify_item<true, false>(sr, pr, mem, 0, len, tab, bt)
                     : ed25519_verify_item<false, false>(sr, pr, mem, 0, len, tab, bt);
  else
    st = (flags & 1) ? ed25519_verify_item<true, true>(sr, pr, mem, 0, len, tab, bt)
                     : ed25519_verify_item<false, true>(sr, pr, mem, 0, len, tab, bt);
  delete[] buf;
  return (int)st;
}

// field op self-check: op 0 mul, 1 sqr, 2 inv, 3 add, 4 sub; mod 0..5; values as 32 BE bytes (plain)
// Raw Montgomery operations on arbitrary limb vectors (no conversion, no canonicalisation):
// op 0 = mont_mul(a, b), 1 = mont_sqr(a). Inputs and output are 9 u32 limbs.
void dc_mont_raw(int op, int mod, const u32* a, const u32* b, u32* out) {
  bounds_enabled() = false;  // raw limb vectors may lie outside the invariants the checker enforces
  fe x, y, r;
  for (int k = 0; k < NL; k++) {
    x.v[k] = a[k];
    y.v[k] = b[k];
  }
#define RAW(F) r = op == 0 ? mont_mul<F>(x, y) : mont_sqr<F>(x);
  switch (mod) {
    case 0: RAW(ModPk1) break;
    case 1: RAW(ModNk1) break;
    case 2: RAW(ModPr1) break;
    case 3: RAW(ModNr1) break;
    case 4: RAW(ModPed) break;
    default: RAW(ModL) break;
  }
#undef RAW
  for (int k = 0; k < NL; k++) out[k] = r.v[k];
  bounds_enabled() = true;
}

// SHA-256 of len bytes placed at byte offset `at` of a word buffer (exercises unaligned starts)
// memory view that records reads outside the words [lo, hi] (quad: the 16-byte chunks [lo, hi])
struct MemChecked {
  const u32* w;
  u64 lo, hi;
  mutable bool outside;
  u32 word(u64 i) const {
    if (i < lo || i > hi) {
      outside = true;
      return 0;
    }
    return w[i];
  }
  void quad(u64 c, u32* r) const {
    if (c < lo || c > hi) {
      outside = true;
      r[0] = r[1] = r[2] = r[3] = 0;
      return;
    }
    memcpy(r, w + 4 * c, 16);
  }
};

// returns -1 if the hash read a word other than those holding message bytes (for an empty message:
// the word at its offset). x4: the 16-byte-load reader, allowed the 16-byte chunks that hold
// message bytes (for an empty message: the chunk at its offset).
int dc_sha256(const unsigned char* msg, unsigned len, unsigned at, unsigned char* out, int x4) {
  u32* buf = new u32[(at + len + 16) / 4 + 8]();
  memcpy((unsigned char*)buf + at, msg, len);
  const unsigned last = at + len - (len != 0);
  MemChecked mem{buf, x4 ? at / 16u : at / 4u, x4 ? last / 16u : last / 4u, false};
  u32 d[8];
  if (x4)
    sha256_msg<true>(mem, at, len, d);
  else
    sha256_msg(mem, at, len, d);
  memcpy(out, d, 32);
  delete[] buf;
  return mem.outside ? -1 : 0;
}

void dc_sha256_64(const unsigned char* in, unsigned char* out) {
  u32 w[16], d[8];
  memcpy(w, in, 64);
  sha256_64(w, d);
  memcpy(out, d, 32);
}

// GLV split of a secp256k1 scalar k (8 little-endian words, k < n): h1, h2 as 5 words each and
// their signs (bit 0: h1 negative, bit 1: h2 negative)
int dc_glv_split(const u32* k, u32* h1, u32* h2) {
  bool n1, n2;
  glv_split(k, h1, h2, n1, n2);
  return (int)n1 | ((int)n2 << 1);
}

void dc_field(int op, int mod, const unsigned char* a, const unsigned char* b, unsigned char* out) {
  u32 ra[8], rb[8], wa[8], wb[8], wo[8], ro[8];
  memcpy(ra, a, 32);
  memcpy(rb, b, 32);
  words_from_be_bytes(ra, wa);
  words_from_be_bytes(rb, wb);
  fe x = fe_from_words(wa), y = fe_from_words(wb), r;
#define DO(F)                                              \
  {                                                        \
    fe xm = to_mont<F>(x), ym = to_mont<F>(y), rm;         \
    switch (op) {                                          \
      case 0: rm = mont_mul<F>(xm, ym); break;             \
      case 1: rm = mont_sqr<F>(xm); break;                 \
      case 2: rm = finv<F>(xm); break;                     \
      case 3: rm = fadd<F>(xm, ym); break;                 \
      case 5: rm = finv_fermat<F>(xm); break;              \
      default: rm = fsub<F>(xm, ym); break;                \
    }                                                      \
    r = from_mont<F>(rm);                                  \
  }
  switch (mod) {
    case 0: DO(ModPk1) break;
    case 1: DO(ModNk1) break;
    case 2: DO(ModPr1) break;
    case 3: DO(ModNr1) break;
    case 4: DO(ModPed) break;
    default: DO(ModL) break;
  }
  fe_to_words(r, wo);
  words_to_be_bytes(wo, ro);
  memcpy(out, ro, 32);
}

// SHA-512(hdr[0..hl) || msg) with hl = 32 or 64 and msg placed at byte offset `at` of a word
// buffer; returns -1 if the hash read a word other than those holding message bytes (for an empty
// message: the word at its offset)
int dc_sha512(const unsigned char* hdr_bytes, int hl, const unsigned char* msg, unsigned len, unsigned at,
              unsigned char* out) {
  u32 hdr[16] = {}, h[16];
  memcpy(hdr, hdr_bytes, hl);
  u32* buf = new u32[(at + len + 16) / 4 + 4]();
  memcpy((unsigned char*)buf + at, msg, len);
  MemChecked mem{buf, at / 4u, (at + len - (len != 0)) / 4u, false};
  if (hl == 64)
    sha512_ed(hdr, mem, at, len, h);
  else
    sha512_hdr<32>(hdr, mem, at, len, h);
  memcpy(out, h, 64);
  delete[] buf;
  return mem.outside ? -1 : 0;
}

void dc_gen_secp(int curve, unsigned long long seed, unsigned long long i, unsigned ppm, unsigned char* sig,
                 unsigned char* dig, unsigned char* pub, unsigned* corrupted) {
  u32 s[16], dg[8], p[16], cor;
  if (curve == 0) gen_secp_item<CurveK1>(seed, i, ppm, GTabHost<ModPk1>{K1_GTABLE}, s, dg, p, cor);
  else gen_secp_item<CurveR1>(seed, i, ppm, GTabHost<ModPr1>{R1_GTABLE}, s, dg, p, cor);
  memcpy(sig, s, 64);
  memcpy(dig, dg, 32);
  memcpy(pub, p, 64);
  *corrupted = cor;
}

struct PutHost {
  unsigned char* p;
  void operator()(u64 off, u32 v) const { p[off] = (unsigned char)v; }
};

// msgbuf: 4-byte aligned with slack; generates message bytes then signs, applies the flip
void dc_gen_ed(unsigned long long seed, unsigned long long i, unsigned ppm, unsigned long long off, unsigned len,
               unsigned char* msgbuf, unsigned char* sig, unsigned char* pk, unsigned* corrupted) {
  PutHost put{msgbuf};
  gen_ed_message(seed, i, off, len, put);
  MemHost mem{(const u32*)msgbuf};
  u32 s[16], p[8], cor, mask;
  u64 foff;
  gen_ed_item(seed, i, ppm, off, len, BTabHost{}, mem, s, p, cor, foff, mask);
  if (mask) msgbuf[foff] ^= (unsigned char)mask;
  memcpy(sig, s, 64);
  memcpy(pk, p, 32);
  *corrupted = cor;
}
}
