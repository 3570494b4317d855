// kernels_impl.hpp -- gfx950 kernels of the signature engine (instantiated by kern_*.hip).
//
// One lane per signature, persistent grid-stride kernels (the grid is sized to the resident
// capacity of the GPU, so a lane's slot index also names its private table in HBM). Per kernel:
//   * the fixed-base table (odd multiples of G or B, uploaded to HBM once per context) is staged
//     into LDS by each workgroup and read with ds_read_b128;
//   * the per-signature variable-base table lives in HBM scratch, laid out [entry][chunk][slot]
//     so that lanes that pick the same entry read contiguous 16-byte chunks;
//   * secp lanes whose fast ladder meets an exceptional addition are appended to a pending list
//     and finished by k_secp_careful, enqueued right after on the same stream (no host round trip).
// Replaces the reference's 19 WGSL entry points (src/wgsl/main/*.wgsl), whose multi-stage variants
// write every intermediate point to memory between dispatches (src/secp256k1_ecdsa.rs:78-193).
#pragma once
#include <hip/hip_runtime.h>

#include "ed25519.hpp"
#include "kernels.hpp"
#include "secp.hpp"

// Waves per SIMD the heavy kernels are compiled for (the register budget per lane is 512 / this).
#ifndef SIGOPS_WAVES_PER_EU
#define SIGOPS_WAVES_PER_EU 2
#endif

namespace sigops {

// ---- table accessors ----
// secp fixed-base table in LDS. CH = 5: entry = 5 x uint4 (x[9], y[9], pad[2]), the format of the
// table in HBM. CH = 7 (secp256k1, staged by stage_gtab): (x[9], y[9], beta x[9], pad), so that
// load_phi reads (beta x, y) = phi(entry) from chunks 2..6, as many loads as load's chunks 0..4.
template <int CH>
struct GTabLdsN {
  const uint4* t;
  template <int A>
  DEV void chunks(u32 idx, u32* w) const {
#pragma unroll
    for (int c = A; c < A + 5; c++) {
      const uint4 q = t[idx * CH + c];
      w[4 * c] = q.x; w[4 * c + 1] = q.y; w[4 * c + 2] = q.z; w[4 * c + 3] = q.w;
    }
  }
  DEV void load(u32 idx, fe& x, fe& y) const {
    u32 w[20];
    chunks<0>(idx, w);
#pragma unroll
    for (int k = 0; k < NL; k++) { x.v[k] = w[k]; y.v[k] = w[9 + k]; }
  }
  DEV void load_phi(u32 idx, fe& bx, fe& y) const {
    static_assert(CH == 7, "phi entries need the 7-chunk layout");
    u32 w[28];
    chunks<2>(idx, w);
#pragma unroll
    for (int k = 0; k < NL; k++) { y.v[k] = w[9 + k]; bx.v[k] = w[18 + k]; }
  }
};
typedef GTabLdsN<5> GTabLds;

// the LDS table type of curve C: secp256k1 (GLV) carries beta x
template <class C>
using GTabOf = GTabLdsN<C::GLV ? 7 : 5>;
template <class C>
constexpr int gtab_chunks() { return C::GLV ? 7 : 5; }

// copies the first N entries of the HBM table g (5 chunks per entry) into LDS s in curve C's
// layout, computing beta x for secp256k1 (one product per entry, N per workgroup); the caller
// syncs afterwards
template <class C, u32 N = GTAB_ENTRIES>
DEV void stage_gtab(uint4* s, const uint4* g) {
  if constexpr (C::GLV) {
    typedef typename C::Fp F;
    const fe beta = fe_from_const(K1_BETA);
    for (u32 e = threadIdx.x; e < N; e += blockDim.x) {
      u32 w[28];
#pragma unroll
      for (int c = 0; c < 5; c++) {
        const uint4 q = g[e * 5 + c];
        w[4 * c] = q.x; w[4 * c + 1] = q.y; w[4 * c + 2] = q.z; w[4 * c + 3] = q.w;
      }
      fe x;
#pragma unroll
      for (int k = 0; k < NL; k++) x.v[k] = w[k];
      const fe bx = mont_mul<F>(x, beta);
#pragma unroll
      for (int k = 0; k < NL; k++) w[18 + k] = bx.v[k];
      w[27] = 0;
#pragma unroll
      for (int c = 0; c < 7; c++) s[e * 7 + c] = make_uint4(w[4 * c], w[4 * c + 1], w[4 * c + 2], w[4 * c + 3]);
    }
  } else {
    for (u32 t = threadIdx.x; t < N * 5; t += blockDim.x) s[t] = g[t];
  }
}

struct BTabLds {  // Ed25519 fixed-base: entry = 7 x uint4 (ypx[9], ymx[9], t2d[9], pad)
  const uint4* t;
  DEV void load(u32 idx, fe& ypx, fe& ymx, fe& t2d) const {
    u32 w[28];
#pragma unroll
    for (int c = 0; c < 7; c++) {
      const uint4 q = t[idx * 7 + c];
      w[4 * c] = q.x; w[4 * c + 1] = q.y; w[4 * c + 2] = q.z; w[4 * c + 3] = q.w;
    }
#pragma unroll
    for (int k = 0; k < NL; k++) { ypx.v[k] = w[k]; ymx.v[k] = w[9 + k]; t2d.v[k] = w[18 + k]; }
  }
};

// per-lane table in HBM: chunk c of entry e of slot s at base[(e * CH + c) * slots + s]
template <int CH>
struct LaneTab {
  uint4* base;
  u32 slots, slot;
  DEV void put(int e, const u32* w) {
#pragma unroll
    for (int c = 0; c < CH; c++)
      base[((u64)e * CH + c) * slots + slot] = make_uint4(w[4 * c], w[4 * c + 1], w[4 * c + 2], w[4 * c + 3]);
  }
  DEV void get(u32 e, u32* w) const { get_n<CH>(e, w); }
  // the first N chunks of entry e
  template <int N>
  DEV void put_n(int e, const u32* w) {
#pragma unroll
    for (int c = 0; c < N; c++)
      base[((u64)e * CH + c) * slots + slot] = make_uint4(w[4 * c], w[4 * c + 1], w[4 * c + 2], w[4 * c + 3]);
  }
  template <int N>
  DEV void get_n(u32 e, u32* w) const { get_range<0, N>(e, w); }
  // chunks A .. A+N-1 of entry e into w[4A ..]
  template <int A, int N>
  DEV void get_range(u32 e, u32* w) const {
#pragma unroll
    for (int c = A; c < A + N; c++) {
      const uint4 q = base[((u64)e * CH + c) * slots + slot];
      w[4 * c] = q.x; w[4 * c + 1] = q.y; w[4 * c + 2] = q.z; w[4 * c + 3] = q.w;
    }
  }
};

// secp R-table (secp_build_table): while building, entry e holds a Jacobian point and the running
// product of the Z coordinates before it (36 words, 9 chunks); afterwards its first 5 chunks hold
// the affine (x, y) the ladder reads, and for secp256k1 chunks 4..6 also hold beta * x, so that
// the phi(R) stream reads (beta x, y) from chunks 2..6 (five loads, as for (x, y))
struct STabDev : LaneTab<9> {
  DEV void store_full(int e, const jac& p, const fe& before) {
    u32 w[36];
#pragma unroll
    for (int k = 0; k < NL; k++) {
      w[k] = p.X.v[k]; w[9 + k] = p.Y.v[k]; w[18 + k] = p.Z.v[k]; w[27 + k] = before.v[k];
    }
    put(e, w);
  }
  DEV void load_full(u32 e, jac& p, fe& before) const {
    u32 w[36];
    get(e, w);
#pragma unroll
    for (int k = 0; k < NL; k++) {
      p.X.v[k] = w[k]; p.Y.v[k] = w[9 + k]; p.Z.v[k] = w[18 + k]; before.v[k] = w[27 + k];
    }
  }
  DEV void store_affine(int e, const fe& x, const fe& y) {
    u32 w[20];
#pragma unroll
    for (int k = 0; k < NL; k++) { w[k] = x.v[k]; w[9 + k] = y.v[k]; }
    w[18] = w[19] = 0;
    put_n<5>(e, w);
  }
  DEV void load_affine(u32 e, fe& x, fe& y) const {
    u32 w[20];
    get_n<5>(e, w);
#pragma unroll
    for (int k = 0; k < NL; k++) { x.v[k] = w[k]; y.v[k] = w[9 + k]; }
  }
  DEV void store_affine_phi(int e, const fe& x, const fe& y, const fe& bx) {
    u32 w[28];
#pragma unroll
    for (int k = 0; k < NL; k++) { w[k] = x.v[k]; w[9 + k] = y.v[k]; w[18 + k] = bx.v[k]; }
    w[27] = 0;
    put_n<7>(e, w);
  }
  DEV void load_affine_phi(u32 e, fe& bx, fe& y) const {
    u32 w[28];
    get_range<2, 5>(e, w);
#pragma unroll
    for (int k = 0; k < NL; k++) { y.v[k] = w[9 + k]; bx.v[k] = w[18 + k]; }
  }
};

struct CTabDev : LaneTab<9> {  // cached Edwards point: 36 words
  DEV void store(int e, const cached& p) {
    u32 w[36];
#pragma unroll
    for (int k = 0; k < NL; k++) {
      w[k] = p.ypx.v[k]; w[9 + k] = p.ymx.v[k]; w[18 + k] = p.z.v[k]; w[27 + k] = p.t2d.v[k];
    }
    put(e, w);
  }
  DEV cached load(u32 e) const {
    u32 w[36];
    get(e, w);
    cached p;
#pragma unroll
    for (int k = 0; k < NL; k++) {
      p.ypx.v[k] = w[k]; p.ymx.v[k] = w[9 + k]; p.z.v[k] = w[18 + k]; p.t2d.v[k] = w[27 + k];
    }
    return p;
  }
};

struct GTabGlobal {
  const uint4* t;
  DEV void load(u32 idx, fe& x, fe& y) const { GTabLds{t}.load(idx, x, y); }
};

struct MemDev {
  const u32* w;
  DEV u32 word(u64 i) const { return w[i]; }
  // 16-byte chunk c (the buffer must be 16-byte aligned)
  DEV void quad(u64 c, u32* r) const {
    const uint4 v = ((const uint4*)w)[c];
    r[0] = v.x;
    r[1] = v.y;
    r[2] = v.z;
    r[3] = v.w;
  }
};

DEV void load_words(const uint4* p, u32* w, int nq) {
  for (int c = 0; c < nq; c++) {
    const uint4 q = p[c];
    w[4 * c] = q.x; w[4 * c + 1] = q.y; w[4 * c + 2] = q.z; w[4 * c + 3] = q.w;
  }
}

// ---- kernels ----
// WGB: the fixed-base window, 8 bits (128 odd multiples of G staged in LDS) or 9 (256; the resident
// table only, a caller's converted table holds 128)
template <class C, int WGB = 8>
__global__ __launch_bounds__(256, SIGOPS_WAVES_PER_EU) void k_secp_recover(const uint4* __restrict__ sigs,
                                                      const uint4* __restrict__ digs, uint4* __restrict__ out,
                                                      u8* __restrict__ status, u32 n,
                                                      const uint4* __restrict__ gtab, uint4* __restrict__ rtab,
                                                      u32 slots, u32* __restrict__ pend_list,
                                                      u32* __restrict__ pend_count) {
  constexpr u32 GN = WGB == 9 ? GTAB9_ENTRIES : GTAB_ENTRIES;
  __shared__ uint4 sg[GN * gtab_chunks<C>()];
  stage_gtab<C, GN>(sg, gtab);
  __syncthreads();
  const GTabOf<C> gt{sg};
  const u32 slot = blockIdx.x * blockDim.x + threadIdx.x;
  STabDev tab;
  tab.base = rtab;
  tab.slots = slots;
  tab.slot = slot;
  const u32 stride = gridDim.x * blockDim.x;
  for (u32 i = slot; i < n; i += stride) {
    u32 sr[16], dr[8], o[16];
    load_words(sigs + 4 * (u64)i, sr, 4);
    load_words(digs + 2 * (u64)i, dr, 2);
    secp_pro pro;
    secp_prologue<C>(sr, dr, pro);
    u32 st = pro.status;
    if (st == ST_OK) {
      bool exc = false;
      const jac Q = secp_ladder<C, WGB>(pro, tab, gt, exc);
      if (exc) {
        st = ST_PENDING;
        pend_list[atomicAdd(pend_count, 1u)] = i;
      } else {
        secp_output<C>(Q, o);
      }
    }
    if (st != ST_OK) {
#pragma unroll
      for (int k = 0; k < 16; k++) o[k] = 0;
    }
#pragma unroll
    for (int c = 0; c < 4; c++) out[4 * (u64)i + c] = make_uint4(o[4 * c], o[4 * c + 1], o[4 * c + 2], o[4 * c + 3]);
    status[i] = (u8)st;
  }
}

template <class C>
__global__ __launch_bounds__(64) void k_secp_careful(const uint4* __restrict__ sigs, const uint4* __restrict__ digs,
                                                     uint4* __restrict__ out, u8* __restrict__ status,
                                                     const uint4* __restrict__ gtab,
                                                     const u32* __restrict__ pend_list,
                                                     const u32* __restrict__ pend_count) {
  const u32 cnt = *pend_count;
  const GTabGlobal gt{gtab};
  for (u32 j = blockIdx.x * blockDim.x + threadIdx.x; j < cnt; j += gridDim.x * blockDim.x) {
    const u32 i = pend_list[j];
    u32 sr[16], dr[8], o[16];
    load_words(sigs + 4 * (u64)i, sr, 4);
    load_words(digs + 2 * (u64)i, dr, 2);
    secp_pro pro;
    secp_prologue<C>(sr, dr, pro);
    const jacf Q = secp_ladder_careful<C>(pro, gt);
    u32 st = ST_OK;
    if (Q.inf) {
      st = ST_INFINITY;
#pragma unroll
      for (int k = 0; k < 16; k++) o[k] = 0;
    } else {
      secp_output<C>(Q.P, o);
    }
#pragma unroll
    for (int c = 0; c < 4; c++) out[4 * (u64)i + c] = make_uint4(o[4 * c], o[4 * c + 1], o[4 * c + 2], o[4 * c + 3]);
    status[i] = (u8)st;
  }
}

// ECDSA verification with a known key (fuel Signature::verify): the recovery kernel with the key
// in place of R and the projective x comparison in place of the affine output. Lanes that meet an
// exceptional addition are finished by k_secp_verify_careful. WGB as in k_secp_recover.
template <class C, int WGB = 8>
__global__ __launch_bounds__(256, SIGOPS_WAVES_PER_EU) void k_secp_verify(
    const uint4* __restrict__ sigs, const uint4* __restrict__ digs, const uint4* __restrict__ pubs,
    u8* __restrict__ status, u32 n, const uint4* __restrict__ gtab, uint4* __restrict__ rtab, u32 slots,
    u32* __restrict__ pend_list, u32* __restrict__ pend_count) {
  constexpr u32 GN = WGB == 9 ? GTAB9_ENTRIES : GTAB_ENTRIES;
  __shared__ uint4 sg[GN * gtab_chunks<C>()];
  stage_gtab<C, GN>(sg, gtab);
  __syncthreads();
  const GTabOf<C> gt{sg};
  const u32 slot = blockIdx.x * blockDim.x + threadIdx.x;
  STabDev tab;
  tab.base = rtab;
  tab.slots = slots;
  tab.slot = slot;
  const u32 stride = gridDim.x * blockDim.x;
  for (u32 i = slot; i < n; i += stride) {
    u32 sr[16], dr[8], pr[16];
    load_words(sigs + 4 * (u64)i, sr, 4);
    load_words(digs + 2 * (u64)i, dr, 2);
    load_words(pubs + 4 * (u64)i, pr, 4);
    secp_pro pro;
    secp_prologue_verify<C>(sr, dr, pr, pro);
    u32 st = pro.status;
    if (st == ST_OK) {
      bool exc = false;
      const jac P = secp_ladder<C, WGB>(pro, tab, gt, exc);
      if (exc) {
        st = ST_PENDING;
        pend_list[atomicAdd(pend_count, 1u)] = i;
      } else {
        load_words(sigs + 4 * (u64)i, sr, 4);
        st = secp_verify_x<C>(P, sr);
      }
    }
    status[i] = (u8)st;
  }
}

template <class C>
__global__ __launch_bounds__(64) void k_secp_verify_careful(const uint4* __restrict__ sigs,
                                                            const uint4* __restrict__ digs,
                                                            const uint4* __restrict__ pubs, u8* __restrict__ status,
                                                            const uint4* __restrict__ gtab,
                                                            const u32* __restrict__ pend_list,
                                                            const u32* __restrict__ pend_count) {
  const u32 cnt = *pend_count;
  const GTabGlobal gt{gtab};
  for (u32 j = blockIdx.x * blockDim.x + threadIdx.x; j < cnt; j += gridDim.x * blockDim.x) {
    const u32 i = pend_list[j];
    u32 sr[16], dr[8], pr[16];
    load_words(sigs + 4 * (u64)i, sr, 4);
    load_words(digs + 2 * (u64)i, dr, 2);
    load_words(pubs + 4 * (u64)i, pr, 4);
    secp_pro pro;
    secp_prologue_verify<C>(sr, dr, pr, pro);
    const jacf P = secp_ladder_careful<C>(pro, gt);
    status[i] = (u8)(P.inf ? ST_INFINITY : secp_verify_x<C>(P.P, sr));
  }
}

static __global__ void k_fill_u32(u32* p, u32 v, u32 n) {
  for (u32 i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}

// one recovery on the fast ladder: the status (ST_PENDING when the ladder met an exceptional
// addition and the careful pass must redo the item) and the key X || Y (zeros unless ST_OK)
// (WGB: the fixed-base window, 8 or 9 bits, which the table gt must match)
template <class C, int WGB = 8, class GT>
DEV u32 secp_item(const u32 sr[16], const u32 dr[8], STabDev& tab, const GT& gt, u32 o[16]) {
  secp_pro pro;
  secp_prologue<C>(sr, dr, pro);
  u32 st = pro.status;
  if (st == ST_OK) {
    bool exc = false;
    const jac Q = secp_ladder<C, WGB>(pro, tab, gt, exc);
    if (exc)
      st = ST_PENDING;
    else
      secp_output<C>(Q, o);
  }
  if (st != ST_OK) {
#pragma unroll
    for (int k = 0; k < 16; k++) o[k] = 0;
  }
  return st;
}

}  // namespace sigops
