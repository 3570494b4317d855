// kern_stream.hip -- persistent launches that drain tiles the host publishes while they run: the
// streamed host batch (one queue, tiles packed into pinned memory chunk by chunk) and the serving
// ring (one queue per scheme, positions reused). Layout and host side in stream_host.hpp; the
// per-item code is the batch kernel's (kernels_impl.hpp).
#include "kernels_impl.hpp"

namespace sigops {

static __global__ void k_zero_words(u32* p, u32 n) {
  for (u32 i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0;
}

struct StreamArgs {
  const uint4* desc;
  const uint4* in;
  const u32* mrel;
  const u32* mlen;
  const uint4* hmsgs;
  uint4* dmsgs;
  uint4* out;
  u8* st;
  u64* tile_done;
  const u64* ready;  // nq words
  const u32* qpos;   // nq x (mask + 1) positions, or null
  const u32* stop;
  u32* err;
  u32* ctr;
  u32* pend_list;
  u32 pend_cap, slots, base, mask, seq, nq, split0, split1;
  u64 poll_ticks, limit_ticks;
  u64 msg_cap;  // bytes of dmsgs (and hmsgs) a tile's message region must lie within
};

// a non-home queue's tile is taken only after this many looks found none in the home queue
constexpr u32 STREAM_STEAL_AFTER = 3;

// Claims a tile: a ticket (the tile's index in its queue) of the wavefront's home queue when that
// has a published tile nobody took yet, else (after STREAM_STEAL_AFTER looks) of the queue with the
// most such tiles; then waits until the ticket's tile is published. Returns the queue, with the
// ticket in `ticket`; -1 when the launch is over (every queue closed and taken, the ticket's queue
// closed below it, or the host moved on to a later launch); -2 when it gave up (the host's stop
// flag, or limit_ticks) after marking the launch failed.
// A ticket is an atomic add on the queue's counter, taken only when a free tile was seen, so a
// claim never retries; wavefronts that race for the last free tile hold tickets of tiles still
// to come and wait for them. The published counts are read from device copies: one wavefront at
// a time, at most once per poll_ticks and only when the copies leave it nothing to do, reads the
// pinned words over PCIe and folds them into the copies (an atomic max: a count only grows, and
// the closed flag is its top bit); the others read the copy of their block's XCD. A waiting
// wavefront sleeps longer the farther its tile is, or the longer it found nothing. (When every
// idle wavefront read host memory itself, the PCIe link had no room left for the tiles' inputs.)
// One lane works; the wavefront follows its answer. The acquire is at system scope: a ring
// position is rewritten by the host while the launch runs, so no cache may keep a line of an
// earlier tile at the position.
DEV int stream_claim(const StreamArgs& a, u32 home, u32& ticket) {
  int res = -1;
  u32 tk = 0;
  if ((threadIdx.x & 63u) == 0) {
    u8* base = (u8*)a.ctr;
    u32* claim = a.ctr;  // claim counters of the queues
    u64* last = (u64*)(base + 64);
    u64* stop = (u64*)(base + 72);
    u64* mirror = (u64*)(base + 256 + 256 * (blockIdx.x % STREAM_MIRRORS));
    const u64 t0 = (u64)wall_clock64();
    u32 looks = 0, idle = 0;
    int q = -1;  // the queue whose ticket tk this wavefront holds
    for (;;) {
      u64 r[STREAM_QUEUES];
      int pick = -1;
      bool over = true, superseded = false;
      for (int pass = 0; pass < 2; pass++) {
        pick = -1;
        over = true;
        superseded = false;
        u32 most = 0;
        for (u32 k = 0; k < a.nq; k++) {
          r[k] = __hip_atomic_load(mirror + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          if (r[k] == 0) {
            over = false;
            continue;
          }
          if ((u32)(r[k] >> 32) != a.seq) {
            superseded = true;
            continue;
          }
          const u32 cnt = (u32)r[k] & ~STREAM_CLOSED;
          const u32 c = __hip_atomic_load(claim + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          if (!((u32)r[k] & STREAM_CLOSED) || c < cnt) over = false;
          if (c < cnt && pick != (int)home && (k == home || cnt - c > most)) {
            pick = (int)k;
            most = cnt - c;
          }
        }
        // nothing to do with the copies as they are: refresh them from pinned memory
        const bool stuck = q >= 0 ? (r[q] == 0 || (!superseded && ((u32)r[q] & ~STREAM_CLOSED) <= tk &&
                                                   !((u32)r[q] & STREAM_CLOSED)))
                                  : (pick < 0 && !over && !superseded);
        if (!stuck || pass) break;
        const u64 now = (u64)wall_clock64();
        u64 prev = __hip_atomic_load(last, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (now - prev < a.poll_ticks ||
            !__hip_atomic_compare_exchange_strong(last, &prev, now, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                  __HIP_MEMORY_SCOPE_AGENT))
          break;
        if (__hip_atomic_load(a.stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM))
          __hip_atomic_store(stop, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        for (u32 k = 0; k < a.nq; k++) {
          const u64 h = __hip_atomic_load(a.ready + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
          for (u32 m = 0; m < STREAM_MIRRORS; m++)
            __hip_atomic_fetch_max((u64*)(base + 256 + 256 * m) + k, h, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
      }
      if (superseded) {
        res = -1;
        break;
      }
      u32 naps = 1;
      if (q < 0) {
        if (pick >= 0 && (pick == (int)home || looks >= STREAM_STEAL_AFTER)) {
          q = pick;
          tk = __hip_atomic_fetch_add(claim + q, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          continue;  // the ticket's tile is mostly published already: look at once
        }
        if (over) {
          res = -1;
          break;
        }
        if (pick >= 0) {
          looks++;
        } else {
          naps = 1u << (idle < 5 ? idle : 5);
          idle++;
        }
      } else if (r[q] != 0) {
        const u32 cnt = (u32)r[q] & ~STREAM_CLOSED;
        if (tk < cnt) {
          res = q;
          break;
        }
        if ((u32)r[q] & STREAM_CLOSED) {  // closed below the ticket
          res = -1;
          break;
        }
        const u32 d = tk - cnt;
        naps = d / 64 < 31 ? 1 + d / 64 : 32;
      }
      const u64 now = (u64)wall_clock64();
      if (__hip_atomic_load(stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) || now - t0 >= a.limit_ticks) {
        __hip_atomic_store(a.err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        res = -2;
        break;
      }
      for (u32 k = 0; k < naps; k++) __builtin_amdgcn_s_sleep(127);
    }
  }
  res = __shfl(res, 0);
  ticket = __shfl(tk, 0);
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "");
  return res;
}

// inputs of lane l of slot t from the tile layout: 16 signature words, 8 digest / key words
DEV void stream_load(const StreamArgs& a, u32 t, u32 l, u32 sr[16], u32 ar[8]) {
  const uint4* in = a.in + (u64)t * (6 * 64) + l;
#pragma unroll
  for (int c = 0; c < 4; c++) {
    const uint4 q = in[c * 64];
    sr[4 * c] = q.x; sr[4 * c + 1] = q.y; sr[4 * c + 2] = q.z; sr[4 * c + 3] = q.w;
  }
#pragma unroll
  for (int c = 0; c < 2; c++) {
    const uint4 q = in[(4 + c) * 64];
    ar[4 * c] = q.x; ar[4 * c + 1] = q.y; ar[4 * c + 2] = q.z; ar[4 * c + 3] = q.w;
  }
}

DEV void stream_store_out(const StreamArgs& a, u32 t, u32 l, const u32 o[16]) {
  uint4* out = a.out + (u64)t * (4 * 64) + l;
#pragma unroll
  for (int c = 0; c < 4; c++) out[c * 64] = make_uint4(o[4 * c], o[4 * c + 1], o[4 * c + 2], o[4 * c + 3]);
}

// Persistent: each wavefront takes the next tile, waits until the host has published it, and
// processes its items (one per lane) straight from pinned memory; an Ed25519 tile first copies its
// messages into HBM with 1 KB-wide loads (the hash reads them word by word), and the results go
// back to pinned memory, after which the tile's done word is set. Secp items whose fast ladder
// meets an exceptional addition go to the pending list for k_stream_careful. The launch ends when
// its tiles run out: the batch's last slot, or the count the host closed. WGB: the secp streams'
// fixed-base window, 8 bits (128 odd multiples of G staged in LDS per curve) or 9 (256).
template <int WGB>
__global__ __launch_bounds__(256, SIGOPS_WAVES_PER_EU) void k_stream_run(StreamArgs a, const uint4* __restrict__ gk1,
                                                                    const uint4* __restrict__ gr1,
                                                                    const uint4* __restrict__ gb,
                                                                    uint4* __restrict__ lane_tab, u32 lane_slots) {
  constexpr u32 GN = WGB == 9 ? GTAB9_ENTRIES : GTAB_ENTRIES;
  __shared__ uint4 s_k1[GN * gtab_chunks<CurveK1>()];
  __shared__ uint4 s_r1[GN * gtab_chunks<CurveR1>()];
  __shared__ uint4 s_b[BTAB_ENTRIES * 7];
  stage_gtab<CurveK1, GN>(s_k1, gk1);
  stage_gtab<CurveR1, GN>(s_r1, gr1);
  for (int t = threadIdx.x; t < BTAB_ENTRIES * 7; t += blockDim.x) s_b[t] = gb[t];
  __syncthreads();
  const u32 lane = threadIdx.x & 63u;
  const u32 slot = blockIdx.x * blockDim.x + threadIdx.x;
  STabDev jt;
  jt.base = lane_tab;
  jt.slots = lane_slots;
  jt.slot = slot;
  CTabDev ct;
  ct.base = lane_tab;
  ct.slots = lane_slots;
  ct.slot = slot;
  // the home queue of this block: the blocks that share a CU share one (consecutive blocks go to
  // different CUs, so block b and b + 256 meet), and each queue gets its share of them
  const u32 hb = blockIdx.x % 256u;
  const u32 home = a.nq == 1 ? 0u : (hb < a.split0 ? 0u : (hb < a.split1 ? 1u : 2u));
  for (u32 n_tiles = 0; n_tiles < a.slots; n_tiles++) {
    u32 t = 0;
    const int q = stream_claim(a, home, t);
    if (q < 0) break;
    // the tile's position in the layout: the queue's position table, or base + t
    const u32 p = __builtin_amdgcn_readfirstlane(
        a.qpos ? __hip_atomic_load(a.qpos + (u64)q * (a.mask + 1) + (t & a.mask), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_SYSTEM)
               : (a.base + t) & a.mask);
    const uint4 dv = a.desc[p];
    const uint4 d = make_uint4(__builtin_amdgcn_readfirstlane(dv.x), __builtin_amdgcn_readfirstlane(dv.y),
                               __builtin_amdgcn_readfirstlane(dv.z), __builtin_amdgcn_readfirstlane(dv.w));
    const u32 info = d.x;
    const u32 cnt = info & 0xFFu, b = info >> 8;
    if (cnt == 0) continue;  // an unused position
    if (b >= BUCKET_ED0) {
      const u64 m0 = d.z;
      // a region outside the message buffer (a descriptor that is not this launch's) is not
      // read: the tile's lanes keep their canary statuses
      const bool fits = m0 + d.y + 16 <= a.msg_cap;
      if (fits && a.hmsgs != a.dmsgs) {
        // the tile's messages, [m0, m0 + mbytes) rounded up to 16 bytes, pinned -> HBM (when the
        // host has not copied them to HBM already)
        const u64 q0 = m0 >> 4, nq = ((u64)d.y + 15) >> 4;
        const uint4* src = a.hmsgs + q0;
        uint4* dst = a.dmsgs + q0;
        u64 q = lane;
        for (; q + 192 < nq; q += 256) {  // four PCIe loads in flight per lane
          const uint4 v0 = src[q], v1 = src[q + 64], v2 = src[q + 128], v3 = src[q + 192];
          dst[q] = v0;
          dst[q + 64] = v1;
          dst[q + 128] = v2;
          dst[q + 192] = v3;
        }
        for (; q < nq; q += 64) dst[q] = src[q];
        // the wave's copy is complete and visible before its lanes hash the messages; the release
        // also writes the lines back, so no L2 keeps a dirty copy that could later overwrite the
        // region another tile at these offsets writes
        __threadfence();
      }
      if (fits && lane < cnt) {
        const u32 rel = a.mrel[(u64)p * 64 + lane], len = a.mlen[(u64)p * 64 + lane];
        if ((u64)rel + len <= d.y) {
          u32 sr[16], ar[8];
          stream_load(a, p, lane, sr, ar);
          const u32 st = ed25519_verify_item<false>(sr, ar, MemDev{(const u32*)a.dmsgs}, m0 + rel, len, ct,
                                                    BTabLds{s_b});
          a.st[(u64)p * 64 + lane] = (u8)st;
        }
      }
    } else if (lane < cnt) {
      u32 sr[16], ar[8], o[16];
      stream_load(a, p, lane, sr, ar);
      const u32 st = b == BUCKET_K1 ? secp_item<CurveK1, WGB>(sr, ar, jt, GTabOf<CurveK1>{s_k1}, o)
                                    : secp_item<CurveR1, WGB>(sr, ar, jt, GTabOf<CurveR1>{s_r1}, o);
      if (st == ST_PENDING) {
        const u32 k = atomicAdd(&a.ctr[STREAM_CTR_PENDING], 1u);
        if (k < a.pend_cap) a.pend_list[k] = p * 64 + lane;
      }
      stream_store_out(a, p, lane, o);
      a.st[(u64)p * 64 + lane] = (u8)st;
    }
    // every lane's results reach host memory before the tile is marked done: the stores into the
    // pinned buffers may sit in L2, so the release is at system scope (waiting for the stores
    // alone left results the host could not yet see)
    __threadfence_system();
    if (lane == 0) __hip_atomic_store(&a.tile_done[p], ((u64)a.seq << 32) | d.w, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

// the careful pass of a streamed launch: pending position lanes, inputs and results in the tile layout
__global__ __launch_bounds__(64) void k_stream_careful(StreamArgs a, const uint4* __restrict__ gk1,
                                                       const uint4* __restrict__ gr1) {
  const u32 cnt = min(a.ctr[STREAM_CTR_PENDING], a.pend_cap);
  for (u32 j = blockIdx.x * blockDim.x + threadIdx.x; j < cnt; j += gridDim.x * blockDim.x) {
    const u32 p = a.pend_list[j];
    const u32 t = p >> 6, l = p & 63u;  // position and lane
    const u32 b = a.desc[t].x >> 8;
    u32 sr[16], ar[8], o[16];
    stream_load(a, t, l, sr, ar);
    secp_pro pro;
    jacf Q;
    if (b == BUCKET_K1) {
      secp_prologue<CurveK1>(sr, ar, pro);
      Q = secp_ladder_careful<CurveK1>(pro, GTabGlobal{gk1});
      if (!Q.inf) secp_output<CurveK1>(Q.P, o);
    } else {
      secp_prologue<CurveR1>(sr, ar, pro);
      Q = secp_ladder_careful<CurveR1>(pro, GTabGlobal{gr1});
      if (!Q.inf) secp_output<CurveR1>(Q.P, o);
    }
    if (Q.inf) {
#pragma unroll
      for (int k = 0; k < 16; k++) o[k] = 0;
    }
    stream_store_out(a, t, l, o);
    a.st[p] = (u8)(Q.inf ? ST_INFINITY : ST_OK);
  }
}

hipError_t launch_stream(const StreamLaunch& L, hipStream_t s) {
  StreamArgs a;
  a.desc = (const uint4*)L.desc;
  a.in = (const uint4*)L.in;
  a.mrel = L.mrel;
  a.mlen = L.mlen;
  a.hmsgs = (const uint4*)L.hmsgs;
  a.dmsgs = (uint4*)L.dmsgs;
  a.out = (uint4*)L.out;
  a.st = L.st;
  a.tile_done = L.tile_done;
  a.ready = L.ready;
  a.qpos = L.qpos;
  a.nq = L.nq ? L.nq : 1;
  a.split0 = L.split0;
  a.split1 = L.split1;
  a.stop = L.stop;
  a.err = L.err;
  a.ctr = L.ctr;
  a.pend_list = L.pend_list;
  a.pend_cap = L.pend_cap;
  a.slots = L.slots;
  a.base = L.base;
  a.mask = L.mask;
  a.seq = L.seq;
  a.poll_ticks = L.poll_ticks;
  a.limit_ticks = L.limit_ticks;
  a.msg_cap = L.msg_cap;
  k_zero_words<<<1, 256, 0, s>>>(L.ctr, STREAM_CTR_BYTES / 4);
  if (L.g_window == 9)
    k_stream_run<9><<<L.grid, 256, 0, s>>>(a, (const uint4*)L.gk1, (const uint4*)L.gr1, (const uint4*)L.gb,
                                           (uint4*)L.lane_tab, L.lane_slots);
  else
    k_stream_run<8><<<L.grid, 256, 0, s>>>(a, (const uint4*)L.gk1, (const uint4*)L.gr1, (const uint4*)L.gb,
                                           (uint4*)L.lane_tab, L.lane_slots);
  k_stream_careful<<<64, 64, 0, s>>>(a, (const uint4*)L.gk1, (const uint4*)L.gr1);
  return hipGetLastError();
}

}  // namespace sigops
