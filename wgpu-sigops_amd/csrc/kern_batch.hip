// kern_batch.hip -- mixed-scheme batch: device-side bucketing plus one persistent kernel.
//
// A batch mixes secp256k1 recoveries, secp256r1 recoveries and Ed25519 verifications with messages
// of any length. The reference has one kernel chain per scheme and fixed 32-byte messages; here:
//   1. k_bucket_hist / k_bucket_scan / k_bucket_scatter sort item indices into buckets (k1, r1, and
//      Ed25519 by SHA-512 block count, so the lanes of a wavefront hash equally long messages), each
//      bucket padded to whole tiles; the buckets are processed in the context's order (its rank
//      table: secp256r1, secp256k1, then Ed25519 by decreasing block count);
//   2. k_batch_run is persistent: every wavefront pulls the next tile (64 x tile_rounds items of
//      one bucket) from an atomic counter, so the scheme is uniform across the wavefront and load
//      balances dynamically; it runs the same per-item code as the plain kernels;
//   3. k_batch_careful finishes secp items whose fast ladder met an exceptional addition.
// The batch's buffers and n are read from device memory (BatchIO), so the captured graph of these
// launches replays for any buffers and any n up to the context capacity.
#include "kernels_impl.hpp"

namespace sigops {

constexpr u32 DUMMY = 0xFFFFFFFFu;

struct BatchArgs {
  const u8* scheme;
  const uint4* sigs;
  const uint4* aux;  // digest (secp) or public key (Ed25519), 32 bytes per item
  const u32* msgs;
  const u64* msg_off;
  uint4* out;
  u8* status;
  u32 n;
  const BatchIO* io;  // the call's buffers and n in device memory; bind_io copies them into the fields above
  u32* counts;      // [NB]
  u32* cursor;      // [NB]
  u32* tile_start;  // [NB + 1] in processing order
  u32* order;       // padded item lists, bucket-major in processing order
  u32* tile_counter;
  u32* pend_list;
  u32* pend_count;
  const u8* rank;   // bucket -> processing rank
  u32 tile_items;   // 64 * tile_rounds
};

// Every kernel starts with bind_io: the per-call buffers come from device memory (written on the
// stream before the launches), the rest of BatchArgs is fixed per context and passed by value, so a
// captured graph of these launches is the same for every call.
DEV void bind_io(BatchArgs& a) {
  const BatchIO& v = *a.io;
  a.scheme = (const u8*)v.scheme;
  a.sigs = (const uint4*)v.sigs;
  a.aux = (const uint4*)v.aux;
  a.msgs = (const u32*)v.msgs;
  a.msg_off = v.msg_off;
  a.out = (uint4*)v.out;
  a.status = (u8*)v.status;
  a.n = v.n;
}

// scheme == NULL: every item is an Ed25519 verification (the plain Ed25519 path hashes through
// the same bucketing, so that its wavefronts hash equally long messages)
DEV u32 item_scheme(const BatchArgs& a, u32 i) { return a.scheme ? a.scheme[i] : 2u; }

DEV u32 item_bucket(const BatchArgs& a, u32 i) {
  const u32 s = item_scheme(a, i);
  if (s == 0) return BUCKET_K1;
  if (s == 1) return BUCKET_R1;
  const u64 len = a.msg_off[i + 1] - a.msg_off[i];
  u64 nb = (64 + len + 17 + 127) >> 7;
  if (nb > ED_BINS) nb = ED_BINS;
  return BUCKET_ED0 + (u32)nb - 1;
}

__global__ __launch_bounds__(256) void k_bucket_hist(BatchArgs a) {
  bind_io(a);
  __shared__ u32 h[NBUCKETS];
  for (u32 t = threadIdx.x; t < NBUCKETS; t += blockDim.x) h[t] = 0;
  __syncthreads();
  const u32 n = a.n;
  for (u32 i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    if (item_scheme(a, i) > 2) continue;
    atomicAdd(&h[item_bucket(a, i)], 1u);
  }
  __syncthreads();
  for (u32 t = threadIdx.x; t < NBUCKETS; t += blockDim.x)
    if (h[t]) atomicAdd(&a.counts[t], h[t]);
}

// one block: tile offsets in processing order, cursors, and DUMMY padding of each bucket's tail
__global__ __launch_bounds__(64) void k_bucket_scan(BatchArgs a) {
  bind_io(a);
  __shared__ u32 start[NBUCKETS];
  if (threadIdx.x == 0) {
    u32 pos_tiles = 0;
    for (u32 r = 0; r < NBUCKETS; r++) {
      u32 b = 0;
      for (u32 q = 0; q < NBUCKETS; q++)
        if (a.rank[q] == r) b = q;
      const u32 c = a.counts[b];
      const u32 tiles = (c + a.tile_items - 1) / a.tile_items;
      a.tile_start[r] = pos_tiles;
      start[b] = pos_tiles * a.tile_items;
      pos_tiles += tiles;
    }
    a.tile_start[NBUCKETS] = pos_tiles;
  }
  __syncthreads();
  for (u32 b = 0; b < NBUCKETS; b++) {
    const u32 c = a.counts[b];
    const u32 padded = (c + a.tile_items - 1) / a.tile_items * a.tile_items;
    for (u32 q = c + threadIdx.x; q < padded; q += blockDim.x) a.order[start[b] + q] = DUMMY;
    if (threadIdx.x == 0) a.cursor[b] = start[b];
  }
}

// Each workgroup ranks its chunk of 1024 items (4 per thread) per bucket with LDS atomics,
// reserves one range per bucket with a single global atomic, then writes the indices: one global
// atomic per bucket and chunk instead of one per item, on a handful of counters shared by every
// workgroup.
__global__ __launch_bounds__(256) void k_bucket_scatter(BatchArgs a) {
  bind_io(a);
  constexpr u32 PER = 4;
  __shared__ u32 cnt[NBUCKETS];
  __shared__ u32 base[NBUCKETS];
  const u32 n = a.n;
  for (u32 c0 = blockIdx.x * blockDim.x * PER; c0 < n; c0 += gridDim.x * blockDim.x * PER) {
    for (u32 t = threadIdx.x; t < NBUCKETS; t += blockDim.x) cnt[t] = 0;
    __syncthreads();
    u32 bq[PER], rq[PER];
#pragma unroll
    for (u32 q = 0; q < PER; q++) {
      const u32 i = c0 + q * blockDim.x + threadIdx.x;
      bq[q] = NBUCKETS;
      rq[q] = 0;
      if (i < n) {
        if (item_scheme(a, i) > 2) {
          a.status[i] = 0xFD;  // unknown scheme: reported as a canary failure by the host
        } else {
          bq[q] = item_bucket(a, i);
          rq[q] = atomicAdd(&cnt[bq[q]], 1u);
        }
      }
    }
    __syncthreads();
    for (u32 t = threadIdx.x; t < NBUCKETS; t += blockDim.x)
      if (cnt[t]) base[t] = atomicAdd(&a.cursor[t], cnt[t]);
    __syncthreads();
#pragma unroll
    for (u32 q = 0; q < PER; q++)
      if (bq[q] < NBUCKETS) a.order[base[bq[q]] + rq[q]] = c0 + q * blockDim.x + threadIdx.x;
    __syncthreads();
  }
}

DEV void store16(uint4* p, const u32 o[16]) {
#pragma unroll
  for (int c = 0; c < 4; c++) p[c] = make_uint4(o[4 * c], o[4 * c + 1], o[4 * c + 2], o[4 * c + 3]);
}

template <class C, int WGB, class GT>
DEV void batch_secp_item(const BatchArgs& a, u32 i, STabDev& tab, const GT& gt) {
  u32 sr[16], dr[8], o[16];
  load_words(a.sigs + 4 * (u64)i, sr, 4);
  load_words(a.aux + 2 * (u64)i, dr, 2);
  const u32 st = secp_item<C, WGB>(sr, dr, tab, gt, o);
  if (st == ST_PENDING) a.pend_list[atomicAdd(a.pend_count, 1u)] = i;
  store16(a.out + 4 * (u64)i, o);
  a.status[i] = (u8)st;
}

// The schemes a persistent launch serves (template mask): the mixed batch serves all three; the
// plain Ed25519 entry points serve only Ed25519 (dalek `verify`, or `verify_strict` rules), so
// they run the same tile loop, with the challenge hashed inside it, and stage only B's table.
constexpr u32 RUN_K1 = 1, RUN_R1 = 2, RUN_ED = 4, RUN_ED_STRICT = 8;
constexpr u32 RUN_MIXED = RUN_K1 | RUN_R1 | RUN_ED;

// WAVES: waves per SIMD the kernel is built for, which sets its register budget per lane (512 /
// WAVES): 2 by default (256 registers, some values spilled to scratch memory); 1 lets the
// compiler keep them in registers (accumulation registers included) with half the waves to hide
// latencies. WGB: the secp streams' fixed-base window, 8 bits (128 odd multiples of G staged in
// LDS) or 9 (256: 63.5 KB of LDS per workgroup, still two per CU). The autotuner measures them.
template <u32 MASK, int WAVES = SIGOPS_WAVES_PER_EU, int WGB = 8>
__global__ __launch_bounds__(256, WAVES) void k_batch_run(BatchArgs a, const uint4* __restrict__ gk1,
                                                   const uint4* __restrict__ gr1, const uint4* __restrict__ gb,
                                                   uint4* __restrict__ lane_tab, u32 slots) {
  static_assert(!((MASK & RUN_ED) && (MASK & RUN_ED_STRICT)), "one Ed25519 rule set per launch");
  constexpr bool K1 = MASK & RUN_K1, R1 = MASK & RUN_R1, ED = MASK & (RUN_ED | RUN_ED_STRICT);
  constexpr u32 GN = WGB == 9 ? GTAB9_ENTRIES : GTAB_ENTRIES;
  // the kernels that serve only Ed25519 run the lattice-reduced check (both fixed-base tables)
  constexpr bool LAT = ED && !K1 && !R1;
  constexpr u32 BN = LAT ? BTAB_UINT4 : BTAB_ENTRIES * 7;
  bind_io(a);
  __shared__ uint4 s_k1[K1 ? GN * gtab_chunks<CurveK1>() : 1];
  __shared__ uint4 s_r1[R1 ? GN * gtab_chunks<CurveR1>() : 1];
  __shared__ uint4 s_b[ED ? BN : 1];
  if constexpr (K1) stage_gtab<CurveK1, GN>(s_k1, gk1);
  if constexpr (R1) stage_gtab<CurveR1, GN>(s_r1, gr1);
  if constexpr (ED)
    for (int t = threadIdx.x; t < BN; t += blockDim.x) s_b[t] = gb[t];
  __syncthreads();
  const u32 slot = blockIdx.x * blockDim.x + threadIdx.x;
  const u32 lane = threadIdx.x & 63u;
  STabDev jt;
  jt.base = lane_tab;
  jt.slots = slots;
  jt.slot = slot;
  CTabDev ct;
  ct.base = lane_tab;
  ct.slots = slots;
  ct.slot = slot;
  const u32 ntiles = a.tile_start[NBUCKETS];
  for (;;) {
    u32 t = 0;
    if (lane == 0) t = atomicAdd(a.tile_counter, 1u);
    t = __shfl(t, 0);
    if (t >= ntiles) break;
    u32 r = 0;
    while (r + 1 < NBUCKETS && a.tile_start[r + 1] <= t) r++;
    u32 b = 0;
    for (u32 q = 0; q < NBUCKETS; q++)
      if (a.rank[q] == r) b = q;
    b = __builtin_amdgcn_readfirstlane(b);
    for (u32 k = 0; k < a.tile_items; k += 64) {
      const u32 i = a.order[(u64)t * a.tile_items + k + lane];
      if (i == DUMMY) continue;
      if (K1 && b == BUCKET_K1) {
        if constexpr (K1) batch_secp_item<CurveK1, WGB>(a, i, jt, GTabOf<CurveK1>{s_k1});
      } else if (R1 && b == BUCKET_R1) {
        if constexpr (R1) batch_secp_item<CurveR1, WGB>(a, i, jt, GTabOf<CurveR1>{s_r1});
      } else if constexpr (ED) {
        // the challenge is hashed here, inside the persistent kernel: the tile's items have equal
        // SHA-512 block counts (bucketed), and the hash's memory latency hides behind the other
        // waves' field arithmetic instead of idling the GPU in a pass of its own
        u32 sr[16], pr[8];
        load_words(a.sigs + 4 * (u64)i, sr, 4);
        load_words(a.aux + 2 * (u64)i, pr, 2);
        const u64 off = a.msg_off[i];
        a.status[i] = (u8)ed25519_verify_item<(MASK & RUN_ED_STRICT) != 0, LAT>(
            sr, pr, MemDev{a.msgs}, off, (u32)(a.msg_off[i + 1] - off), ct, BTabLds{s_b});
        if (a.out) {
          const u32 zero[16] = {0};
          store16(a.out + 4 * (u64)i, zero);
        }
      }
    }
  }
}

__global__ __launch_bounds__(64) void k_batch_careful(BatchArgs a, const uint4* __restrict__ gk1,
                                                      const uint4* __restrict__ gr1) {
  bind_io(a);
  const u32 cnt = *a.pend_count;
  for (u32 j = blockIdx.x * blockDim.x + threadIdx.x; j < cnt; j += gridDim.x * blockDim.x) {
    const u32 i = a.pend_list[j];
    u32 sr[16], dr[8], o[16];
    load_words(a.sigs + 4 * (u64)i, sr, 4);
    load_words(a.aux + 2 * (u64)i, dr, 2);
    secp_pro pro;
    jacf Q;
    if (a.scheme[i] == 0) {
      secp_prologue<CurveK1>(sr, dr, pro);
      Q = secp_ladder_careful<CurveK1>(pro, GTabGlobal{gk1});
      if (!Q.inf) secp_output<CurveK1>(Q.P, o);
    } else {
      secp_prologue<CurveR1>(sr, dr, pro);
      Q = secp_ladder_careful<CurveR1>(pro, GTabGlobal{gr1});
      if (!Q.inf) secp_output<CurveR1>(Q.P, o);
    }
    if (Q.inf) {
#pragma unroll
      for (int k = 0; k < 16; k++) o[k] = 0;
    }
    store16(a.out + 4 * (u64)i, o);
    a.status[i] = (u8)(Q.inf ? ST_INFINITY : ST_OK);
  }
}

static __global__ void k_set_io(BatchIO* d, BatchIO v) { *d = v; }

hipError_t launch_set_batch_io(BatchIO* d_io, const BatchIO& v, hipStream_t s) {
  k_set_io<<<1, 1, 0, s>>>(d_io, v);
  return hipGetLastError();
}

static __global__ void k_zero_u32(u32* p, u32 n) {
  for (u32 i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0;
}

static BatchArgs batch_args(const BatchLaunch& L) {
  BatchArgs a{};
  a.io = L.io;
  a.counts = L.work;                    // NBUCKETS
  a.cursor = L.work + NBUCKETS;         // NBUCKETS
  a.tile_start = L.work + 2 * NBUCKETS;  // NBUCKETS + 1
  a.tile_counter = L.work + 3 * NBUCKETS + 1;
  a.pend_count = L.work + 3 * NBUCKETS + 2;
  a.pend_list = L.pend_list;
  a.order = L.order;
  a.rank = L.rank;
  a.tile_items = 64u * L.tile_rounds;
  return a;
}

// plain Ed25519 verification (io->scheme == NULL, every item Ed25519): the bucketing by SHA-512
// block count, then the persistent kernel serving only Ed25519 tiles, each of which hashes its
// items' challenges and verifies them (dalek `verify`, or `verify_strict` rules)
hipError_t launch_ed_verify(const BatchLaunch& L, bool strict, hipStream_t s) {
  const BatchArgs a = batch_args(L);
  k_zero_u32<<<1, 256, 0, s>>>(L.work, BATCH_WORK_WORDS);
  k_bucket_hist<<<L.aux_grid, 256, 0, s>>>(a);
  k_bucket_scan<<<1, 64, 0, s>>>(a);
  k_bucket_scatter<<<L.aux_grid, 256, 0, s>>>(a);
  if (strict)
    k_batch_run<RUN_ED_STRICT><<<L.grid, 256, 0, s>>>(a, nullptr, nullptr, (const uint4*)L.gb, (uint4*)L.lane_tab,
                                                      L.slots);
  else
    k_batch_run<RUN_ED><<<L.grid, 256, 0, s>>>(a, nullptr, nullptr, (const uint4*)L.gb, (uint4*)L.lane_tab, L.slots);
  return hipGetLastError();
}

hipError_t launch_batch(const BatchLaunch& L, hipStream_t s) {
  const BatchArgs a = batch_args(L);
  // stage events exist only on the eager (timed) path, never under graph capture
  auto mark = [&](int i) -> hipError_t { return L.stage_ev ? hipEventRecord(L.stage_ev[i], s) : hipSuccess; };
  hipError_t e;
  if ((e = mark(0)) != hipSuccess) return e;
  // counters are cleared with a kernel (not a memset node) so the sequence is safe to capture
  k_zero_u32<<<1, 256, 0, s>>>(L.work, BATCH_WORK_WORDS);
  k_bucket_hist<<<L.aux_grid, 256, 0, s>>>(a);
  k_bucket_scan<<<1, 64, 0, s>>>(a);
  k_bucket_scatter<<<L.aux_grid, 256, 0, s>>>(a);
  if ((e = mark(1)) != hipSuccess) return e;
  // (the Ed25519 challenges are hashed inside k_batch_run; the stage stays empty, so that the
  // stage indices of earlier measurements keep their meaning)
  if ((e = mark(2)) != hipSuccess) return e;
  if (L.waves == 1)
    k_batch_run<RUN_MIXED, 1><<<L.grid, 256, 0, s>>>(a, (const uint4*)L.gk1, (const uint4*)L.gr1, (const uint4*)L.gb,
                                                    (uint4*)L.lane_tab, L.slots);
  else if (L.g_window == 9)
    k_batch_run<RUN_MIXED, 2, 9><<<L.grid, 256, 0, s>>>(a, (const uint4*)L.gk1, (const uint4*)L.gr1,
                                                       (const uint4*)L.gb, (uint4*)L.lane_tab, L.slots);
  else
    k_batch_run<RUN_MIXED><<<L.grid, 256, 0, s>>>(a, (const uint4*)L.gk1, (const uint4*)L.gr1, (const uint4*)L.gb,
                                                 (uint4*)L.lane_tab, L.slots);
  if ((e = mark(3)) != hipSuccess) return e;
  k_batch_careful<<<64, 64, 0, s>>>(a, (const uint4*)L.gk1, (const uint4*)L.gr1);
  if ((e = mark(4)) != hipSuccess) return e;
  return hipGetLastError();
}

}  // namespace sigops
