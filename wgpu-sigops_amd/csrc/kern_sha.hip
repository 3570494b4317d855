// kern_sha.hip -- SHA-256 on the device: Fuel's Message::new over raw data, and addresses of
// recovered public keys (SHA-256 of the 64-byte X || Y).
#include "kernels_impl.hpp"
#include "sha256.hpp"

namespace sigops {

// Longest messages first. Lanes take message positions by grid stride, so the launch ends with
// the last round of positions; in index order a lane that got a long message in that round kept
// the launch running, alone, for its remaining blocks. A counting sort by block count therefore
// orders the positions longest first: bin 0 holds messages of MSG_BINS or more blocks, bin b > 0
// those of MSG_BINS - b blocks. The work words hold the bin counts, the bin cursors and a flag
// that is set when more than one bin is used (with a single block count the index order is kept
// and the scatter does nothing).
constexpr u32 MW_COUNT = 0, MW_CURSOR = MSG_BINS, MW_FLAG = 2 * MSG_BINS;

DEV u32 msg_bin(const u64* off, u32 i) {
  const u64 nb = (off[i + 1] - off[i] + 8) / 64 + 1;  // SHA-256 blocks, padding included
  return nb >= MSG_BINS ? 0u : MSG_BINS - (u32)nb;
}

static __global__ __launch_bounds__(256) void k_msg_hist(const u64* __restrict__ off, u32 n, u32* __restrict__ work) {
  __shared__ u32 h[MSG_BINS];
  for (u32 t = threadIdx.x; t < MSG_BINS; t += blockDim.x) h[t] = 0;
  __syncthreads();
  for (u32 i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    atomicAdd(&h[msg_bin(off, i)], 1u);
  __syncthreads();
  for (u32 t = threadIdx.x; t < MSG_BINS; t += blockDim.x)
    if (h[t]) atomicAdd(&work[MW_COUNT + t], h[t]);
}

// one lane: exclusive scan of the bin counts into the cursors, and the flag
static __global__ __launch_bounds__(64) void k_msg_scan(u32* __restrict__ work) {
  if (threadIdx.x != 0) return;
  u32 pos = 0, used = 0;
  for (u32 b = 0; b < MSG_BINS; b++) {
    const u32 c = work[MW_COUNT + b];
    work[MW_CURSOR + b] = pos;
    pos += c;
    used += c != 0;
  }
  work[MW_FLAG] = used > 1;
}

// ranks within each workgroup's chunk of 1024 messages (4 per thread) by LDS atomics, then one
// global atomic per bin and chunk reserves the chunk's range in that bin (the chunks are large
// because those atomics, on 66 counters shared by every workgroup, are what the scatter waits on)
static __global__ __launch_bounds__(256) void k_msg_scatter(const u64* __restrict__ off, u32 n, u32* __restrict__ work,
                                                           u32* __restrict__ order) {
  constexpr u32 PER = 4;
  __shared__ u32 cnt[MSG_BINS];
  __shared__ u32 base[MSG_BINS];
  if (work[MW_FLAG] == 0) return;
  for (u32 c0 = blockIdx.x * blockDim.x * PER; c0 < n; c0 += gridDim.x * blockDim.x * PER) {
    for (u32 t = threadIdx.x; t < MSG_BINS; t += blockDim.x) cnt[t] = 0;
    __syncthreads();
    u32 bq[PER], rq[PER];
#pragma unroll
    for (u32 q = 0; q < PER; q++) {
      const u32 i = c0 + q * blockDim.x + threadIdx.x;
      bq[q] = MSG_BINS;
      rq[q] = 0;
      if (i < n) {
        bq[q] = msg_bin(off, i);
        rq[q] = atomicAdd(&cnt[bq[q]], 1u);
      }
    }
    __syncthreads();
    for (u32 t = threadIdx.x; t < MSG_BINS; t += blockDim.x)
      if (cnt[t]) base[t] = atomicAdd(&work[MW_CURSOR + t], cnt[t]);
    __syncthreads();
#pragma unroll
    for (u32 q = 0; q < PER; q++)
      if (bq[q] < MSG_BINS) order[base[bq[q]] + rq[q]] = c0 + q * blockDim.x + threadIdx.x;
    __syncthreads();
  }
}

// digest i = SHA-256(msgs[off[i] .. off[i+1])); msgs 4-byte aligned with slack after the last byte.
// Messages differ in length, so no lane owns one message for the whole launch: each loop trip
// compresses one block per lane, and a lane whose message is done takes its next position by grid
// stride and the message at that position of the longest-first order. In that order each lane
// takes one message from every round of similar lengths, so the lanes stay balanced. (A shared
// work counter handing out positions instead was the bottleneck: one contended atomic on every
// message's critical path, 2.4x slower on 2^20 64-byte messages.) X4: msgs is 16-byte aligned.
template <bool X4>
static __global__ __launch_bounds__(256) void k_message_new(const u32* __restrict__ msgs,
                                                           const u64* __restrict__ off, u32 n,
                                                           uint4* __restrict__ out, u32* __restrict__ work,
                                                           const u32* __restrict__ order) {
  const MemDev mem{msgs};
  const bool sorted = work[MW_FLAG] != 0;
  const u32 stride = gridDim.x * blockDim.x;
  u32 next = blockIdx.x * blockDim.x + threadIdx.x;
  u32 i = 0, b = 0, nblk = 0, len = 0;
  u64 o = 0;
  u32 h[8];
  for (;;) {
    const bool need = b == nblk;
    if (__ballot(need) != 0) {
      if (need) {
        const u32 j = next;
        next += stride;
        if (j >= n) break;
        i = sorted ? order[j] : j;
        o = off[i];
        len = (u32)(off[i + 1] - o);
        nblk = sha256_nblocks(len);
        b = 0;
#pragma unroll
        for (int k = 0; k < 8; k++) h[k] = SHA256_IV[k];
      }
    }
    u32 w[16];
    if constexpr (X4)
      sha256_block_words_x4(mem, o, len, b, nblk, w);
    else
      sha256_block_words(mem, o, len, b, nblk, w);
    sha256_compress(h, w);
    if (++b == nblk) {
      out[2 * (u64)i] = make_uint4(bswap32(h[0]), bswap32(h[1]), bswap32(h[2]), bswap32(h[3]));
      out[2 * (u64)i + 1] = make_uint4(bswap32(h[4]), bswap32(h[5]), bswap32(h[6]), bswap32(h[7]));
    }
  }
}

// address i = SHA-256(pub i), or 32 zero bytes where status[i] != 0 (a failed recovery)
static __global__ __launch_bounds__(256) void k_fuel_address(const uint4* __restrict__ pubs,
                                                            const u8* __restrict__ status, u32 n,
                                                            uint4* __restrict__ out) {
  for (u32 i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    u32 pk[16], d[8];
    load_words(pubs + 4 * (u64)i, pk, 4);
    sha256_64(pk, d);
    const u32 keep = (status == nullptr || status[i] == 0) ? ~0u : 0u;
    out[2 * (u64)i] = make_uint4(d[0] & keep, d[1] & keep, d[2] & keep, d[3] & keep);
    out[2 * (u64)i + 1] = make_uint4(d[4] & keep, d[5] & keep, d[6] & keep, d[7] & keep);
  }
}

static u32 grid_for(u32 n, u32 grid) {
  const u32 want = (n + 255) / 256;
  return want < grid ? want : grid;
}

hipError_t launch_message_new(const void* msgs, const u64* off, u32 n, void* out, u32* work, u32* order, u32 grid,
                              hipStream_t s) {
  if (n == 0) return hipSuccess;
  k_fill_u32<<<1, 256, 0, s>>>(work, 0u, MSG_WORK_WORDS);
  k_msg_hist<<<grid_for(n, grid), 256, 0, s>>>(off, n, work);
  k_msg_scan<<<1, 64, 0, s>>>(work);
  k_msg_scatter<<<grid_for(n, grid), 256, 0, s>>>(off, n, work, order);
  if (((uintptr_t)msgs & 15u) == 0)
    k_message_new<true><<<grid_for(n, grid), 256, 0, s>>>((const u32*)msgs, off, n, (uint4*)out, work, order);
  else
    k_message_new<false><<<grid_for(n, grid), 256, 0, s>>>((const u32*)msgs, off, n, (uint4*)out, work, order);
  return hipGetLastError();
}

hipError_t launch_fuel_address(const void* pubs, const void* status, u32 n, void* out, u32 grid, hipStream_t s) {
  if (n == 0) return hipSuccess;
  k_fuel_address<<<grid_for(n, grid), 256, 0, s>>>((const uint4*)pubs, (const u8*)status, n, (uint4*)out);
  return hipGetLastError();
}

}  // namespace sigops
