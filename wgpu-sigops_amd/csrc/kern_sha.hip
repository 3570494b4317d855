// kern_sha.hip -- SHA-256 on the device: Fuel's Message::new over raw data, and addresses of
// recovered public keys (SHA-256 of the 64-byte X || Y).
#include "kernels_impl.hpp"
#include "sha256.hpp"

namespace sigops {

// digest i = SHA-256(msgs[off[i] .. off[i+1])); msgs 4-byte aligned with slack after the last byte.
// Messages differ in length, so no lane owns one message for the whole launch: each loop trip
// compresses one block per lane, and a lane whose message is done takes the next index from the
// queue counter (the lanes of a wave that need one in the same trip take consecutive indices with
// one atomic). A wave stays busy until the queue runs dry instead of idling behind its longest
// message. queue[0] must be 0 at launch; every lane leaves once an index >= n comes out.
static __global__ __launch_bounds__(256) void k_message_new(const u32* __restrict__ msgs,
                                                           const u64* __restrict__ off, u32 n,
                                                           uint4* __restrict__ out, u32* __restrict__ queue) {
  const MemDev mem{msgs};
  u32 i = 0, b = 0, nblk = 0, len = 0;
  u64 o = 0;
  u32 h[8];
  for (;;) {
    const bool need = b == nblk;
    const u64 want = __ballot(need);
    if (want != 0) {
      const u32 leader = (u32)__ffsll((unsigned long long)want) - 1;
      u32 first = 0;
      if (__lane_id() == leader) first = atomicAdd(queue, (u32)__popcll(want));
      first = __builtin_amdgcn_readlane(first, leader);
      if (need) {
        i = first + __builtin_amdgcn_mbcnt_hi((u32)(want >> 32), __builtin_amdgcn_mbcnt_lo((u32)want, 0u));
        if (i >= n) break;
        o = off[i];
        len = (u32)(off[i + 1] - o);
        nblk = sha256_nblocks(len);
        b = 0;
#pragma unroll
        for (int k = 0; k < 8; k++) h[k] = SHA256_IV[k];
      }
    }
    u32 w[16];
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

hipError_t launch_message_new(const void* msgs, const u64* off, u32 n, void* out, u32* queue, u32 grid,
                              hipStream_t s) {
  if (n == 0) return hipSuccess;
  k_fill_u32<<<1, 64, 0, s>>>(queue, 0u, 1u);
  k_message_new<<<grid_for(n, grid), 256, 0, s>>>((const u32*)msgs, off, n, (uint4*)out, queue);
  return hipGetLastError();
}

hipError_t launch_fuel_address(const void* pubs, const void* status, u32 n, void* out, u32 grid, hipStream_t s) {
  if (n == 0) return hipSuccess;
  k_fuel_address<<<grid_for(n, grid), 256, 0, s>>>((const uint4*)pubs, (const u8*)status, n, (uint4*)out);
  return hipGetLastError();
}

}  // namespace sigops
