// kern_sha.hip -- SHA-256 on the device: Fuel's Message::new over raw data, and addresses of
// recovered public keys (SHA-256 of the 64-byte X || Y).
#include "kernels_impl.hpp"
#include "sha256.hpp"

namespace sigops {

// digest i = SHA-256(msgs[off[i] .. off[i+1])); msgs 4-byte aligned with slack after the last byte
static __global__ __launch_bounds__(256) void k_message_new(const u32* __restrict__ msgs,
                                                           const u64* __restrict__ off, u32 n,
                                                           uint4* __restrict__ out) {
  const MemDev mem{msgs};
  for (u32 i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const u64 o = off[i];
    u32 d[8];
    sha256_msg(mem, o, (u32)(off[i + 1] - o), d);
    out[2 * (u64)i] = make_uint4(d[0], d[1], d[2], d[3]);
    out[2 * (u64)i + 1] = make_uint4(d[4], d[5], d[6], d[7]);
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

hipError_t launch_message_new(const void* msgs, const u64* off, u32 n, void* out, u32 grid, hipStream_t s) {
  if (n == 0) return hipSuccess;
  k_message_new<<<grid_for(n, grid), 256, 0, s>>>((const u32*)msgs, off, n, (uint4*)out);
  return hipGetLastError();
}

hipError_t launch_fuel_address(const void* pubs, const void* status, u32 n, void* out, u32 grid, hipStream_t s) {
  if (n == 0) return hipSuccess;
  k_fuel_address<<<grid_for(n, grid), 256, 0, s>>>((const uint4*)pubs, (const u8*)status, n, (uint4*)out);
  return hipGetLastError();
}

}  // namespace sigops
