// kern_gen.hip -- seeded workload generation kernels (device signing; see gen.hpp) and launchers.
#include "gen.hpp"
#include "kernels_impl.hpp"

namespace sigops {

template <class C>
__global__ __launch_bounds__(256) void k_gen_secp(u64 seed, u32 n, u32 ppm, const uint4* __restrict__ gtab,
                                                  uint4* __restrict__ sigs, uint4* __restrict__ digs,
                                                  uint4* __restrict__ pubs, u8* __restrict__ corrupted) {
  __shared__ uint4 sg[GTAB_ENTRIES * 5];
  for (int t = threadIdx.x; t < GTAB_ENTRIES * 5; t += blockDim.x) sg[t] = gtab[t];
  __syncthreads();
  const GTabLds gt{sg};
  for (u32 i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    u32 sig[16], dig[8], pub[16], cor;
    gen_secp_item<C>(seed, i, ppm, gt, sig, dig, pub, cor);
#pragma unroll
    for (int c = 0; c < 4; c++) {
      sigs[4 * (u64)i + c] = make_uint4(sig[4 * c], sig[4 * c + 1], sig[4 * c + 2], sig[4 * c + 3]);
      pubs[4 * (u64)i + c] = make_uint4(pub[4 * c], pub[4 * c + 1], pub[4 * c + 2], pub[4 * c + 3]);
    }
#pragma unroll
    for (int c = 0; c < 2; c++) digs[2 * (u64)i + c] = make_uint4(dig[4 * c], dig[4 * c + 1], dig[4 * c + 2], dig[4 * c + 3]);
    if (corrupted) corrupted[i] = (u8)cor;
  }
}

struct PutByteDev {
  u8* p;
  DEV void operator()(u64 off, u32 v) const { p[off] = (u8)v; }
};

__global__ __launch_bounds__(256) void k_gen_ed_msgs(u64 seed, u32 n, const u64* __restrict__ msg_off, u8* msgs) {
  PutByteDev put{msgs};
  for (u32 i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const u64 off = msg_off[i];
    gen_ed_message(seed, i, off, (u32)(msg_off[i + 1] - off), put);
  }
}

__global__ __launch_bounds__(256) void k_gen_ed_sign(u64 seed, u32 n, u32 ppm, const uint4* __restrict__ btab,
                                                     const u64* __restrict__ msg_off, u32* msgs,
                                                     uint4* __restrict__ sigs, uint4* __restrict__ pks,
                                                     u8* __restrict__ corrupted) {
  __shared__ uint4 sb[BTAB_ENTRIES * 7];
  for (int t = threadIdx.x; t < BTAB_ENTRIES * 7; t += blockDim.x) sb[t] = btab[t];
  __syncthreads();
  const BTabLds bt{sb};
  const MemDev mem{msgs};
  for (u32 i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const u64 off = msg_off[i];
    const u32 len = (u32)(msg_off[i + 1] - off);
    u32 sig[16], pk[8], cor, mask;
    u64 foff;
    gen_ed_item(seed, i, ppm, off, len, bt, mem, sig, pk, cor, foff, mask);
#pragma unroll
    for (int c = 0; c < 4; c++) sigs[4 * (u64)i + c] = make_uint4(sig[4 * c], sig[4 * c + 1], sig[4 * c + 2], sig[4 * c + 3]);
#pragma unroll
    for (int c = 0; c < 2; c++) pks[2 * (u64)i + c] = make_uint4(pk[4 * c], pk[4 * c + 1], pk[4 * c + 2], pk[4 * c + 3]);
    if (corrupted) corrupted[i] = (u8)cor;
    // the flip is applied after this lane hashed its message; neighbouring lanes only read their
    // own bytes of a shared word, so the atomic does not disturb them
    if (mask) atomicXor(msgs + (foff >> 2), mask << (8 * (u32)(foff & 3u)));
  }
}

hipError_t launch_gen_secp(int curve, u64 seed, u32 n, u32 ppm, const void* gtab, void* sigs, void* digs, void* pubs,
                           void* corrupted, u32 grid, hipStream_t s) {
  if (n == 0) return hipSuccess;
  const u32 want = (n + 255) / 256;
  const u32 g = want < grid ? want : grid;
  if (curve == 0)
    k_gen_secp<CurveK1><<<g, 256, 0, s>>>(seed, n, ppm, (const uint4*)gtab, (uint4*)sigs, (uint4*)digs, (uint4*)pubs,
                                         (u8*)corrupted);
  else
    k_gen_secp<CurveR1><<<g, 256, 0, s>>>(seed, n, ppm, (const uint4*)gtab, (uint4*)sigs, (uint4*)digs, (uint4*)pubs,
                                         (u8*)corrupted);
  return hipGetLastError();
}

hipError_t launch_gen_ed25519(u64 seed, u32 n, u32 ppm, const void* btab, const u64* msg_off, void* msgs, void* sigs,
                              void* pks, void* corrupted, u32 grid, hipStream_t s) {
  if (n == 0) return hipSuccess;
  const u32 want = (n + 255) / 256;
  const u32 g = want < grid ? want : grid;
  k_gen_ed_msgs<<<g, 256, 0, s>>>(seed, n, msg_off, (u8*)msgs);
  k_gen_ed_sign<<<g, 256, 0, s>>>(seed, n, ppm, (const uint4*)btab, msg_off, (u32*)msgs, (uint4*)sigs, (uint4*)pks,
                                  (u8*)corrupted);
  return hipGetLastError();
}

}  // namespace sigops
