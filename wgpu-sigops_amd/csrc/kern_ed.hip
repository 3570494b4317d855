// kern_ed.hip -- instantiates the Ed25519 verification kernels and provides their launch helper
// (the challenges come from the bucketed hash pass of kern_batch.hip).
#include "kernels_impl.hpp"

namespace sigops {

// verify_strict: the rules on R of every item into the top bits of its k (ed25519_strict_r_checks)
static __global__ __launch_bounds__(256) void k_ed25519_strict(const uint4* __restrict__ sigs, u32 n,
                                                               uint4* __restrict__ kbuf) {
  for (u32 i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    u32 sr[16];
    load_words(sigs + 4 * (u64)i, sr, 4);
    const u32 st = ed25519_strict_r_checks(sr);
    uint4 top = kbuf[2 * (u64)i + 1];
    top.w |= st << ED_K_STRICT_SHIFT;
    kbuf[2 * (u64)i + 1] = top;
  }
}

// the challenges are in kbuf already (launch_ed_hash_bucketed); strict adds verify_strict's pass
hipError_t launch_ed25519_verify(const void* sigs, const void* pks, void* status, u32 n, const void* btab, void* ctab,
                                 u32 slots, void* kbuf, u32 grid, hipStream_t s, bool strict) {
  if (n == 0) return hipSuccess;
  const u32 want = (n + 255) / 256;
  const u32 g = want < grid ? want : grid;
  if (strict) {
    k_ed25519_strict<<<g, 256, 0, s>>>((const uint4*)sigs, n, (uint4*)kbuf);
    k_ed25519_verify<true><<<g, 256, 0, s>>>((const uint4*)sigs, (const uint4*)pks, (const uint4*)kbuf, (u8*)status,
                                            n, (const uint4*)btab, (uint4*)ctab, slots);
  } else {
    k_ed25519_verify<false><<<g, 256, 0, s>>>((const uint4*)sigs, (const uint4*)pks, (const uint4*)kbuf, (u8*)status,
                                             n, (const uint4*)btab, (uint4*)ctab, slots);
  }
  return hipGetLastError();
}

}  // namespace sigops
