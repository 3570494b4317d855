// kern_ed.hip -- instantiates the Ed25519 verification kernels and provides their launch helper
// (the challenges come from the bucketed hash pass of kern_batch.hip).
#include "kernels_impl.hpp"

namespace sigops {

// the challenges are in kbuf already (launch_ed_hash_bucketed); strict adds verify_strict's pass
hipError_t launch_ed25519_verify(const void* sigs, const void* pks, void* status, u32 n, const void* btab, void* ctab,
                                 u32 slots, void* kbuf, u32 grid, hipStream_t s, bool strict) {
  if (n == 0) return hipSuccess;
  const u32 want = (n + 255) / 256;
  const u32 g = want < grid ? want : grid;
  if (strict) {
    k_ed25519_strict<<<g, 256, 0, s>>>((const uint4*)sigs, (const uint4*)pks, n, (uint4*)kbuf);
    k_ed25519_verify<true><<<g, 256, 0, s>>>((const uint4*)sigs, (const uint4*)pks, (const uint4*)kbuf, (u8*)status,
                                            n, (const uint4*)btab, (uint4*)ctab, slots);
  } else {
    k_ed25519_verify<false><<<g, 256, 0, s>>>((const uint4*)sigs, (const uint4*)pks, (const uint4*)kbuf, (u8*)status,
                                             n, (const uint4*)btab, (uint4*)ctab, slots);
  }
  return hipGetLastError();
}

}  // namespace sigops
