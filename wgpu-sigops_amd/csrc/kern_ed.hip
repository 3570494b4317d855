// kern_ed.hip -- instantiates the Ed25519 verification kernel and provides its launch helper.
#include "kernels_impl.hpp"

namespace sigops {

hipError_t launch_ed25519_verify(const void* sigs, const void* pks, const void* msgs, const u64* msg_off,
                                 void* status, u32 n, const void* btab, void* ctab, u32 slots, void* kbuf, u32 grid,
                                 hipStream_t s, bool strict) {
  if (n == 0) return hipSuccess;
  const u32 want = (n + 255) / 256;
  const u32 g = want < grid ? want : grid;
  if (strict)
    k_ed25519_hash<true><<<g, 256, 0, s>>>((const uint4*)sigs, (const uint4*)pks, (const u32*)msgs, msg_off, nullptr,
                                          nullptr, n, (uint4*)kbuf);
  else
    k_ed25519_hash<false><<<g, 256, 0, s>>>((const uint4*)sigs, (const uint4*)pks, (const u32*)msgs, msg_off, nullptr,
                                           nullptr, n, (uint4*)kbuf);
  if (strict)
    k_ed25519_verify<true><<<g, 256, 0, s>>>((const uint4*)sigs, (const uint4*)pks, (const uint4*)kbuf, (u8*)status,
                                            n, (const uint4*)btab, (uint4*)ctab, slots);
  else
    k_ed25519_verify<false><<<g, 256, 0, s>>>((const uint4*)sigs, (const uint4*)pks, (const uint4*)kbuf, (u8*)status,
                                             n, (const uint4*)btab, (uint4*)ctab, slots);
  return hipGetLastError();
}

}  // namespace sigops
