// kern_secp.hip -- instantiates the secp recovery kernels for one curve (SIGOPS_CURVE = CurveK1
// or CurveR1, set by the build) and provides its launch helper.
#include "kernels_impl.hpp"

#ifndef SIGOPS_CURVE
#error "build with -DSIGOPS_CURVE=CurveK1 or -DSIGOPS_CURVE=CurveR1"
#endif
#define SIGOPS_CAT2(a, b) a##b
#define SIGOPS_CAT(a, b) SIGOPS_CAT2(a, b)

namespace sigops {

hipError_t SIGOPS_CAT(launch_secp_recover_, SIGOPS_CURVE)(const void* sigs, const void* digs, void* out, void* status,
                                                         u32 n, const void* gtab, void* rtab, u32 slots,
                                                         u32* pend_list, u32* pend_count, u32 grid, u32 g_window,
                                                         hipStream_t s) {
  if (n == 0) return hipSuccess;
  const u32 want = (n + 255) / 256;
  const u32 g = want < grid ? want : grid;
  k_fill_u32<<<1, 64, 0, s>>>(pend_count, 0u, 1u);
  if (g_window == 9)
    k_secp_recover<SIGOPS_CURVE, 9><<<g, 256, 0, s>>>((const uint4*)sigs, (const uint4*)digs, (uint4*)out,
                                                     (u8*)status, n, (const uint4*)gtab, (uint4*)rtab, slots,
                                                     pend_list, pend_count);
  else
    k_secp_recover<SIGOPS_CURVE, 8><<<g, 256, 0, s>>>((const uint4*)sigs, (const uint4*)digs, (uint4*)out,
                                                     (u8*)status, n, (const uint4*)gtab, (uint4*)rtab, slots,
                                                     pend_list, pend_count);
  k_secp_careful<SIGOPS_CURVE><<<64, 64, 0, s>>>((const uint4*)sigs, (const uint4*)digs, (uint4*)out, (u8*)status,
                                                (const uint4*)gtab, pend_list, pend_count);
  return hipGetLastError();
}

hipError_t SIGOPS_CAT(launch_secp_verify_, SIGOPS_CURVE)(const void* sigs, const void* digs, const void* pubs,
                                                        void* status, u32 n, const void* gtab, void* rtab, u32 slots,
                                                        u32* pend_list, u32* pend_count, u32 grid, u32 g_window,
                                                        hipStream_t s) {
  if (n == 0) return hipSuccess;
  const u32 want = (n + 255) / 256;
  const u32 g = want < grid ? want : grid;
  k_fill_u32<<<1, 64, 0, s>>>(pend_count, 0u, 1u);
  if (g_window == 9)
    k_secp_verify<SIGOPS_CURVE, 9><<<g, 256, 0, s>>>((const uint4*)sigs, (const uint4*)digs, (const uint4*)pubs,
                                                    (u8*)status, n, (const uint4*)gtab, (uint4*)rtab, slots,
                                                    pend_list, pend_count);
  else
    k_secp_verify<SIGOPS_CURVE, 8><<<g, 256, 0, s>>>((const uint4*)sigs, (const uint4*)digs, (const uint4*)pubs,
                                                    (u8*)status, n, (const uint4*)gtab, (uint4*)rtab, slots,
                                                    pend_list, pend_count);
  k_secp_verify_careful<SIGOPS_CURVE><<<64, 64, 0, s>>>((const uint4*)sigs, (const uint4*)digs, (const uint4*)pubs,
                                                       (u8*)status, (const uint4*)gtab, pend_list, pend_count);
  return hipGetLastError();
}

}  // namespace sigops
