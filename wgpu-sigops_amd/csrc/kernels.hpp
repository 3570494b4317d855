// kernels.hpp -- launch helpers shared by the runtime (internal to libsigops).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace sigops {

typedef uint8_t u8;

constexpr uint32_t GTAB_ENTRIES = 1u << (8 - 1);  // odd multiples 1..255 of G
constexpr uint32_t BTAB_ENTRIES = 1u << (8 - 1);  // odd multiples 1..255 of B
constexpr uint32_t SECP_LANE_TAB_BYTES = 8 * 7 * 16;  // 8 Jacobian entries x 7 uint4
constexpr uint32_t ED_LANE_TAB_BYTES = 8 * 9 * 16;    // 8 cached entries x 9 uint4

hipError_t launch_secp_recover_CurveK1(const void* sigs, const void* digs, void* out, void* status, uint32_t n,
                                      const void* gtab, void* rtab, uint32_t slots, uint32_t* pend_list,
                                      uint32_t* pend_count, uint32_t grid, hipStream_t s);
hipError_t launch_secp_recover_CurveR1(const void* sigs, const void* digs, void* out, void* status, uint32_t n,
                                      const void* gtab, void* rtab, uint32_t slots, uint32_t* pend_list,
                                      uint32_t* pend_count, uint32_t grid, hipStream_t s);

hipError_t launch_ed25519_verify(const void* sigs, const void* pks, const void* msgs, const uint64_t* msg_off,
                                 void* status, uint32_t n, const void* btab, void* ctab, uint32_t slots,
                                 uint32_t grid, hipStream_t s);

}  // namespace sigops
