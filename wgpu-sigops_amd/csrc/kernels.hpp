// kernels.hpp -- launch helpers shared by the runtime (internal to libsigops).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace sigops {

typedef uint8_t u8;

constexpr uint32_t GTAB_ENTRIES = 1u << (8 - 1);  // odd multiples 1..255 of G
constexpr uint32_t BTAB_ENTRIES = 1u << (8 - 1);  // odd multiples 1..255 of B
constexpr uint32_t SECP_LANE_TAB_BYTES = 8 * 9 * 16;  // 8 entries x 9 uint4 (Jacobian + Z product, then affine)
constexpr uint32_t ED_LANE_TAB_BYTES = 8 * 9 * 16;    // 8 cached entries x 9 uint4

hipError_t launch_secp_recover_CurveK1(const void* sigs, const void* digs, void* out, void* status, uint32_t n,
                                      const void* gtab, void* rtab, uint32_t slots, uint32_t* pend_list,
                                      uint32_t* pend_count, uint32_t grid, hipStream_t s);
hipError_t launch_secp_recover_CurveR1(const void* sigs, const void* digs, void* out, void* status, uint32_t n,
                                      const void* gtab, void* rtab, uint32_t slots, uint32_t* pend_list,
                                      uint32_t* pend_count, uint32_t grid, hipStream_t s);

// verification with a known public key (64 bytes X || Y per item); status only
hipError_t launch_secp_verify_CurveK1(const void* sigs, const void* digs, const void* pubs, void* status, uint32_t n,
                                     const void* gtab, void* rtab, uint32_t slots, uint32_t* pend_list,
                                     uint32_t* pend_count, uint32_t grid, hipStream_t s);
hipError_t launch_secp_verify_CurveR1(const void* sigs, const void* digs, const void* pubs, void* status, uint32_t n,
                                     const void* gtab, void* rtab, uint32_t slots, uint32_t* pend_list,
                                     uint32_t* pend_count, uint32_t grid, hipStream_t s);

// Ed25519 verification given the challenges in kbuf (launch_ed_hash_bucketed writes them);
// strict: ed25519-dalek verify_strict instead of verify
hipError_t launch_ed25519_verify(const void* sigs, const void* pks, void* status, uint32_t n, const void* btab,
                                 void* ctab, uint32_t slots, void* kbuf, uint32_t grid, hipStream_t s,
                                 bool strict = false);

// ---- mixed batch ----
constexpr uint32_t BUCKET_K1 = 0, BUCKET_R1 = 1, BUCKET_ED0 = 2;
constexpr uint32_t ED_BINS = 40;  // Ed25519 buckets by SHA-512 block count 1..40 (40 = "40 or more")
constexpr uint32_t NBUCKETS = 2 + ED_BINS;
constexpr uint32_t BATCH_WORK_WORDS = 3 * NBUCKETS + 3;  // counts, cursors, tile starts, tile ctr, pending ctr
constexpr uint32_t MAX_TILE_ROUNDS = 4;

struct BatchLaunch {
  const void* scheme;
  const void* sigs;
  const void* aux;
  const void* msgs;
  const uint64_t* msg_off;
  void* out;
  void* status;
  const uint32_t* n_ptr;
  uint32_t* work;       // BATCH_WORK_WORDS
  uint32_t* order;      // capacity max_batch + NBUCKETS * 64 * tile_rounds
  uint32_t* pend_list;  // capacity max_batch
  const uint8_t* rank;  // NBUCKETS
  const void* gk1;
  const void* gr1;
  const void* gb;
  void* lane_tab;
  void* kbuf;  // n x 32 bytes: Ed25519 challenges
  uint32_t slots, grid, aux_grid, tile_rounds;
  hipEvent_t* stage_ev;  // null, or 5 events recorded before, between and after the 4 stages
};

hipError_t launch_batch(const BatchLaunch& L, hipStream_t s);
// bucketing and hash passes only, for an all-Ed25519 batch (L.scheme == NULL): challenges into L.kbuf
hipError_t launch_ed_hash_bucketed(const BatchLaunch& L, hipStream_t s);

// SHA-256: digests of messages (Fuel Message::new) and addresses of 64-byte public keys.
// work: MSG_WORK_WORDS device words (cleared by the launch); order: n words, the longest-first
// processing order of the messages
constexpr uint32_t MSG_BINS = 66;  // message ordering bins by SHA-256 block count (66 = "66 or more")
constexpr uint32_t MSG_WORK_WORDS = 1 + 2 * MSG_BINS;
hipError_t launch_message_new(const void* msgs, const uint64_t* off, uint32_t n, void* out, uint32_t* work,
                              uint32_t* order, uint32_t grid, hipStream_t s);
hipError_t launch_fuel_address(const void* pubs, const void* status, uint32_t n, void* out, uint32_t grid,
                               hipStream_t s);

// caller-supplied generator tables in the reference's layout -> this engine's table (flags[0] = 0
// when the table passes its checks); lbits in 11..15
hipError_t launch_table_secp(int curve, const uint32_t* limbs, uint32_t lbits, void* gtab_out, uint32_t* flags,
                             hipStream_t s);
hipError_t launch_table_ed(const uint32_t* limbs, uint32_t lbits, void* btab_out, uint32_t* flags, hipStream_t s);

hipError_t launch_gen_secp(int curve, uint64_t seed, uint32_t n, uint32_t ppm, const void* gtab, void* sigs,
                           void* digs, void* pubs, void* corrupted, uint32_t grid, hipStream_t s);
hipError_t launch_gen_ed25519(uint64_t seed, uint32_t n, uint32_t ppm, const void* btab, const uint64_t* msg_off,
                              void* msgs, void* sigs, void* pks, void* corrupted, uint32_t grid, hipStream_t s);

}  // namespace sigops
