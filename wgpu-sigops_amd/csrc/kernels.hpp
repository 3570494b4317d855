// kernels.hpp -- launch helpers shared by the runtime (internal to libsigops).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace sigops {

typedef uint8_t u8;

constexpr uint32_t GTAB_ENTRIES = 1u << (8 - 1);  // odd multiples 1..255 of G
constexpr uint32_t GTAB9_ENTRIES = 1u << (9 - 1);  // odd multiples 1..511 of G (9-bit windows; the HBM table's size)
constexpr uint32_t BTAB_ENTRIES = 1u << (8 - 1);  // odd multiples 1..255 of B (and as many of [2^136] B)
// the Ed25519 fixed-base tables as held in HBM and LDS: odd multiples of B, then of [2^136] B,
// 7 uint4 per entry
constexpr uint32_t BTAB_UINT4 = 2 * BTAB_ENTRIES * 7;
constexpr uint32_t SECP_LANE_TAB_BYTES = 8 * 9 * 16;  // 8 entries x 9 uint4 (Jacobian + Z product, then affine)
constexpr uint32_t ED_LANE_TAB_BYTES = 16 * 9 * 16;   // 16 cached entries (-A, then -R) x 9 uint4

// g_window: the fixed-base window, 8 or 9 bits (9 needs a table of 256 odd multiples of G: the
// resident one)
hipError_t launch_secp_recover_CurveK1(const void* sigs, const void* digs, void* out, void* status, uint32_t n,
                                      const void* gtab, void* rtab, uint32_t slots, uint32_t* pend_list,
                                      uint32_t* pend_count, uint32_t grid, uint32_t g_window, hipStream_t s);
hipError_t launch_secp_recover_CurveR1(const void* sigs, const void* digs, void* out, void* status, uint32_t n,
                                      const void* gtab, void* rtab, uint32_t slots, uint32_t* pend_list,
                                      uint32_t* pend_count, uint32_t grid, uint32_t g_window, hipStream_t s);

// verification with a known public key (64 bytes X || Y per item); status only
hipError_t launch_secp_verify_CurveK1(const void* sigs, const void* digs, const void* pubs, void* status, uint32_t n,
                                     const void* gtab, void* rtab, uint32_t slots, uint32_t* pend_list,
                                     uint32_t* pend_count, uint32_t grid, uint32_t g_window, hipStream_t s);
hipError_t launch_secp_verify_CurveR1(const void* sigs, const void* digs, const void* pubs, void* status, uint32_t n,
                                     const void* gtab, void* rtab, uint32_t slots, uint32_t* pend_list,
                                     uint32_t* pend_count, uint32_t grid, uint32_t g_window, hipStream_t s);

// ---- mixed batch ----
constexpr uint32_t BUCKET_K1 = 0, BUCKET_R1 = 1, BUCKET_ED0 = 2;
constexpr uint32_t ED_BINS = 40;  // Ed25519 buckets by SHA-512 block count 1..40 (40 = "40 or more")
constexpr uint32_t NBUCKETS = 2 + ED_BINS;
constexpr uint32_t BATCH_WORK_WORDS = 3 * NBUCKETS + 3;  // counts, cursors, tile starts, tile ctr, pending ctr
constexpr uint32_t MAX_TILE_ROUNDS = 4;

// The buffers of one batch call and its size. The host writes them into device memory
// (launch_set_batch_io, stream-ordered) before the launches, and the kernels read them from
// there, so one captured graph per context serves every call whatever buffers it passes.
struct BatchIO {
  const void* scheme;  // NULL: every item is an Ed25519 verification
  const void* sigs;
  const void* aux;
  const void* msgs;
  const uint64_t* msg_off;
  void* out;  // NULL for the plain Ed25519 path
  void* status;
  uint32_t n;
};
hipError_t launch_set_batch_io(BatchIO* d_io, const BatchIO& v, hipStream_t s);

struct BatchLaunch {
  const BatchIO* io;    // device copy of the call's buffers
  uint32_t* work;       // BATCH_WORK_WORDS
  uint32_t* order;      // capacity max_batch + NBUCKETS * 64 * tile_rounds
  uint32_t* pend_list;  // capacity max_batch
  const uint8_t* rank;  // NBUCKETS
  const void* gk1;
  const void* gr1;
  const void* gb;
  void* lane_tab;
  uint32_t slots, grid, aux_grid, tile_rounds;
  uint32_t waves;        // mixed batch: waves per SIMD the persistent kernel is built for (1 or 2)
  uint32_t g_window;     // mixed batch: fixed-base window of the secp streams, 8 or 9 bits (9: waves 2 only)
  hipEvent_t* stage_ev;  // null, or 5 events recorded before, between and after the 4 stages
};

hipError_t launch_batch(const BatchLaunch& L, hipStream_t s);
// plain Ed25519 verification of an all-Ed25519 batch (io->scheme == NULL, io->out == NULL): the
// bucketing, then the persistent kernel on Ed25519 tiles with the challenges hashed inside it;
// strict: ed25519-dalek verify_strict rules; L.gb is the fixed-base table
hipError_t launch_ed_verify(const BatchLaunch& L, bool strict, hipStream_t s);

// ---- streamed tiles: the host packs tiles into pinned, device-mapped memory while one persistent
// launch drains them (layout in stream_host.hpp). A streamed batch fills positions 0.. of the
// layout once; a serving ring reuses positions base + t (mod positions) for the launch's tile t ----
constexpr uint32_t STREAM_CLOSED = 1u << 31;  // in the published count: no tile beyond it will come
// A launch drains up to STREAM_QUEUES queues of tiles (a serving ring: one per scheme; a batch: one).
// Device words of a streamed launch, cleared by the launch: u32 [q] the claim counter of queue q,
// [STREAM_CTR_PENDING] the pending count; u64 at byte 64 the wall-clock time of the last read of
// the pinned counts, at byte 72 the stop flag seen there; STREAM_MIRRORS copies of the pinned
// counts (u64 per queue) from byte 256 on, 256 bytes apart
constexpr uint32_t STREAM_QUEUES = 3;
constexpr uint32_t STREAM_CTR_PENDING = 4;
constexpr uint32_t STREAM_MIRRORS = 8;
constexpr uint32_t STREAM_CTR_BYTES = 256 + STREAM_MIRRORS * 256;
struct StreamLaunch {
  const void* desc;      // 16-byte tile descriptors, one per position (pinned, or an HBM copy)
  const void* in;        // inputs, [position][6][64] 16-byte chunks (pinned, or an HBM copy)
  const uint32_t* mrel;  // per position lane, message start relative to the tile's
  const uint32_t* mlen;  // per position lane, message length
  const void* hmsgs;     // the tiles' messages: pinned, copied by each tile to dmsgs; or equal to dmsgs
                         // when the host has copied them to HBM before publishing the tile
  void* dmsgs;           // device: the messages the hash reads (same offsets)
  uint64_t msg_cap;      // bytes of hmsgs and dmsgs: a tile's region [m0, m0 + bytes + 16) lies within
  void* out;             // pinned: results, [position][4][64] 16-byte chunks
  uint8_t* st;           // pinned: statuses, [position][64]
  uint64_t* tile_done;   // pinned: per position, seq << 32 | the descriptor's tag once the tile's results are written
  const uint64_t* ready; // pinned, per queue: seq << 32 | tiles published so far, | STREAM_CLOSED at the end
  const uint32_t* qpos;  // pinned: per queue, mask + 1 entries: position of the queue's tile t (t & mask);
                         // null: one queue whose tile t sits at position (base + t) & mask
  uint32_t nq;           // queues (1 .. STREAM_QUEUES)
  uint32_t split0, split1;  // home queue of block b: 0 if b % 256 < split0, 1 if < split1, else 2
  const uint32_t* stop;  // pinned: nonzero makes waiting waves give up
  uint32_t* err;         // pinned: set to 1 by a wave that gave up waiting
  uint32_t* ctr;         // device: STREAM_CTR_BYTES of counters and polling state (cleared by the launch)
  uint32_t* pend_list;   // device: position lanes (position * 64 + lane) the careful pass redoes
  uint32_t pend_cap;     // entries pend_list holds (further pending lanes are not recorded)
  const void* gk1;
  const void* gr1;
  const void* gb;
  void* lane_tab;
  uint32_t slots;        // tiles a wavefront may take at most
  uint32_t base, mask;   // see qpos (a batch: 0 and ~0)
  uint32_t seq, lane_slots, grid;
  uint32_t g_window;     // fixed-base window of the secp items, 8 or 9 bits
  uint64_t poll_ticks;   // wall-clock ticks between reads of the pinned count (device-wide)
  uint64_t limit_ticks;  // a wave that waits this long for a tile gives up
};
hipError_t launch_stream(const StreamLaunch& L, hipStream_t s);
// the polling interval (2 us) and the give-up time (15 s) of a streamed launch in the device's
// wall-clock ticks
inline void stream_poll_ticks(int device, StreamLaunch& L) {
  int khz = 0;
  if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device) != hipSuccess || khz <= 0) khz = 100000;
  L.poll_ticks = (uint64_t)khz * 2 / 1000;
  L.limit_ticks = (uint64_t)khz * 15000;
}

// what a serving ring borrows from a context: the resident tables, the lane tables and the launch
// shape they are sized for, the processing order of the buckets, and the context's stream
struct RingResources {
  const void* gk1;
  const void* gr1;
  const void* gb;
  void* lane_tab;
  uint32_t lane_slots, grid;
  uint32_t g_window;             // fixed-base window of the secp items, 8 or 9 bits
  const uint32_t* bucket_order;  // processing rank -> bucket, NBUCKETS entries
  hipStream_t stream;
};


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

// the resources of a context that a serving ring launches with (sigops_api.hip); SIGOPS_OK or an
// error code
struct sigops_ctx;
int sigops_ring_resources(sigops_ctx* c, sigops::RingResources* r);

