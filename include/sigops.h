/*
 * sigops.h -- C ABI of the MI355X (gfx950) batch signature engine (libsigops.so).
 *
 * This boundary replaces the reference's wgpu plumbing (src/gpu.rs:5-186) and WGSL assembly
 * (src/shader.rs:20-647) behind its public entry points:
 *   secp256k1_ecdsa::ecrecover / ecrecover_single_shader   src/secp256k1_ecdsa.rs:61-66, :215-219
 *   secp256r1_ecdsa::ecrecover / ecrecover_single_shader   src/secp256r1_ecdsa.rs:62-67, :216-220
 *   ed25519_eddsa::ecverify / ecverify_single               src/ed25519_eddsa.rs:67-73, :259-264
 * Plain pointers and sizes only. The caller owns every host buffer; the library owns device memory.
 * One context may be used by one host thread at a time. No function throws or unwinds.
 *
 * Return values: 0 on success, a negative SIGOPS_E_* code on failure (the reference maps its only
 * failure, a missing `success` canary, to ShaderFailureError: src/lib.rs:12-14). Per-item results
 * never fail the call; invalid signatures are reported per item.
 */
#ifndef SIGOPS_H
#define SIGOPS_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* call status */
#define SIGOPS_OK 0
#define SIGOPS_E_ARG (-1)     /* bad argument (NULL pointer with n > 0, n too large, bad scheme) */
#define SIGOPS_E_DEVICE (-2)  /* HIP runtime error; sigops_last_error() has the text */
#define SIGOPS_E_NOMEM (-3)   /* device or pinned-host allocation failed */
#define SIGOPS_E_CANARY (-4)  /* a kernel did not write every result (ShaderFailureError) */

/* per-item status for secp recovery (same codes as the oracle). On any non-zero status the 64
   output bytes are zero, which is what a drop-in wrapper returns (the reference has no per-item
   error channel). */
#define SIGOPS_ST_OK 0
#define SIGOPS_ST_BAD_R 1     /* r == 0 or r >= n */
#define SIGOPS_ST_BAD_S 2     /* s == 0 or s >= n after clearing the recovery bit */
#define SIGOPS_ST_NO_POINT 3  /* r is not the x-coordinate of a curve point */
#define SIGOPS_ST_INFINITY 4  /* the recovered key is the point at infinity (verification: u1 G + u2 Q is) */
/* verification with a known public key only */
#define SIGOPS_ST_HIGH_S 5    /* secp256k1: s > (n-1)/2 (libsecp256k1 verifies lower-S signatures only) */
#define SIGOPS_ST_BAD_KEY 6   /* the public key is not a curve point with X, Y < p */
#define SIGOPS_ST_MISMATCH 7  /* x(u1 G + u2 Q) mod n != r */

/* per-item status for Ed25519 verification (valid[i] = status == 0) */
#define SIGOPS_ED_OK 0
#define SIGOPS_ED_BAD_S 1     /* s >= l (non-canonical scalar) */
#define SIGOPS_ED_BAD_KEY 2   /* public key does not decompress */
#define SIGOPS_ED_MISMATCH 3  /* [s]B - [k]A does not compress to R */
/* verify_strict only */
#define SIGOPS_ED_BAD_R 4       /* R does not decompress */
#define SIGOPS_ED_SMALL_ORDER 5 /* R or A has small order ([8]P is the identity) */

/* schemes of the mixed batch */
#define SIGOPS_SCHEME_SECP256K1 0
#define SIGOPS_SCHEME_SECP256R1 1
#define SIGOPS_SCHEME_ED25519 2

typedef struct sigops_ctx sigops_ctx;

typedef struct {
  int device;               /* HIP device ordinal */
  uint32_t blocks_per_cu;   /* resident 256-thread blocks per CU for the persistent kernels (0: default 2) */
  uint32_t max_batch;       /* largest n accepted per call (0: 1 << 22, the reference's limit,
                               src/benchmarks/mod.rs:45) */
  uint32_t tile_rounds;     /* mixed batch: a wavefront takes 64 * tile_rounds items of one bucket per
                               queue pop (1..4, 0: default 1) */
  uint32_t no_graph;        /* mixed batch: 1 = plain stream launches instead of the captured graph */
  uint32_t waves_per_simd;  /* mixed batch: the persistent kernel built for 2 waves per SIMD (256 registers
                               per lane; 0: default 2) or 1 (512: no spills to scratch memory, half the
                               waves); blocks_per_cu defaults to this */
  uint32_t bucket_order;    /* mixed batch: processing order of the buckets, 0 (default) secp256r1,
                               secp256k1, Ed25519 by decreasing SHA-512 block count; 1 secp256r1, Ed25519,
                               secp256k1; 2 secp256k1, secp256r1, Ed25519 */
  uint32_t g_window;        /* every secp kernel: fixed-base window of the secp256k1 / secp256r1 G streams,
                               8 or 9 bits (256 odd multiples of G in LDS, 4 fewer additions per
                               secp256k1 item and 3 per secp256r1 item); 9 needs waves_per_simd 2; calls
                               with a caller's generator table use 8. 0: 9 with two waves per SIMD
                               (measured faster on the MI355X), 8 with one */
} sigops_config;

/* Creates a context on one GPU: uploads the fixed-base generator tables to HBM once (the
   reference re-uploads them on every call, src/secp256k1_ecdsa.rs:109) and allocates per-lane
   scratch. cfg may be NULL. Returns NULL on failure. */
sigops_ctx* sigops_create(const sigops_config* cfg);
void sigops_destroy(sigops_ctx* ctx);
/* text of the last error on this context ("" if none); valid until the next call */
const char* sigops_last_error(const sigops_ctx* ctx);

/* ---- plain per-scheme entry points: host buffers in and out, transfers included ---- */

/* secp256k1 public-key recovery, fuel-crypto format.
   sigs: n x 64 bytes r||s, big-endian, recovery parity in bit 7 of byte 32;
   digests: n x 32 bytes (message digests, taken mod n);
   out_pub: n x 64 bytes X||Y big-endian (zeros for a failed item); status: n bytes or NULL.
   Replaces secp256k1_ecdsa::ecrecover / ecrecover_single_shader (src/secp256k1_ecdsa.rs:61-274). */
int sigops_secp256k1_recover(sigops_ctx* ctx, const uint8_t* sigs, const uint8_t* digests, size_t n,
                             uint8_t* out_pub, uint8_t* status);

/* secp256r1 (P-256) recovery with the same layout.
   Replaces secp256r1_ecdsa::ecrecover / ecrecover_single_shader (src/secp256r1_ecdsa.rs:62-275). */
int sigops_secp256r1_recover(sigops_ctx* ctx, const uint8_t* sigs, const uint8_t* digests, size_t n,
                             uint8_t* out_pub, uint8_t* status);

/* Ed25519 verification with ed25519-dalek `verify` rules.
   sigs: n x 64 bytes R||s; pubkeys: n x 32 bytes; message i is msgs[msg_off[i] .. msg_off[i+1]);
   valid: n bytes (1 valid, 0 invalid); status: n bytes (SIGOPS_ED_*) or NULL.
   Replaces ed25519_eddsa::ecverify / ecverify_single (src/ed25519_eddsa.rs:67-324), whose
   messages are fixed 32-byte digests (a special case: msg_off[i] = 32 i). */
int sigops_ed25519_verify(sigops_ctx* ctx, const uint8_t* sigs, const uint8_t* pubkeys, const uint8_t* msgs,
                          const uint64_t* msg_off, size_t n, uint8_t* valid, uint8_t* status);

/* Ed25519 verification with ed25519-dalek 2.1 `VerifyingKey::verify_strict` rules: `verify`, plus
   R must decompress and neither R nor A may have small order (the adjacent Fuel predicate the
   reference names, src/curve_algos/ed25519_eddsa.rs:197-198). Same layout as sigops_ed25519_verify;
   status adds SIGOPS_ED_BAD_R and SIGOPS_ED_SMALL_ORDER. No counterpart among the reference's
   entry points. */
int sigops_ed25519_verify_strict(sigops_ctx* ctx, const uint8_t* sigs, const uint8_t* pubkeys, const uint8_t* msgs,
                                 const uint64_t* msg_off, size_t n, uint8_t* valid, uint8_t* status);

/* ECDSA verification with a known public key, fuel `Signature::verify(&public_key, &message)`
   (the predicate next to recovery; fuel-crypto 0.49 verifies secp256k1 with libsecp256k1, which
   accepts lower-S signatures only; secp256r1 follows p256 0.13 / ecdsa 0.16 `verify_prehash`).
   sigs, digests as for recovery (the recid bit is ignored); pubkeys: n x 64 bytes X||Y big-endian;
   valid: n bytes (1 valid, 0 invalid); status: n bytes (SIGOPS_ST_*) or NULL.
   No counterpart among the reference's entry points: its tests verify by recovering and comparing
   keys (src/tests/secp256k1_ecdsa.rs:27-30). */
int sigops_secp256k1_verify(sigops_ctx* ctx, const uint8_t* sigs, const uint8_t* digests, const uint8_t* pubkeys,
                            size_t n, uint8_t* valid, uint8_t* status);
int sigops_secp256r1_verify(sigops_ctx* ctx, const uint8_t* sigs, const uint8_t* digests, const uint8_t* pubkeys,
                            size_t n, uint8_t* valid, uint8_t* status);

/* Caller-supplied fixed-base tables (`table_limbs` / `log_limb_size` of the reference's ecrecover
   and ecverify, src/secp256k1_ecdsa.rs:61-66, src/secp256r1_ecdsa.rs:62-67, src/ed25519_eddsa.rs:67-73).
   The table is the reference's layout (src/precompute.rs:12-69): [P, 2P, ..., 16P] affine, each
   coordinate x * 2^(nl*l) mod p in nl = 256 / l + 1 little-endian limbs of l = log_limb_size bits
   (11..15); secp entries (x, y), Ed25519 entries (x, y, x*y). The reference's stage-1 shaders use it
   as the fixed base, so the result is u1 * P + u2 * R (secp) or [s]P - [k]A (Ed25519) for the
   table's point P; these entry points convert it on the device (a 128-lane kernel that also builds
   the odd multiples) and run the same kernels on it. table_len is the number of u32 limbs. A table
   that is not [P, ..., 16P] for a curve point P (for Ed25519: of prime order) fails the call with
   SIGOPS_E_ARG; for the tables precompute::*_bases builds (P = G, B) the results equal the plain
   entry points'. */
int sigops_secp256k1_recover_table(sigops_ctx* ctx, const uint8_t* sigs, const uint8_t* digests, size_t n,
                                   const uint32_t* table_limbs, size_t table_len, uint32_t log_limb_size,
                                   uint8_t* out_pub, uint8_t* status);
int sigops_secp256r1_recover_table(sigops_ctx* ctx, const uint8_t* sigs, const uint8_t* digests, size_t n,
                                   const uint32_t* table_limbs, size_t table_len, uint32_t log_limb_size,
                                   uint8_t* out_pub, uint8_t* status);
int sigops_ed25519_verify_table(sigops_ctx* ctx, const uint8_t* sigs, const uint8_t* pubkeys, const uint8_t* msgs,
                                const uint64_t* msg_off, size_t n, const uint32_t* table_limbs, size_t table_len,
                                uint32_t log_limb_size, uint8_t* valid, uint8_t* status);

/* ---- device-resident variants: pointers are device memory, work is enqueued on `stream`
   (a hipStream_t, NULL for the default stream) and the call returns without synchronising.
   The context's scratch is shared, so calls on one context must not overlap. ---- */
int sigops_secp256k1_recover_dev(sigops_ctx* ctx, const void* d_sigs, const void* d_digests, size_t n,
                                 void* d_out_pub, void* d_status, void* stream);
int sigops_secp256r1_recover_dev(sigops_ctx* ctx, const void* d_sigs, const void* d_digests, size_t n,
                                 void* d_out_pub, void* d_status, void* stream);
/* d_status: n bytes SIGOPS_ST_* (0 = valid) */
int sigops_secp256k1_verify_dev(sigops_ctx* ctx, const void* d_sigs, const void* d_digests, const void* d_pubkeys,
                                size_t n, void* d_status, void* stream);
int sigops_secp256r1_verify_dev(sigops_ctx* ctx, const void* d_sigs, const void* d_digests, const void* d_pubkeys,
                                size_t n, void* d_status, void* stream);
/* d_msgs must be non-NULL and 4-byte aligned (SIGOPS_E_ARG otherwise) and stay readable for 16
   bytes past its last message byte (also when every message is empty) */
int sigops_ed25519_verify_dev(sigops_ctx* ctx, const void* d_sigs, const void* d_pubkeys, const void* d_msgs,
                              const uint64_t* d_msg_off, size_t n, void* d_status, void* stream);
int sigops_ed25519_verify_strict_dev(sigops_ctx* ctx, const void* d_sigs, const void* d_pubkeys, const void* d_msgs,
                                     const uint64_t* d_msg_off, size_t n, void* d_status, void* stream);

/* ---- mixed-scheme batch: one call for secp256k1 / secp256r1 recoveries and Ed25519 verifications
   in any order. No single counterpart in the reference (one entry point per scheme, 32-byte
   messages); item i is processed exactly as the per-scheme entry points would process it.
   schemes: n bytes (SIGOPS_SCHEME_*); sigs: n x 64; aux: n x 32 (the digest of a secp item, the
   public key of an Ed25519 item); message of item i = msgs[msg_off[i] .. msg_off[i+1]) (empty for
   secp items); out: n x 64 (recovered key of a secp item, zeros otherwise); status: n bytes
   (SIGOPS_ST_* for secp items, SIGOPS_ED_* for Ed25519 items). An unknown scheme byte fails the
   call with SIGOPS_E_ARG. ---- */
int sigops_batch(sigops_ctx* ctx, const uint8_t* schemes, const uint8_t* sigs, const uint8_t* aux,
                 const uint8_t* msgs, const uint64_t* msg_off, size_t n, uint8_t* out, uint8_t* status);
/* How sigops_batch moves a host batch (no counterpart in the reference). Streamed (the default,
   for n >= 4096): the host packs the batch into pinned, device-mapped memory chunk by chunk while
   one persistent launch already processes the published chunks, reading them over PCIe, and the
   results come back the same way as tiles finish, so packing, transfers and compute overlap.
   Off: the whole batch is copied to the device, processed, and copied back. Both give the same
   results; a streamed launch that cannot take every chunk (its waits are bounded) is redone the
   other way. */
int sigops_set_streaming(sigops_ctx* ctx, int on);
/* phases of the last streamed batch on this context, in milliseconds from the call's start (for
   profiling): [0] first chunks packed and the launch enqueued, [1] every chunk packed and
   published, [2] every chunk's results unpacked, [3] the launch (and its careful pass) complete,
   [4] the call done; [5] the launch's own duration on the GPU (HIP events). */
#define SIGOPS_HOST_PHASES 6
int sigops_batch_host_ms(sigops_ctx* ctx, float* ms);

/* device pointers (d_msgs non-NULL, 4-byte aligned and readable 16 bytes past the last message), enqueued
   on `stream`; the context's bucketing and lane scratch are shared, so calls on one context must
   not overlap on the device */
int sigops_batch_dev(sigops_ctx* ctx, const void* d_schemes, const void* d_sigs, const void* d_aux,
                     const void* d_msgs, const uint64_t* d_msg_off, size_t n, void* d_out, void* d_status,
                     void* stream);

/* ---- Fuel hashing on the device (SHA-256), the steps either side of the path ----
   Message::new (the step before recovery: fuel `Message::new(data)` = SHA-256(data), as the
   reference's tests and benchmarks build their digests, src/tests/ed25519_eddsa.rs:19,
   src/benchmarks/secp256k1_ecdsa.rs:162): digest i = SHA-256(data[off[i] .. off[i+1])).
   Addresses (the step after it): address i = SHA-256(X || Y) of a 64-byte public key, 32 zero
   bytes where the recovery failed (status != 0). No counterpart in the reference's API. */
int sigops_message_new(sigops_ctx* ctx, const uint8_t* data, const uint64_t* off, size_t n, uint8_t* digests);
/* device pointers, d_data 4-byte aligned and readable 16 bytes past its last byte; n <= max_batch.
   Messages are hashed longest first (a counting sort by block count on the device). The context's
   hashing work words and order are shared: calls on one context must not overlap on the device. */
int sigops_message_new_dev(sigops_ctx* ctx, const void* d_data, const uint64_t* d_off, size_t n, void* d_digests,
                           void* stream);
/* Message::new with no allocation. No counterpart in the reference: it is for callers that cannot
   allocate device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
   hipMalloc). Same work and digests as sigops_message_new_dev, but its kernels' scratch (the
   counting sort's work words and the longest-first order) is a fixed array in the library's device
   memory, sized for SIGOPS_NOALLOC_MAX_MESSAGES messages (4 MiB + 1 KiB), which every GPU the
   library runs on holds until the process ends; the call allocates nothing. n above the maximum
   fails with SIGOPS_E_ARG. Every call on one GPU, from any context, uses that GPU's one array, so
   calls that overlap on one GPU race. */
#define SIGOPS_NOALLOC_MAX_MESSAGES (1u << 20)
int sigops_message_new_noalloc(sigops_ctx* ctx, const void* d_data, const uint64_t* d_off, size_t n,
                               void* d_digests, void* stream);
/* d_status may be NULL (every key hashed) */
int sigops_address_from_pubkeys_dev(sigops_ctx* ctx, const void* d_pubs, const void* d_status, size_t n,
                                    void* d_addr, void* stream);
/* recovery followed by the address, both on the device; only n x 32 bytes come back */
int sigops_secp256k1_recover_address(sigops_ctx* ctx, const uint8_t* sigs, const uint8_t* digests, size_t n,
                                     uint8_t* out_addr, uint8_t* status);
int sigops_secp256r1_recover_address(sigops_ctx* ctx, const uint8_t* sigs, const uint8_t* digests, size_t n,
                                     uint8_t* out_addr, uint8_t* status);

/* ---- per-stage timing of the mixed batch (no counterpart in the reference; for profiling).
   With timing on, sigops_batch / sigops_batch_dev launch the stages eagerly (no graph) with HIP
   events recorded on the launch stream between them. sigops_batch_stage_ms waits for the last
   batch and writes SIGOPS_STAGES durations in milliseconds: [0] bucketing (clear, histogram,
   scan, scatter), [1] a separate Ed25519 hashing pass, now always empty (the persistent kernel
   hashes each Ed25519 tile's challenges itself; the slot keeps the indices stable), [2] the
   persistent verification kernel, [3] the careful pass for exceptional additions. ---- */
#define SIGOPS_STAGES 4
int sigops_set_stage_timing(sigops_ctx* ctx, int on);
int sigops_batch_stage_ms(sigops_ctx* ctx, float* ms);

/* ---- fault injection, for tests of the error path. The reference's only failure is the
   `success` canary a shader did not write (ShaderFailureError, src/lib.rs:12-14,
   src/secp256k1_ecdsa.rs:203-205); with SIGOPS_FAULT_SKIP_COMPUTE set, every recovery,
   verification and batch call on this context transfers its inputs and outputs but launches no
   compute kernel, so the host entry points find their status canaries unwritten and return
   SIGOPS_E_CANARY. flags = 0 turns it off. Test use only. ---- */
#define SIGOPS_FAULT_SKIP_COMPUTE 1u
int sigops_set_fault(sigops_ctx* ctx, uint32_t flags);

/* ---- serving: continuous batching of verification requests. No counterpart in the reference.
   Requests (each a small mixed batch in the sigops_batch layout) are queued on submit and leave the
   queue in arrival order while the bytes of the requests in flight stay within the HBM budget.
   Two modes:
   SIGOPS_SERVE_SLOTS: requests are packed into the next free slot (a context with its own stream
     and scratch) while the batch stays within max_batch items; arrivals during a launch go out
     with the next batch, and consecutive batches overlap on the device.
   SIGOPS_SERVE_RING: one persistent launch drains tile positions in pinned, device-mapped
     memory: the items of admitted requests are packed, by bucket, into tiles of up to 64 items
     that mix requests, and published to the running launch through one queue per scheme; each
     wavefront claims published tiles from its home scheme's queue (its CU's, so that a CU runs
     one scheme's code), and from the others only when its own has none. Full tiles go out at
     once, partial ones when wavefronts would otherwise idle. A launch ends after idle_close_us
     without work (or when a tile needs the careful pass, which runs after it) and the next
     request starts another.
   Thread-safe: any thread may submit or wait. ---- */
#define SIGOPS_SERVE_SLOTS 0
#define SIGOPS_SERVE_RING 1
typedef struct sigops_server sigops_server;
typedef struct {
  int device;
  uint32_t slots;          /* slot mode: batches in flight at once (0: 2) */
  uint64_t hbm_budget;     /* bytes of the requests in flight (0: 1 GiB) */
  uint32_t max_batch;      /* items per request, and per launched batch in slot mode (0: 1 << 20) */
  uint32_t blocks_per_cu;  /* slot mode: per-slot context setting (0: default) */
  uint32_t mode;           /* SIGOPS_SERVE_SLOTS (0) or SIGOPS_SERVE_RING */
  uint32_t ring_tiles;     /* ring mode: tile positions, a power of two in 64 .. 2^20 (0: 4096) */
  uint64_t ring_msg_bytes; /* ring mode: message bytes, pinned and in HBM, split evenly over the tile
                              positions (0: 256 MiB; at least 1 KiB per position, below 4 GiB); a
                              request holding a message longer than a position's share less 132
                              bytes is refused */
  uint32_t idle_close_us;  /* ring mode: a launch with no work for this long ends (0: 1000) */
} sigops_server_config;

sigops_server* sigops_server_create(const sigops_server_config* cfg);
/* finishes the queued and in-flight work, then releases everything */
void sigops_server_destroy(sigops_server* srv);
/* queues one request (inputs are copied before returning); returns a ticket >= 0, or SIGOPS_E_ARG
   for a bad request or one larger than max_batch items or the HBM budget */
int64_t sigops_server_submit(sigops_server* srv, const uint8_t* schemes, const uint8_t* sigs, const uint8_t* aux,
                             const uint8_t* msgs, const uint64_t* msg_off, size_t n);
/* waits for a ticket (timeout_ms < 0: no limit); on completion copies out (n x 64) and status (n),
   gives the submit and completion times (steady clock, ns) and releases the ticket.
   Returns 0, 1 on timeout (the ticket stays valid), or a negative SIGOPS_E_* code. */
int sigops_server_wait(sigops_server* srv, int64_t ticket, uint8_t* out, uint8_t* status, int timeout_ms,
                       uint64_t* t_submit_ns, uint64_t* t_done_ns);
/* batches launched (ring mode: persistent launches) and items processed so far */
int sigops_server_stats(sigops_server* srv, uint64_t* batches, uint64_t* items);
/* Load replay for latency measurement: native clients (4 threads, request r on thread r mod 4)
   submit request r (items req_off[r] .. req_off[r+1] of the given arrays, messages through the
   global msg_off) at arrival_ns[r] after the start; every result is collected into out / status
   (may be NULL) and latency_ns[r] = completion - scheduled arrival, so a client behind its
   schedule adds to the latency. */
int sigops_server_replay(sigops_server* srv, const uint8_t* schemes, const uint8_t* sigs, const uint8_t* aux,
                         const uint8_t* msgs, const uint64_t* msg_off, const uint64_t* req_off,
                         const uint64_t* arrival_ns, size_t n_requests, uint64_t* latency_ns, uint8_t* out,
                         uint8_t* status);

/* ---- synthetic workload generation on the GPU. No counterpart in the reference, which signs its
   benchmark inputs serially on the CPU with fuel-crypto / p256 / dalek
   (src/benchmarks/secp256k1_ecdsa.rs:150-172, src/benchmarks/ed25519_eddsa.rs:140-166); this signs
   millions of inputs in milliseconds for benchmarks and serving experiments. Item i is a pure
   function of (seed, i): sk and nonce uniform in [1, n-1], digest 32 uniform bytes, fuel-format
   low-s signature; Ed25519 keys from a 32-byte seed per item (RFC 8032). corrupt_ppm items (per
   million) get one uniform bit flipped in r/R, s or the digest/message; d_corrupted (n bytes, may
   be NULL) marks them. Device pointers, enqueued on `stream`. ---- */
int sigops_gen_secp_dev(sigops_ctx* ctx, int curve, uint64_t seed, size_t n, uint32_t corrupt_ppm, void* d_sigs,
                        void* d_digests, void* d_pubs, void* d_corrupted, void* stream);
/* messages are written at d_msgs + d_msg_off[i] (lengths from the offsets); d_msgs must be 4-byte
   aligned with 16 readable bytes after the last message */
int sigops_gen_ed25519_dev(sigops_ctx* ctx, uint64_t seed, size_t n, uint32_t corrupt_ppm, const uint64_t* d_msg_off,
                           void* d_msgs, void* d_sigs, void* d_pubkeys, void* d_corrupted, void* stream);

/* ---- VALU roofline probe. No counterpart in the reference. Measures, on `device`, the
   whole-GPU issue rate of v_mad_u64_u32 (the multiply-accumulate of the field arithmetic) and of
   v_fma_f64, in operations per second: the peak against which the signature kernels are
   reported. Takes a few milliseconds. ---- */
int sigops_probe_valu(int device, double* mad_per_s, double* fma64_per_s);

#ifdef __cplusplus
}
#endif
#endif
