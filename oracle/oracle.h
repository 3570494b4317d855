/*
 * oracle.h -- CPU restatement of the batch signature path, used ONLY as a checker.
 *
 * TEST INFRASTRUCTURE. Nothing in the product library links, loads or calls this code.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use it, and only
 * to check results or to time the CPU baseline.
 *
 * What it restates:
 *   - secp256k1 public-key recovery with fuel-crypto 0.49 / libsecp256k1 acceptance rules
 *     (reference call sites: src/tests/secp256k1_ecdsa.rs:27-30, src/benchmarks/secp256k1_ecdsa.rs:118-122;
 *      algorithm flow src/wgsl/secp256k1_ecdsa.wgsl:7-130, src/curve_algos/secp256k1_ecdsa.rs:66-128).
 *   - secp256r1 public-key recovery with fuel-crypto secp256r1::p256 / ecdsa 0.16 rules
 *     (src/tests/secp256r1_ecdsa.rs:29-35, src/curve_algos/secp256r1_ecdsa.rs:16-117).
 *   - Ed25519 verification with ed25519-dalek 2.1 `VerifyingKey::verify` rules
 *     (src/tests/ed25519_eddsa.rs:22-26, src/curve_algos/ed25519_eddsa.rs:16-184).
 *   - SHA-512 (FIPS 180-4; src/wgsl/sha512.wgsl:4-194 is the reference's one-block version)
 *     and SHA-256 (fuel `Message::new`, src/tests/ed25519_eddsa.rs:19).
 *   - Seeded signing (ECDSA with fuel's recid-in-s encoding, RFC 8032 Ed25519) to make inputs.
 *
 * Status codes match include/sigops.h (SIGOPS_ST_*).
 */
#ifndef SIGOPS_ORACLE_H
#define SIGOPS_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { OR_CURVE_K1 = 0, OR_CURVE_R1 = 1 };

/* status codes (must equal SIGOPS_ST_* in include/sigops.h) */
enum {
  OR_ST_OK = 0,
  OR_ST_BAD_R = 1,       /* r == 0 or r >= n */
  OR_ST_BAD_S = 2,       /* s == 0 or s >= n (after clearing the recid bit) */
  OR_ST_NO_POINT = 3,    /* x = r is not the x-coordinate of a curve point */
  OR_ST_INFINITY = 4,    /* recovered key (verification: u1*G + u2*Q) is the point at infinity */
  OR_ST_HIGH_S = 5,      /* verification, secp256k1: s > (n-1)/2 */
  OR_ST_BAD_KEY = 6,     /* verification: the public key is not a curve point with X, Y < p */
  OR_ST_MISMATCH = 7     /* verification: x(u1*G + u2*Q) mod n != r */
};

/* ---- hashes ---- */
void or_sha512(const uint8_t* msg, size_t len, uint8_t out[64]);
void or_sha256(const uint8_t* msg, size_t len, uint8_t out[32]);

/* ---- single-item primitives ---- */
int or_secp_recover(int curve, const uint8_t sig[64], const uint8_t digest[32], uint8_t out_pub[64]);
/* pub = sk*G as X||Y big-endian; returns 0 on success, -1 if sk == 0 or sk >= n */
int or_secp_pubkey(int curve, const uint8_t sk[32], uint8_t out_pub[64]);
/* fuel-format signature (r || s, parity of R.y in bit 7 of byte 32), low-s normalised.
   nonce must be in [1, n-1]; returns 0 on success, -1 if the nonce is unusable. */
int or_secp_sign(int curve, const uint8_t sk[32], const uint8_t digest[32], const uint8_t nonce[32],
                 uint8_t out_sig[64]);

/* ECDSA verification with a known public key X||Y (fuel Signature::verify); returns OR_ST_* */
int or_secp_verify(int curve, const uint8_t sig[64], const uint8_t digest[32], const uint8_t pub[64]);

int or_ed25519_verify(const uint8_t sig[64], const uint8_t pk[32], const uint8_t* msg, size_t len);
/* dalek VerifyingKey::verify_strict: verify plus R must decompress and neither R nor A may have
   small order (returns 1 if valid) */
int or_ed25519_verify_strict(const uint8_t sig[64], const uint8_t pk[32], const uint8_t* msg, size_t len);
/* 1 if the encoding decompresses (dalek rules) to a point of order dividing 8 */
int or_ed25519_is_small_order(const uint8_t enc[32]);
void or_ed25519_pubkey(const uint8_t seed[32], uint8_t out_pk[32]);
void or_ed25519_sign(const uint8_t seed[32], const uint8_t* msg, size_t len, uint8_t out_sig[64]);

/* ---- field / group primitives exposed so tests can pin them against closed forms ---- */
/* moduli: 0 = secp256k1 p, 1 = secp256k1 n, 2 = P-256 p, 3 = P-256 n, 4 = 2^255-19, 5 = l */
void or_mod_mul(int mod, const uint8_t a[32], const uint8_t b[32], uint8_t out[32]); /* BE, inputs < m */
void or_mod_inv(int mod, const uint8_t a[32], uint8_t out[32]);
void or_mod_pow(int mod, const uint8_t a[32], const uint8_t e[32], uint8_t out[32]);
/* sqrt for p = 3 mod 4 moduli (0, 2): returns 1 and the root if a is a square, else 0 */
int or_mod_sqrt(int mod, const uint8_t a[32], uint8_t out[32]);
/* 512-bit little-endian value mod l, output 32 bytes little-endian */
void or_reduce512_l(const uint8_t in[64], uint8_t out[32]);
/* k*P on secp curve; P affine X||Y BE (must be on the curve), returns 1 if result is
   infinity (out zeroed), else 0 */
int or_secp_mul(int curve, const uint8_t P[64], const uint8_t k[32], uint8_t out[64]);
/* u1*G + u2*P; same conventions */
int or_secp_mul2(int curve, const uint8_t P[64], const uint8_t u1[32], const uint8_t u2[32], uint8_t out[64]);
/* ed25519: decompress (dalek rules), returns 1 on success with affine x||y little-endian */
int or_ed25519_decompress(const uint8_t in[32], uint8_t out_xy[64]);

/* ---- batch wrappers (threaded; nthreads <= 0 means all online CPUs) ---- */
void or_secp_recover_batch(int curve, const uint8_t* sigs, const uint8_t* digests, size_t n,
                           uint8_t* out_pub, uint8_t* status, int nthreads);
void or_ed25519_verify_batch(const uint8_t* sigs, const uint8_t* pks, const uint8_t* msgs,
                             const uint64_t* msg_off, size_t n, uint8_t* valid, int nthreads);
void or_ed25519_verify_strict_batch(const uint8_t* sigs, const uint8_t* pks, const uint8_t* msgs,
                                    const uint64_t* msg_off, size_t n, uint8_t* valid, int nthreads);
void or_secp_verify_batch(int curve, const uint8_t* sigs, const uint8_t* digests, const uint8_t* pubs, size_t n,
                          uint8_t* status, int nthreads);

/* ---- seeded workload generation (deterministic, independent of thread count) ----
 * Item i draws from a per-item stream derived from (seed, i).
 * secp: sk uniform in [1,n-1]; digest 32 uniform bytes; fuel-format low-s signature.
 *       corrupt_ppm of the items are corrupted by flipping one uniform bit among r (256 bits),
 *       s (256 bits, recid bit included) and the digest (256 bits). corrupted[i] reports it.
 * ed25519: seed 32 uniform bytes -> RFC 8032 key; message bytes uniform with lengths given by
 *       the caller (msg_off); corrupt_ppm items get one bit flipped in R, s or M (field uniform,
 *       re-picked when M is empty). */
void or_gen_secp(int curve, uint64_t seed, size_t n, uint32_t corrupt_ppm, uint8_t* sigs,
                 uint8_t* digests, uint8_t* pubs, uint8_t* corrupted, int nthreads);
void or_gen_ed25519(uint64_t seed, size_t n, const uint64_t* msg_off, uint32_t corrupt_ppm,
                    uint8_t* sigs, uint8_t* pks, uint8_t* msgs, uint8_t* corrupted, int nthreads);

/* per-item stream: splitmix64 seeded with seed ^ (i * 0x9E3779B97F4A7C15) */
void or_item_random(uint64_t seed, uint64_t i, uint64_t tag, uint8_t* out, size_t len);

#ifdef __cplusplus
}
#endif
#endif
