/*
 * ossl_baseline.c -- CPU baseline timing with OpenSSL libcrypto, a tuned third-party
 * implementation, beside the oracle's plain restatement.
 *
 * BENCHMARK INFRASTRUCTURE. Nothing in the product library links, loads or calls this code; only
 * bench.py's cpu_baseline leg and the CPU tests use it.
 *
 * The reference's CPU path (fuel-crypto over libsecp256k1, p256, ed25519-dalek; timed serially in
 * src/benchmarks/secp256k1_ecdsa.rs:116-123) cannot be built here (Rust crates, no network), and
 * OpenSSL has no public-key recovery. So this times the closest operation OpenSSL offers for each
 * item on all the threads asked for:
 *   - secp256k1 / P-256: ECDSA_do_verify of (r, s) on the digest under the item's public key (a
 *     recovery costs one verification plus a square root);
 *   - Ed25519: EVP_DigestVerify (RFC 8032 verification of R || s on the message).
 * Each item builds its own key object, as a batch of independent signers needs. valid[i] gets
 * the verification result, so a caller can check it against the expected outcome.
 */
#define OPENSSL_SUPPRESS_DEPRECATED
#include <openssl/bn.h>
#include <openssl/ec.h>
#include <openssl/ecdsa.h>
#include <openssl/evp.h>
#include <openssl/obj_mac.h>
#include <pthread.h>
#include <stdint.h>
#include <string.h>
#include <time.h>

typedef struct {
  int curve;
  const uint8_t *sigs, *digs, *pubs, *msgs;
  const uint64_t* off;
  size_t a, b;
  uint8_t* valid;
} job_t;

static void* ecdsa_job(void* p) {
  job_t* j = (job_t*)p;
  EC_KEY* key = EC_KEY_new_by_curve_name(j->curve == 0 ? NID_secp256k1 : NID_X9_62_prime256v1);
  const EC_GROUP* grp = EC_KEY_get0_group(key);
  EC_POINT* pt = EC_POINT_new(grp);
  uint8_t enc[65];
  enc[0] = 4;
  for (size_t i = j->a; i < j->b; i++) {
    memcpy(enc + 1, j->pubs + 64 * i, 64);
    int ok = 0;
    if (EC_POINT_oct2point(grp, pt, enc, 65, NULL) == 1 && EC_KEY_set_public_key(key, pt) == 1) {
      uint8_t s[32];
      memcpy(s, j->sigs + 64 * i + 32, 32);
      s[0] &= 0x7F; /* fuel format: the recovery parity sits in bit 7 of byte 32 */
      ECDSA_SIG* sig = ECDSA_SIG_new();
      ECDSA_SIG_set0(sig, BN_bin2bn(j->sigs + 64 * i, 32, NULL), BN_bin2bn(s, 32, NULL));
      ok = ECDSA_do_verify(j->digs + 32 * i, 32, sig, key) == 1;
      ECDSA_SIG_free(sig);
    }
    j->valid[i] = (uint8_t)ok;
  }
  EC_POINT_free(pt);
  EC_KEY_free(key);
  return NULL;
}

static void* ed25519_job(void* p) {
  job_t* j = (job_t*)p;
  EVP_MD_CTX* ctx = EVP_MD_CTX_new();
  for (size_t i = j->a; i < j->b; i++) {
    int ok = 0;
    EVP_PKEY* pk = EVP_PKEY_new_raw_public_key(EVP_PKEY_ED25519, NULL, j->pubs + 32 * i, 32);
    if (pk && EVP_DigestVerifyInit(ctx, NULL, NULL, NULL, pk) == 1)
      ok = EVP_DigestVerify(ctx, j->sigs + 64 * i, 64, j->msgs + j->off[i], (size_t)(j->off[i + 1] - j->off[i])) == 1;
    EVP_PKEY_free(pk);
    EVP_MD_CTX_reset(ctx);
    j->valid[i] = (uint8_t)ok;
  }
  EVP_MD_CTX_free(ctx);
  return NULL;
}

static double run(job_t base, size_t n, int nthreads, void* (*fn)(void*)) {
  if (nthreads < 1) nthreads = 1;
  if (nthreads > 256) nthreads = 256;
  pthread_t th[256];
  job_t jobs[256];
  struct timespec t0, t1;
  clock_gettime(CLOCK_MONOTONIC, &t0);
  for (int t = 0; t < nthreads; t++) {
    jobs[t] = base;
    jobs[t].a = n * (size_t)t / (size_t)nthreads;
    jobs[t].b = n * (size_t)(t + 1) / (size_t)nthreads;
    pthread_create(&th[t], NULL, fn, &jobs[t]);
  }
  for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
  clock_gettime(CLOCK_MONOTONIC, &t1);
  return (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
}

/* seconds to verify n ECDSA signatures (curve 0 secp256k1, 1 P-256) on nthreads threads */
double ossl_ecdsa_verify_time(int curve, const uint8_t* sigs, const uint8_t* digs, const uint8_t* pubs, size_t n,
                              int nthreads, uint8_t* valid) {
  job_t j = {curve, sigs, digs, pubs, NULL, NULL, 0, 0, valid};
  return run(j, n, nthreads, ecdsa_job);
}

/* seconds to verify n Ed25519 signatures; message i = msgs[off[i] .. off[i+1]) */
double ossl_ed25519_verify_time(const uint8_t* sigs, const uint8_t* pubs, const uint8_t* msgs, const uint64_t* off,
                                size_t n, int nthreads, uint8_t* valid) {
  job_t j = {0, sigs, NULL, pubs, msgs, off, 0, 0, valid};
  return run(j, n, nthreads, ed25519_job);
}
