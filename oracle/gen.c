/*
 * gen.c -- threaded batch wrappers and seeded workload generation for the oracle
 * (test infrastructure only).
 *
 * The batch wrappers run the single-item oracle over a batch, one pthread per slice; they are what
 * the CPU baseline times (the reference times its CPU path serially, one signature after another,
 * src/benchmarks/secp256k1_ecdsa.rs:116-123, src/benchmarks/ed25519_eddsa.rs; here the thread count
 * is a parameter and 1 reproduces that method).
 *
 * The generators make the synthetic inputs the tests check against: each item's randomness comes
 * from a stream keyed by (seed, item index, purpose tag), so the output does not depend on the
 * number of threads and a GPU generator can reproduce it item by item.
 */
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>
#include "oracle_internal.h"

uint64_t splitmix64(uint64_t* s) {
  uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

/* stream state for (seed, i, tag) */
static uint64_t stream_init(uint64_t seed, uint64_t i, uint64_t tag) {
  uint64_t s = seed ^ (i * 0x9E3779B97F4A7C15ull) ^ (tag * 0xD1B54A32D192ED03ull);
  (void)splitmix64(&s);
  return s;
}

void or_item_random(uint64_t seed, uint64_t i, uint64_t tag, uint8_t* out, size_t len) {
  uint64_t s = stream_init(seed, i, tag);
  for (size_t off = 0; off < len; off += 8) {
    uint64_t w = splitmix64(&s);
    for (size_t j = 0; j < 8 && off + j < len; j++) out[off + j] = (uint8_t)(w >> (8 * j));
  }
}

typedef void (*item_fn)(void* ctx, size_t i);
typedef struct { item_fn fn; void* ctx; size_t n; size_t begin, end; } slice_t;

static void* run_slice(void* arg) {
  slice_t* s = (slice_t*)arg;
  for (size_t i = s->begin; i < s->end; i++) s->fn(s->ctx, i);
  return NULL;
}

static void parallel_for(size_t n, int nthreads, item_fn fn, void* ctx) {
  if (nthreads <= 0) nthreads = (int)sysconf(_SC_NPROCESSORS_ONLN);
  if (nthreads < 1) nthreads = 1;
  if ((size_t)nthreads > n) nthreads = (int)(n ? n : 1);
  if (nthreads == 1) {
    for (size_t i = 0; i < n; i++) fn(ctx, i);
    return;
  }
  pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)nthreads);
  slice_t* sl = (slice_t*)malloc(sizeof(slice_t) * (size_t)nthreads);
  size_t per = (n + (size_t)nthreads - 1) / (size_t)nthreads;
  for (int t = 0; t < nthreads; t++) {
    sl[t].fn = fn; sl[t].ctx = ctx; sl[t].n = n;
    sl[t].begin = (size_t)t * per < n ? (size_t)t * per : n;
    sl[t].end = sl[t].begin + per < n ? sl[t].begin + per : n;
    pthread_create(&th[t], NULL, run_slice, &sl[t]);
  }
  for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
  free(th);
  free(sl);
}

/* ---- batch recover / verify ---- */
typedef struct { int curve; const uint8_t *sigs, *digests; uint8_t *out, *status; } rec_ctx;
static void rec_item(void* c, size_t i) {
  rec_ctx* r = (rec_ctx*)c;
  int st = or_secp_recover(r->curve, r->sigs + 64 * i, r->digests + 32 * i, r->out + 64 * i);
  if (r->status) r->status[i] = (uint8_t)st;
}
void or_secp_recover_batch(int curve, const uint8_t* sigs, const uint8_t* digests, size_t n,
                           uint8_t* out_pub, uint8_t* status, int nthreads) {
  rec_ctx c = {curve, sigs, digests, out_pub, status};
  parallel_for(n, nthreads, rec_item, &c);
}

typedef struct { const uint8_t *sigs, *pks, *msgs; const uint64_t* off; uint8_t* valid; } ver_ctx;
static void ver_item(void* c, size_t i) {
  ver_ctx* v = (ver_ctx*)c;
  v->valid[i] = (uint8_t)or_ed25519_verify(v->sigs + 64 * i, v->pks + 32 * i, v->msgs + v->off[i],
                                           (size_t)(v->off[i + 1] - v->off[i]));
}
void or_ed25519_verify_batch(const uint8_t* sigs, const uint8_t* pks, const uint8_t* msgs,
                             const uint64_t* msg_off, size_t n, uint8_t* valid, int nthreads) {
  ver_ctx c = {sigs, pks, msgs, msg_off, valid};
  parallel_for(n, nthreads, ver_item, &c);
}

static void ver_strict_item(void* c, size_t i) {
  ver_ctx* v = (ver_ctx*)c;
  v->valid[i] = (uint8_t)or_ed25519_verify_strict(v->sigs + 64 * i, v->pks + 32 * i, v->msgs + v->off[i],
                                                  (size_t)(v->off[i + 1] - v->off[i]));
}
void or_ed25519_verify_strict_batch(const uint8_t* sigs, const uint8_t* pks, const uint8_t* msgs,
                                    const uint64_t* msg_off, size_t n, uint8_t* valid, int nthreads) {
  ver_ctx c = {sigs, pks, msgs, msg_off, valid};
  parallel_for(n, nthreads, ver_strict_item, &c);
}

typedef struct { int curve; const uint8_t *sigs, *digests, *pubs; uint8_t* status; } sver_ctx;
static void sver_item(void* c, size_t i) {
  sver_ctx* v = (sver_ctx*)c;
  v->status[i] = (uint8_t)or_secp_verify(v->curve, v->sigs + 64 * i, v->digests + 32 * i, v->pubs + 64 * i);
}
void or_secp_verify_batch(int curve, const uint8_t* sigs, const uint8_t* digests, const uint8_t* pubs, size_t n,
                          uint8_t* status, int nthreads) {
  sver_ctx c = {curve, sigs, digests, pubs, status};
  parallel_for(n, nthreads, sver_item, &c);
}

/* ---- generation ---- */
enum { TAG_SK = 1, TAG_DIGEST = 2, TAG_NONCE = 3, TAG_CORRUPT = 4, TAG_SEED = 5, TAG_MSG = 6 };

/* uniform draw in [1, n-1] by rejection on 256-bit words from the item stream */
static void draw_scalar(const modctx* Fn, uint64_t seed, uint64_t i, uint64_t tag, uint8_t out[32]) {
  for (uint64_t attempt = 0;; attempt++) {
    or_item_random(seed, i, tag + 16 * attempt, out, 32);
    u256 v;
    u256_from_be(&v, out);
    if (!u256_is_zero(&v) && u256_cmp(&v, &Fn->m) < 0) return;
  }
}

/* flip one bit chosen by the corruption stream; field 0 = r/R, 1 = s, 2 = digest/message */
static uint64_t corrupt_draw(uint64_t seed, uint64_t i, uint32_t ppm, int* hit) {
  uint64_t s = stream_init(seed, i, TAG_CORRUPT);
  uint64_t w = splitmix64(&s);
  *hit = (w % 1000000ull) < ppm;
  return splitmix64(&s);
}

typedef struct {
  int curve; uint64_t seed; uint32_t ppm;
  uint8_t *sigs, *digests, *pubs, *corrupted;
} gen_secp_ctx;

static void gen_secp_item(void* c, size_t i) {
  gen_secp_ctx* g = (gen_secp_ctx*)c;
  const curve_t* C = or_curve(g->curve);
  uint8_t sk[32], nonce[32];
  uint8_t* sig = g->sigs + 64 * i;
  uint8_t* dig = g->digests + 32 * i;
  draw_scalar(C->fn, g->seed, i, TAG_SK, sk);
  or_item_random(g->seed, i, TAG_DIGEST, dig, 32);
  or_secp_pubkey(g->curve, sk, g->pubs + 64 * i);
  for (uint64_t attempt = 0;; attempt++) {
    draw_scalar(C->fn, g->seed, i, TAG_NONCE + 256 * attempt, nonce);
    if (or_secp_sign(g->curve, sk, dig, nonce, sig) == 0) break;
  }
  int hit;
  uint64_t w = corrupt_draw(g->seed, i, g->ppm, &hit);
  if (g->corrupted) g->corrupted[i] = (uint8_t)hit;
  if (hit) {
    unsigned field = (unsigned)(w % 3), bit = (unsigned)((w >> 8) % 256);
    uint8_t* target = field == 0 ? sig : field == 1 ? sig + 32 : dig;
    target[bit >> 3] ^= (uint8_t)(1u << (bit & 7));
  }
}

void or_gen_secp(int curve, uint64_t seed, size_t n, uint32_t corrupt_ppm, uint8_t* sigs,
                 uint8_t* digests, uint8_t* pubs, uint8_t* corrupted, int nthreads) {
  gen_secp_ctx c = {curve, seed, corrupt_ppm, sigs, digests, pubs, corrupted};
  parallel_for(n, nthreads, gen_secp_item, &c);
}

typedef struct {
  uint64_t seed; uint32_t ppm; const uint64_t* off;
  uint8_t *sigs, *pks, *msgs, *corrupted;
} gen_ed_ctx;

static void gen_ed_item(void* c, size_t i) {
  gen_ed_ctx* g = (gen_ed_ctx*)c;
  uint8_t seed32[32];
  or_item_random(g->seed, i, TAG_SEED, seed32, 32);
  size_t len = (size_t)(g->off[i + 1] - g->off[i]);
  uint8_t* msg = g->msgs + g->off[i];
  or_item_random(g->seed, i, TAG_MSG, msg, len);
  uint8_t* sig = g->sigs + 64 * i;
  or_ed25519_pubkey(seed32, g->pks + 32 * i);
  or_ed25519_sign(seed32, msg, len, sig);
  int hit;
  uint64_t w = corrupt_draw(g->seed, i, g->ppm, &hit);
  if (g->corrupted) g->corrupted[i] = (uint8_t)hit;
  if (hit) {
    unsigned field = (unsigned)(w % 3);
    if (field == 2 && len == 0) field = (unsigned)((w >> 40) % 2);
    if (field == 2) {
      uint64_t bit = (w >> 8) % (8 * (uint64_t)len);
      msg[bit >> 3] ^= (uint8_t)(1u << (bit & 7));
    } else {
      unsigned bit = (unsigned)((w >> 8) % 256);
      uint8_t* target = field == 0 ? sig : sig + 32;
      target[bit >> 3] ^= (uint8_t)(1u << (bit & 7));
    }
  }
}

void or_gen_ed25519(uint64_t seed, size_t n, const uint64_t* msg_off, uint32_t corrupt_ppm,
                    uint8_t* sigs, uint8_t* pks, uint8_t* msgs, uint8_t* corrupted, int nthreads) {
  gen_ed_ctx c = {seed, corrupt_ppm, msg_off, sigs, pks, msgs, corrupted};
  parallel_for(n, nthreads, gen_ed_item, &c);
}
