// sigops_api.hip -- host runtime behind include/sigops.h.
//
// Replaces the reference's per-call wgpu life cycle (src/gpu.rs:5-186: adapter + device creation,
// WGSL compilation, buffer creation, dispatch, staging readback and device.destroy on every call,
// src/secp256k1_ecdsa.rs:79-80) with a context that is created once: code objects are loaded with the library, the
// generator tables are uploaded to HBM once, and per-lane scratch plus I/O buffers are reused.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <memory>
#include <mutex>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/sigops.h"
#include "consts.hpp"
#include "host_xfer.hpp"
#include "kernels.hpp"
#include "stream_host.hpp"

// ---- scratch of sigops_message_new_noalloc ----
// The hashing kernels' work words (the first NOALLOC_WORK_BYTES) and the longest-first order of up
// to SIGOPS_NOALLOC_MAX_MESSAGES messages, in a fixed array of the library's device memory instead
// of context buffers; every GPU the library runs on holds its own copy until the process ends. On
// the host the array's name is only a stand-in, never a device address: sigops_create looks up the
// array's address on its device with hipGetSymbolAddress (once per device, that device current),
// and the entry point passes that address to the kernels.
constexpr size_t NOALLOC_WORK_BYTES = 1024;
static_assert(sigops::MSG_WORK_WORDS * sizeof(uint32_t) <= NOALLOC_WORK_BYTES, "work words fit");
static __device__ __align__(256) uint8_t scratch[NOALLOC_WORK_BYTES + (size_t)SIGOPS_NOALLOC_MAX_MESSAGES * 4];

namespace {

constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void* g_scratch[NOALLOC_MAX_DEVICES];  // device address of `scratch` per device, 0 until looked up

constexpr uint8_t CANARY = 0xEE;
// largest caller table: 16 entries x 3 coordinates x 24 limbs (l = 11)
constexpr size_t TABLE_LIMBS_MAX = 16 * 3 * 24;

struct DevBuf {
  void* p = nullptr;
  size_t cap = 0;
};

}  // namespace

struct sigops_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  uint32_t grid = 0;
  uint32_t slots = 0;
  uint32_t max_batch = 1u << 22;
  void* d_gtab_k1 = nullptr;
  void* d_gtab_r1 = nullptr;
  void* d_btab = nullptr;
  void* d_lane_tab = nullptr;
  uint32_t* d_pend = nullptr;  // [0] = count, [1 ..] = list
  // mixed batch
  uint32_t tile_rounds = 1;
  uint32_t waves = 2;  // waves per SIMD the mixed batch's persistent kernel is built for
  uint32_t g_window = 8;  // the mixed batch's fixed-base window (secp streams), 8 or 9 bits
  bool use_graph = true;
  uint8_t* d_rank = nullptr;
  uint32_t* d_work = nullptr;
  uint32_t* d_order = nullptr;
  sigops::BatchIO* d_io = nullptr;  // the current batch call's buffers, read by its kernels
  uint32_t* d_hq = nullptr;      // work words of the message-hashing launch (MSG_WORK_WORDS)
  uint32_t* d_horder = nullptr;  // max_batch words: longest-first order of the messages
  // caller-supplied tables in the reference's layout: limbs in, converted table out, check flags
  uint32_t* d_tlimbs = nullptr;
  void* d_ctab = nullptr;
  uint32_t* d_tflags = nullptr;
  // the captured mixed-batch launches; they read the call's buffers from d_io, so one graph serves
  // every call on this context
  hipGraphExec_t graph = nullptr;
  DevBuf in, aux, msg, off, out, st, sch;
  uint32_t fault = 0;  // SIGOPS_FAULT_* test switch
  // streamed host batches (stream_host.hpp): one pinned, device-mapped block carved into the
  // layout's regions, a pinned control block, the device counters, the batch tag
  bool streaming = true;
  uint32_t bucket_order[sigops::NBUCKETS];  // processing rank -> bucket (the rank table inverted)
  uint8_t* s_pin = nullptr;
  size_t s_cap = 0;
  uint8_t* s_ctl = nullptr;  // u64 ready, u32 stop, u32 err
  uint32_t* d_sctr = nullptr;
  uint32_t s_seq = 0;
  hipEvent_t s_ev = nullptr;
  hipEvent_t s_tev[2] = {};  // timing events around the streamed launch
  float s_ms[SIGOPS_HOST_PHASES] = {};  // phases of the last streamed batch
  std::vector<uint32_t> s_perm;
  std::vector<uint32_t> s_sorted;  // the batch in processing order (caller indices)
  std::vector<uint8_t> s_bkt;      // bucket of every item
  void* noalloc_scratch = nullptr;  // device address of `scratch` on this context's device
  bool stage_timing = false;
  bool stage_recorded = false;
  hipEvent_t stage_ev[SIGOPS_STAGES + 1] = {};
  std::vector<uint8_t> host_status;
  // host entry points: pinned staging of caller buffers, packed and unpacked by a thread pool
  std::unique_ptr<sigops_host::Pool> pool;
  sigops_host::Pinned h_up, h_down;
  std::vector<hipEvent_t> xfer_ev;
  std::string err;
};

namespace {

int fail(sigops_ctx* c, int code, const std::string& what) {
  if (c) c->err = what;
  return code;
}

int hip_fail(sigops_ctx* c, hipError_t e, const char* where) {
  return fail(c, SIGOPS_E_DEVICE, std::string(where) + ": " + hipGetErrorString(e));
}

#define HIPCHK(c, expr)                                     \
  do {                                                      \
    hipError_t e_ = (expr);                                 \
    if (e_ != hipSuccess) return hip_fail((c), e_, #expr); \
  } while (0)

int ensure(sigops_ctx* c, DevBuf& b, size_t bytes) {
  if (b.cap >= bytes) return SIGOPS_OK;
  if (b.p) (void)hipFree(b.p);
  b.p = nullptr;
  b.cap = 0;
  const size_t want = std::max<size_t>(bytes, 4096);
  hipError_t e = hipMalloc(&b.p, want);
  if (e != hipSuccess) return fail(c, SIGOPS_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  b.cap = want;
  return SIGOPS_OK;
}

sigops_host::Pool& host_pool(sigops_ctx* c) {
  if (!c->pool) c->pool = std::make_unique<sigops_host::Pool>(sigops_host::host_threads());
  return *c->pool;
}

// a failed transfer: SIGOPS_E_NOMEM when the pinned staging could not grow, else SIGOPS_E_DEVICE
int xfer_fail(sigops_ctx* c, hipError_t e, const char* where) {
  if (e == hipErrorOutOfMemory)
    return fail(c, SIGOPS_E_NOMEM, std::string(where) + ": pinned staging: " + hipGetErrorString(e));
  return hip_fail(c, e, where);
}

// caller buffers -> device buffers, queued on the context's stream
int put(sigops_ctx* c, const std::vector<sigops_host::Up>& segs) {
  const hipError_t e = sigops_host::upload(host_pool(c), c->h_up, segs, c->stream);
  return e == hipSuccess ? SIGOPS_OK : xfer_fail(c, e, "upload");
}

// device buffers -> caller buffers after the work queued on the context's stream; returns when
// they are in place
int get(sigops_ctx* c, const std::vector<sigops_host::Down>& segs) {
  const hipError_t e = sigops_host::download(host_pool(c), c->h_down, segs, c->stream, c->xfer_ev);
  return e == hipSuccess ? SIGOPS_OK : xfer_fail(c, e, "download");
}

// pred(i) for every i in [0, n), evaluated by the pool; false if any is false
template <class P>
bool all_of(sigops_ctx* c, size_t n, P pred) {
  std::atomic<bool> ok{true};
  host_pool(c).for_range(n, [&](size_t a, size_t b) {
    for (size_t i = a; i < b; i++)
      if (!pred(i)) {
        ok.store(false, std::memory_order_relaxed);
        return;
      }
  });
  return ok.load();
}

// message offsets: non-decreasing, each length below 2^32
bool offsets_ok(sigops_ctx* c, const uint64_t* off, size_t n) {
  return all_of(c, n, [&](size_t i) { return off[i + 1] >= off[i] && off[i + 1] - off[i] <= 0xFFFFFFFFull; });
}

__global__ void k_fill_bytes(uint8_t* p, uint8_t v, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// gtab: the fixed-base table to use (NULL: the resident table of G, with the context's window; a
// caller's converted table holds 128 entries, so 8-bit windows)
int secp_recover_dev(sigops_ctx* c, int curve, const void* d_sigs, const void* d_digs, size_t n, void* d_out,
                     void* d_status, hipStream_t s, const void* gtab = nullptr) {
  if (n > c->max_batch) return fail(c, SIGOPS_E_ARG, "n exceeds max_batch");
  if (n == 0) return SIGOPS_OK;
  if (c->fault & SIGOPS_FAULT_SKIP_COMPUTE) return SIGOPS_OK;
  const uint32_t w = gtab ? 8u : c->g_window;
  if (curve == 0)
    HIPCHK(c, sigops::launch_secp_recover_CurveK1(d_sigs, d_digs, d_out, d_status, (uint32_t)n,
                                                 gtab ? gtab : c->d_gtab_k1, c->d_lane_tab, c->slots, c->d_pend + 1,
                                                 c->d_pend, c->grid, w, s));
  else
    HIPCHK(c, sigops::launch_secp_recover_CurveR1(d_sigs, d_digs, d_out, d_status, (uint32_t)n,
                                                 gtab ? gtab : c->d_gtab_r1, c->d_lane_tab, c->slots, c->d_pend + 1,
                                                 c->d_pend, c->grid, w, s));
  return SIGOPS_OK;
}

// A caller-supplied table in the reference's layout (which: 0 k1, 1 r1, 2 ed25519) converted into
// c->d_ctab on the context's stream; fails with SIGOPS_E_ARG unless it is [P, 2P, ..., 16P] for a
// curve point P (Ed25519: of prime order) in l-bit Montgomery limbs.
int load_caller_table(sigops_ctx* c, int which, const uint32_t* table_limbs, size_t table_len, uint32_t lbits) {
  if (lbits < 11 || lbits > 15) return fail(c, SIGOPS_E_ARG, "log_limb_size must be in 11..15");
  const size_t coords = which == 2 ? 3 : 2;
  const size_t want = 16 * coords * (256 / lbits + 1);
  if (!table_limbs || table_len != want)
    return fail(c, SIGOPS_E_ARG, "table_limbs must hold 16 x " + std::to_string(coords) + " x " +
                                     std::to_string(256 / lbits + 1) + " limbs");
  hipStream_t s = c->stream;
  HIPCHK(c, hipMemcpyAsync(c->d_tlimbs, table_limbs, want * sizeof(uint32_t), hipMemcpyHostToDevice, s));
  if (which == 2)
    HIPCHK(c, sigops::launch_table_ed(c->d_tlimbs, lbits, c->d_ctab, c->d_tflags, s));
  else
    HIPCHK(c, sigops::launch_table_secp(which, c->d_tlimbs, lbits, c->d_ctab, c->d_tflags, s));
  uint32_t flags = 0;
  HIPCHK(c, hipMemcpyAsync(&flags, c->d_tflags, sizeof flags, hipMemcpyDeviceToHost, s));
  HIPCHK(c, hipStreamSynchronize(s));
  if (flags)
    return fail(c, SIGOPS_E_ARG,
                std::string("table_limbs is not [P, 2P, ..., 16P] for a curve point P") +
                    ((flags & 1) ? " (a limb, value or curve equation is wrong)" : "") +
                    ((flags & 2) ? " (an entry is not the next multiple)" : "") +
                    ((flags & 8) ? " (P is not in the prime-order subgroup)" : ""));
  return SIGOPS_OK;
}

// out_pub: n x 64 recovered keys; or, with out_addr set instead, n x 32 addresses SHA-256(X || Y)
// computed on the device after the recovery (only the addresses travel back)
int secp_recover_host(sigops_ctx* c, int curve, const uint8_t* sigs, const uint8_t* digs, size_t n,
                      uint8_t* out_pub, uint8_t* status, uint8_t* out_addr = nullptr, const void* gtab = nullptr) {
  if (!c) return SIGOPS_E_ARG;
  c->err.clear();
  if (n == 0) return SIGOPS_OK;
  if (!sigs || !digs || !(out_pub || out_addr)) return fail(c, SIGOPS_E_ARG, "NULL buffer");
  if (n > c->max_batch) return fail(c, SIGOPS_E_ARG, "n exceeds max_batch");
  int rc;
  if ((rc = ensure(c, c->in, n * 64)) || (rc = ensure(c, c->aux, n * 32)) || (rc = ensure(c, c->out, n * 64)) ||
      (rc = ensure(c, c->st, n)))
    return rc;
  hipStream_t s = c->stream;
  if ((rc = put(c, {{c->in.p, sigs, n * 64}, {c->aux.p, digs, n * 32}}))) return rc;
  k_fill_bytes<<<64, 256, 0, s>>>((uint8_t*)c->st.p, CANARY, n);
  HIPCHK(c, hipGetLastError());
  if ((rc = secp_recover_dev(c, curve, c->in.p, c->aux.p, n, c->out.p, c->st.p, s, gtab))) return rc;
  uint8_t* st = status;
  if (!st) {
    c->host_status.resize(n);
    st = c->host_status.data();
  }
  if (out_addr) {
    // the digests in aux are consumed by the recovery launched before on this stream
    HIPCHK(c, sigops::launch_fuel_address(c->out.p, c->st.p, (uint32_t)n, c->aux.p, c->grid, s));
    rc = get(c, {{out_addr, c->aux.p, n * 32}, {st, c->st.p, n}});
  } else {
    rc = get(c, {{out_pub, c->out.p, n * 64}, {st, c->st.p, n}});
  }
  if (rc) return rc;
  if (!all_of(c, n, [&](size_t i) { return st[i] <= SIGOPS_ST_INFINITY; }))
    return fail(c, SIGOPS_E_CANARY, "a recovery result was not written");
  return SIGOPS_OK;
}

int secp_verify_dev(sigops_ctx* c, int curve, const void* d_sigs, const void* d_digs, const void* d_pubs, size_t n,
                    void* d_status, hipStream_t s) {
  if (n > c->max_batch) return fail(c, SIGOPS_E_ARG, "n exceeds max_batch");
  if (n == 0) return SIGOPS_OK;
  if (!d_sigs || !d_digs || !d_pubs || !d_status) return fail(c, SIGOPS_E_ARG, "NULL buffer");
  if (c->fault & SIGOPS_FAULT_SKIP_COMPUTE) return SIGOPS_OK;
  if (curve == 0)
    HIPCHK(c, sigops::launch_secp_verify_CurveK1(d_sigs, d_digs, d_pubs, d_status, (uint32_t)n, c->d_gtab_k1,
                                                c->d_lane_tab, c->slots, c->d_pend + 1, c->d_pend, c->grid,
                                                c->g_window, s));
  else
    HIPCHK(c, sigops::launch_secp_verify_CurveR1(d_sigs, d_digs, d_pubs, d_status, (uint32_t)n, c->d_gtab_r1,
                                                c->d_lane_tab, c->slots, c->d_pend + 1, c->d_pend, c->grid,
                                                c->g_window, s));
  return SIGOPS_OK;
}

int secp_verify_host(sigops_ctx* c, int curve, const uint8_t* sigs, const uint8_t* digs, const uint8_t* pubs,
                     size_t n, uint8_t* valid, uint8_t* status) {
  if (!c) return SIGOPS_E_ARG;
  c->err.clear();
  if (n == 0) return SIGOPS_OK;
  if (!sigs || !digs || !pubs || !valid) return fail(c, SIGOPS_E_ARG, "NULL buffer");
  if (n > c->max_batch) return fail(c, SIGOPS_E_ARG, "n exceeds max_batch");
  int rc;
  if ((rc = ensure(c, c->in, n * 64)) || (rc = ensure(c, c->aux, n * 32)) || (rc = ensure(c, c->out, n * 64)) ||
      (rc = ensure(c, c->st, n)))
    return rc;
  hipStream_t s = c->stream;
  if ((rc = put(c, {{c->in.p, sigs, n * 64}, {c->aux.p, digs, n * 32}, {c->out.p, pubs, n * 64}}))) return rc;
  k_fill_bytes<<<64, 256, 0, s>>>((uint8_t*)c->st.p, CANARY, n);
  HIPCHK(c, hipGetLastError());
  if ((rc = secp_verify_dev(c, curve, c->in.p, c->aux.p, c->out.p, n, c->st.p, s))) return rc;
  uint8_t* st = status;
  if (!st) {
    c->host_status.resize(n);
    st = c->host_status.data();
  }
  if ((rc = get(c, {{st, c->st.p, n}}))) return rc;
  if (!all_of(c, n, [&](size_t i) {
        valid[i] = st[i] == SIGOPS_ST_OK;
        return st[i] <= SIGOPS_ST_MISMATCH;
      }))
    return fail(c, SIGOPS_E_CANARY, "a verification result was not written");
  return SIGOPS_OK;
}

}  // namespace

extern "C" {

sigops_ctx* sigops_create(const sigops_config* cfg) {
  sigops_ctx* c = new (std::nothrow) sigops_ctx();
  if (!c) return nullptr;
  sigops_config def{};
  if (!cfg) cfg = &def;
  c->device = cfg->device;
  if (cfg->max_batch) c->max_batch = cfg->max_batch;
  auto bad = [&](hipError_t e, const char* what) {
    fprintf(stderr, "sigops_create: %s: %s\n", what, hipGetErrorString(e));
    sigops_destroy(c);
    return (sigops_ctx*)nullptr;
  };
  hipError_t e;
  if ((e = hipSetDevice(c->device)) != hipSuccess) return bad(e, "hipSetDevice");
  if (c->device >= 0 && c->device < NOALLOC_MAX_DEVICES) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[c->device]) {
      void* p = nullptr;
      if ((e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch))) != hipSuccess) return bad(e, "hipGetSymbolAddress");
      g_scratch[c->device] = p;
    }
    c->noalloc_scratch = g_scratch[c->device];
  }
  int cus = 0;
  if ((e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device)) != hipSuccess)
    return bad(e, "hipDeviceGetAttribute");
  if (cfg->waves_per_simd > 2 || cfg->bucket_order > 2 || (cfg->g_window && cfg->g_window != 8 && cfg->g_window != 9) ||
      (cfg->g_window == 9 && cfg->waves_per_simd == 1)) {
    fprintf(stderr,
            "sigops_create: waves_per_simd must be 0..2, bucket_order 0..2 and g_window 0, 8 or 9 (9 with 2 waves)\n");
    sigops_destroy(c);
    return nullptr;
  }
  c->waves = cfg->waves_per_simd ? cfg->waves_per_simd : 2;
  // 9-bit windows measured faster on the MI355X (and chosen by the autotuner); the one-wave kernel
  // is built with 8-bit ones only
  c->g_window = cfg->g_window ? cfg->g_window : (c->waves == 1 ? 8 : 9);
  const uint32_t bpc = cfg->blocks_per_cu ? cfg->blocks_per_cu : c->waves;
  c->grid = (uint32_t)cus * bpc;
  c->slots = c->grid * 256;
  if ((e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) != hipSuccess) return bad(e, "stream");
  if ((e = hipMalloc(&c->d_gtab_k1, sizeof(sigops::K1_GTABLE))) != hipSuccess) return bad(e, "hipMalloc");
  if ((e = hipMalloc(&c->d_gtab_r1, sizeof(sigops::R1_GTABLE))) != hipSuccess) return bad(e, "hipMalloc");
  if ((e = hipMalloc(&c->d_btab, sizeof(sigops::ED_BTABLE))) != hipSuccess) return bad(e, "hipMalloc");
  if ((e = hipMemcpy(c->d_gtab_k1, sigops::K1_GTABLE, sizeof(sigops::K1_GTABLE), hipMemcpyHostToDevice)) !=
      hipSuccess)
    return bad(e, "upload");
  if ((e = hipMemcpy(c->d_gtab_r1, sigops::R1_GTABLE, sizeof(sigops::R1_GTABLE), hipMemcpyHostToDevice)) !=
      hipSuccess)
    return bad(e, "upload");
  if ((e = hipMemcpy(c->d_btab, sigops::ED_BTABLE, sizeof(sigops::ED_BTABLE), hipMemcpyHostToDevice)) != hipSuccess)
    return bad(e, "upload");
  const size_t lane_bytes =
      (size_t)c->slots * std::max(sigops::SECP_LANE_TAB_BYTES, sigops::ED_LANE_TAB_BYTES);
  if ((e = hipMalloc(&c->d_lane_tab, lane_bytes)) != hipSuccess) return bad(e, "hipMalloc lane tables");
  if ((e = hipMalloc((void**)&c->d_pend, ((size_t)c->max_batch + 1) * sizeof(uint32_t))) != hipSuccess)
    return bad(e, "hipMalloc pending list");
  if (cfg->tile_rounds) c->tile_rounds = std::min<uint32_t>(cfg->tile_rounds, sigops::MAX_TILE_ROUNDS);
  c->use_graph = cfg->no_graph == 0;
  {
    // processing order of the buckets (higher key first). Default: secp256r1, then secp256k1, then
    // Ed25519 by decreasing SHA-512 block count. Measured on MI355X, alternating builds on one box,
    // for the 2^20 mixed batch: 16.30 ms in this order; 16.55 ms for an estimated-cost order that
    // put the few longest Ed25519 buckets first and the short ones after secp256k1; 16.67 ms with
    // secp256k1 before secp256r1 (order 2); about 17 ms with the Ed25519 buckets first. Ending on
    // the cheapest tiles (secp256r1, Ed25519, then secp256k1: order 1) shortens the final tiles
    // but took 16.15-16.30 ms against 15.87-15.88 ms for the default (4 alternations on one box).
    // The autotuner measures the three again per device.
    static const double KEYS[3][2] = {{2.0, 3.0}, {0.5, 3.0}, {3.0, 2.0}};  // (secp256k1, secp256r1)
    std::vector<std::pair<double, uint32_t>> key;
    key.push_back({KEYS[cfg->bucket_order][0], sigops::BUCKET_K1});
    key.push_back({KEYS[cfg->bucket_order][1], sigops::BUCKET_R1});
    for (uint32_t b = 0; b < sigops::ED_BINS; b++) key.push_back({1.0 + 0.001 * (b + 1), sigops::BUCKET_ED0 + b});
    std::stable_sort(key.begin(), key.end(), [](auto& x, auto& y) { return x.first > y.first; });
    uint8_t rank[sigops::NBUCKETS];
    for (uint32_t r = 0; r < sigops::NBUCKETS; r++) rank[key[r].second] = (uint8_t)r;
    for (uint32_t r = 0; r < sigops::NBUCKETS; r++) c->bucket_order[r] = key[r].second;
    if ((e = hipMalloc((void**)&c->d_rank, sizeof rank)) != hipSuccess) return bad(e, "hipMalloc rank");
    if ((e = hipMemcpy(c->d_rank, rank, sizeof rank, hipMemcpyHostToDevice)) != hipSuccess) return bad(e, "upload");
  }
  if ((e = hipMalloc((void**)&c->d_work, sigops::BATCH_WORK_WORDS * sizeof(uint32_t))) != hipSuccess)
    return bad(e, "hipMalloc batch work");
  if ((e = hipMalloc((void**)&c->d_order, ((size_t)c->max_batch + (size_t)sigops::NBUCKETS * 64 *
                                                                      sigops::MAX_TILE_ROUNDS) * sizeof(uint32_t))) !=
      hipSuccess)
    return bad(e, "hipMalloc batch order");
  if ((e = hipMalloc((void**)&c->d_io, sizeof(sigops::BatchIO))) != hipSuccess) return bad(e, "hipMalloc batch io");
  if ((e = hipMalloc((void**)&c->d_hq, sigops::MSG_WORK_WORDS * sizeof(uint32_t))) != hipSuccess ||
      (e = hipMalloc((void**)&c->d_horder, (size_t)c->max_batch * sizeof(uint32_t))) != hipSuccess)
    return bad(e, "hipMalloc message hashing");
  if ((e = hipMalloc((void**)&c->d_tlimbs, TABLE_LIMBS_MAX * sizeof(uint32_t))) != hipSuccess ||
      (e = hipMalloc(&c->d_ctab, sizeof(sigops::ED_BTABLE))) != hipSuccess ||
      (e = hipMalloc((void**)&c->d_tflags, sizeof(uint32_t))) != hipSuccess)
    return bad(e, "hipMalloc caller tables");
  return c;
}

void sigops_destroy(sigops_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->graph) (void)hipGraphExecDestroy(c->graph);
  if (c->s_ev) (void)hipEventDestroy(c->s_ev);
  for (hipEvent_t e : c->s_tev)
    if (e) (void)hipEventDestroy(e);
  if (c->s_pin) (void)hipHostFree(c->s_pin);
  if (c->s_ctl) (void)hipHostFree(c->s_ctl);
  if (c->d_sctr) (void)hipFree(c->d_sctr);
  for (hipEvent_t e : c->stage_ev)
    if (e) (void)hipEventDestroy(e);
  for (void* p : {c->d_gtab_k1, c->d_gtab_r1, c->d_btab, c->d_lane_tab, (void*)c->d_pend, c->in.p, c->aux.p,
                  c->msg.p, c->off.p, c->out.p, c->st.p, c->sch.p, (void*)c->d_rank, (void*)c->d_work,
                  (void*)c->d_order, (void*)c->d_io, (void*)c->d_hq, (void*)c->d_horder, (void*)c->d_tlimbs,
                  c->d_ctab, (void*)c->d_tflags})
    if (p) (void)hipFree(p);
  c->pool.reset();
  sigops_host::pinned_free(c->h_up);
  sigops_host::pinned_free(c->h_down);
  for (hipEvent_t e : c->xfer_ev) (void)hipEventDestroy(e);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

const char* sigops_last_error(const sigops_ctx* c) { return c ? c->err.c_str() : "no context"; }

int sigops_secp256k1_recover(sigops_ctx* c, const uint8_t* sigs, const uint8_t* digests, size_t n,
                             uint8_t* out_pub, uint8_t* status) {
  return secp_recover_host(c, 0, sigs, digests, n, out_pub, status);
}

int sigops_secp256r1_recover(sigops_ctx* c, const uint8_t* sigs, const uint8_t* digests, size_t n,
                             uint8_t* out_pub, uint8_t* status) {
  return secp_recover_host(c, 1, sigs, digests, n, out_pub, status);
}

int sigops_secp256k1_recover_dev(sigops_ctx* c, const void* d_sigs, const void* d_digests, size_t n,
                                 void* d_out_pub, void* d_status, void* stream) {
  if (!c) return SIGOPS_E_ARG;
  return secp_recover_dev(c, 0, d_sigs, d_digests, n, d_out_pub, d_status, (hipStream_t)stream);
}

int sigops_secp256r1_recover_dev(sigops_ctx* c, const void* d_sigs, const void* d_digests, size_t n,
                                 void* d_out_pub, void* d_status, void* stream) {
  if (!c) return SIGOPS_E_ARG;
  return secp_recover_dev(c, 1, d_sigs, d_digests, n, d_out_pub, d_status, (hipStream_t)stream);
}

int sigops_secp256k1_verify(sigops_ctx* c, const uint8_t* sigs, const uint8_t* digests, const uint8_t* pubkeys,
                            size_t n, uint8_t* valid, uint8_t* status) {
  return secp_verify_host(c, 0, sigs, digests, pubkeys, n, valid, status);
}

int sigops_secp256r1_verify(sigops_ctx* c, const uint8_t* sigs, const uint8_t* digests, const uint8_t* pubkeys,
                            size_t n, uint8_t* valid, uint8_t* status) {
  return secp_verify_host(c, 1, sigs, digests, pubkeys, n, valid, status);
}

int sigops_secp256k1_verify_dev(sigops_ctx* c, const void* d_sigs, const void* d_digests, const void* d_pubkeys,
                                size_t n, void* d_status, void* stream) {
  if (!c) return SIGOPS_E_ARG;
  return secp_verify_dev(c, 0, d_sigs, d_digests, d_pubkeys, n, d_status, (hipStream_t)stream);
}

int sigops_secp256r1_verify_dev(sigops_ctx* c, const void* d_sigs, const void* d_digests, const void* d_pubkeys,
                                size_t n, void* d_status, void* stream) {
  if (!c) return SIGOPS_E_ARG;
  return secp_verify_dev(c, 1, d_sigs, d_digests, d_pubkeys, n, d_status, (hipStream_t)stream);
}

// Ed25519 verification, dalek `verify` rules or (strict) `verify_strict` rules; btab: the
// fixed-base table (NULL: the resident table of B). The items are bucketed by SHA-512 block count
// and a persistent kernel pops tiles of one bucket, hashes their challenges and verifies them, so
// a wavefront's lanes hash equally long messages.
static int ed_verify_dev(sigops_ctx* c, const void* d_sigs, const void* d_pubkeys, const void* d_msgs,
                         const uint64_t* d_msg_off, size_t n, void* d_status, void* stream, bool strict,
                         const void* btab = nullptr) {
  if (!c) return SIGOPS_E_ARG;
  if (n > c->max_batch) return fail(c, SIGOPS_E_ARG, "n exceeds max_batch");
  if (n == 0) return SIGOPS_OK;
  // the SHA-512 reader loads the word at each message's offset even for an empty message
  if (!d_sigs || !d_pubkeys || !d_msgs || !d_msg_off || !d_status) return fail(c, SIGOPS_E_ARG, "NULL buffer");
  if (((uintptr_t)d_msgs & 3u) != 0) return fail(c, SIGOPS_E_ARG, "d_msgs must be 4-byte aligned");
  hipStream_t s = (hipStream_t)stream;
  HIPCHK(c, sigops::launch_set_batch_io(c->d_io, {nullptr, d_sigs, d_pubkeys, d_msgs, d_msg_off, nullptr, d_status,
                                                  (uint32_t)n}, s));
  sigops::BatchLaunch L{};
  L.io = c->d_io;
  L.work = c->d_work;
  L.order = c->d_order;
  L.pend_list = c->d_pend + 1;
  L.rank = c->d_rank;
  L.gb = btab ? btab : c->d_btab;
  L.lane_tab = c->d_lane_tab;
  L.slots = c->slots;
  // every bucket may end in a partial tile: at most NBUCKETS tiles beyond n / 64, 4 tiles per block
  L.grid = std::min<uint32_t>(c->grid, (uint32_t)((n + 255) / 256 + (sigops::NBUCKETS + 3) / 4));
  L.aux_grid = std::min<uint32_t>(c->grid, (uint32_t)((n + 255) / 256));
  L.tile_rounds = 1;
  if (c->fault & SIGOPS_FAULT_SKIP_COMPUTE) return SIGOPS_OK;
  HIPCHK(c, sigops::launch_ed_verify(L, strict, s));
  return SIGOPS_OK;
}

int sigops_ed25519_verify_dev(sigops_ctx* c, const void* d_sigs, const void* d_pubkeys, const void* d_msgs,
                              const uint64_t* d_msg_off, size_t n, void* d_status, void* stream) {
  return ed_verify_dev(c, d_sigs, d_pubkeys, d_msgs, d_msg_off, n, d_status, stream, false);
}

int sigops_ed25519_verify_strict_dev(sigops_ctx* c, const void* d_sigs, const void* d_pubkeys, const void* d_msgs,
                                     const uint64_t* d_msg_off, size_t n, void* d_status, void* stream) {
  return ed_verify_dev(c, d_sigs, d_pubkeys, d_msgs, d_msg_off, n, d_status, stream, true);
}

static int ed_verify_host(sigops_ctx* c, const uint8_t* sigs, const uint8_t* pubkeys, const uint8_t* msgs,
                          const uint64_t* msg_off, size_t n, uint8_t* valid, uint8_t* status, bool strict,
                          const void* btab = nullptr) {
  if (!c) return SIGOPS_E_ARG;
  c->err.clear();
  if (n == 0) return SIGOPS_OK;
  if (!sigs || !pubkeys || !msg_off || !valid) return fail(c, SIGOPS_E_ARG, "NULL buffer");
  if (n > c->max_batch) return fail(c, SIGOPS_E_ARG, "n exceeds max_batch");
  if (!offsets_ok(c, msg_off, n))
    return fail(c, SIGOPS_E_ARG, "msg_off must be non-decreasing with lengths < 2^32");
  const uint64_t base = msg_off[0];
  const uint64_t total = msg_off[n] - base;
  if (total && !msgs) return fail(c, SIGOPS_E_ARG, "NULL msgs");
  int rc;
  if ((rc = ensure(c, c->in, n * 64)) || (rc = ensure(c, c->aux, n * 32)) ||
      (rc = ensure(c, c->msg, total + 16)) || (rc = ensure(c, c->off, (n + 1) * 8)) || (rc = ensure(c, c->st, n)))
    return rc;
  hipStream_t s = c->stream;
  if ((rc = put(c, {{c->in.p, sigs, n * 64},
                    {c->aux.p, pubkeys, n * 32},
                    {c->msg.p, total ? msgs + base : nullptr, total},
                    {c->off.p, msg_off, (n + 1) * 8, base}})))
    return rc;
  k_fill_bytes<<<1, 64, 0, s>>>((uint8_t*)c->msg.p + total, 0, 16);
  k_fill_bytes<<<64, 256, 0, s>>>((uint8_t*)c->st.p, CANARY, n);
  HIPCHK(c, hipGetLastError());
  if ((rc = ed_verify_dev(c, c->in.p, c->aux.p, c->msg.p, (const uint64_t*)c->off.p, n, c->st.p, s, strict, btab)))
    return rc;
  uint8_t* st = status;
  if (!st) {
    c->host_status.resize(n);
    st = c->host_status.data();
  }
  if ((rc = get(c, {{st, c->st.p, n}}))) return rc;
  const uint8_t last = strict ? SIGOPS_ED_SMALL_ORDER : SIGOPS_ED_MISMATCH;
  if (!all_of(c, n, [&](size_t i) {
        valid[i] = st[i] == SIGOPS_ED_OK;
        return st[i] <= last;
      }))
    return fail(c, SIGOPS_E_CANARY, "a verification result was not written");
  return SIGOPS_OK;
}

int sigops_ed25519_verify(sigops_ctx* c, const uint8_t* sigs, const uint8_t* pubkeys, const uint8_t* msgs,
                          const uint64_t* msg_off, size_t n, uint8_t* valid, uint8_t* status) {
  return ed_verify_host(c, sigs, pubkeys, msgs, msg_off, n, valid, status, false);
}

int sigops_ed25519_verify_strict(sigops_ctx* c, const uint8_t* sigs, const uint8_t* pubkeys, const uint8_t* msgs,
                                 const uint64_t* msg_off, size_t n, uint8_t* valid, uint8_t* status) {
  return ed_verify_host(c, sigs, pubkeys, msgs, msg_off, n, valid, status, true);
}

int sigops_ed25519_verify_table(sigops_ctx* c, const uint8_t* sigs, const uint8_t* pubkeys, const uint8_t* msgs,
                                const uint64_t* msg_off, size_t n, const uint32_t* table_limbs, size_t table_len,
                                uint32_t log_limb_size, uint8_t* valid, uint8_t* status) {
  if (!c) return SIGOPS_E_ARG;
  c->err.clear();
  int rc;
  if (n == 0) return SIGOPS_OK;  // nothing to do, as the reference returns early (src/secp256k1_ecdsa.rs:71-73)
  if ((rc = load_caller_table(c, 2, table_limbs, table_len, log_limb_size))) return rc;
  return ed_verify_host(c, sigs, pubkeys, msgs, msg_off, n, valid, status, false, c->d_ctab);
}

int sigops_secp256k1_recover_table(sigops_ctx* c, const uint8_t* sigs, const uint8_t* digests, size_t n,
                                   const uint32_t* table_limbs, size_t table_len, uint32_t log_limb_size,
                                   uint8_t* out_pub, uint8_t* status) {
  if (!c) return SIGOPS_E_ARG;
  c->err.clear();
  int rc;
  if (n == 0) return SIGOPS_OK;  // nothing to do, as the reference returns early (src/secp256k1_ecdsa.rs:71-73)
  if ((rc = load_caller_table(c, 0, table_limbs, table_len, log_limb_size))) return rc;
  return secp_recover_host(c, 0, sigs, digests, n, out_pub, status, nullptr, c->d_ctab);
}

int sigops_secp256r1_recover_table(sigops_ctx* c, const uint8_t* sigs, const uint8_t* digests, size_t n,
                                   const uint32_t* table_limbs, size_t table_len, uint32_t log_limb_size,
                                   uint8_t* out_pub, uint8_t* status) {
  if (!c) return SIGOPS_E_ARG;
  c->err.clear();
  int rc;
  if (n == 0) return SIGOPS_OK;  // nothing to do, as the reference returns early (src/secp256k1_ecdsa.rs:71-73)
  if ((rc = load_caller_table(c, 1, table_limbs, table_len, log_limb_size))) return rc;
  return secp_recover_host(c, 1, sigs, digests, n, out_pub, status, nullptr, c->d_ctab);
}

int sigops_gen_secp_dev(sigops_ctx* c, int curve, uint64_t seed, size_t n, uint32_t corrupt_ppm, void* d_sigs,
                        void* d_digests, void* d_pubs, void* d_corrupted, void* stream) {
  if (!c || (curve != 0 && curve != 1)) return SIGOPS_E_ARG;
  if (n > c->max_batch) return fail(c, SIGOPS_E_ARG, "n exceeds max_batch");
  if (n && (!d_sigs || !d_digests || !d_pubs)) return fail(c, SIGOPS_E_ARG, "NULL buffer");
  HIPCHK(c, sigops::launch_gen_secp(curve, seed, (uint32_t)n, corrupt_ppm, curve == 0 ? c->d_gtab_k1 : c->d_gtab_r1,
                                    d_sigs, d_digests, d_pubs, d_corrupted, c->grid, (hipStream_t)stream));
  return SIGOPS_OK;
}

int sigops_gen_ed25519_dev(sigops_ctx* c, uint64_t seed, size_t n, uint32_t corrupt_ppm, const uint64_t* d_msg_off,
                           void* d_msgs, void* d_sigs, void* d_pubkeys, void* d_corrupted, void* stream) {
  if (!c) return SIGOPS_E_ARG;
  if (n > c->max_batch) return fail(c, SIGOPS_E_ARG, "n exceeds max_batch");
  if (n && (!d_msg_off || !d_msgs || !d_sigs || !d_pubkeys)) return fail(c, SIGOPS_E_ARG, "NULL buffer");
  if (((uintptr_t)d_msgs & 3u) != 0) return fail(c, SIGOPS_E_ARG, "d_msgs must be 4-byte aligned");
  HIPCHK(c, sigops::launch_gen_ed25519(seed, (uint32_t)n, corrupt_ppm, c->d_btab, d_msg_off, d_msgs, d_sigs, d_pubkeys,
                                       d_corrupted, c->grid, (hipStream_t)stream));
  return SIGOPS_OK;
}

int sigops_batch_dev(sigops_ctx* c, const void* d_schemes, const void* d_sigs, const void* d_aux, const void* d_msgs,
                     const uint64_t* d_msg_off, size_t n, void* d_out, void* d_status, void* stream) {
  if (!c) return SIGOPS_E_ARG;
  if (n > c->max_batch) return fail(c, SIGOPS_E_ARG, "n exceeds max_batch");
  if (n == 0) return SIGOPS_OK;
  if (!d_schemes || !d_sigs || !d_aux || !d_msgs || !d_msg_off || !d_out || !d_status)
    return fail(c, SIGOPS_E_ARG, "NULL buffer");
  if (((uintptr_t)d_msgs & 3u) != 0) return fail(c, SIGOPS_E_ARG, "d_msgs must be 4-byte aligned");
  hipStream_t s = (hipStream_t)stream;
  // the call's buffers and n go to device memory first (stream-ordered, by value at enqueue time),
  // so one captured graph serves every call and back-to-back enqueues do not race
  HIPCHK(c, sigops::launch_set_batch_io(c->d_io, {d_schemes, d_sigs, d_aux, d_msgs, d_msg_off, d_out, d_status,
                                                  (uint32_t)n}, s));
  if (c->fault & SIGOPS_FAULT_SKIP_COMPUTE) return SIGOPS_OK;
  sigops::BatchLaunch L{};
  L.io = c->d_io;
  L.work = c->d_work;
  L.order = c->d_order;
  L.pend_list = c->d_pend + 1;
  L.rank = c->d_rank;
  L.gk1 = c->d_gtab_k1;
  L.gr1 = c->d_gtab_r1;
  L.gb = c->d_btab;
  L.lane_tab = c->d_lane_tab;
  L.slots = c->slots;
  L.grid = c->grid;
  L.aux_grid = std::min<uint32_t>(c->grid, (uint32_t)((n + 255) / 256));
  L.tile_rounds = c->tile_rounds;
  L.waves = c->waves;
  L.g_window = c->g_window;
  L.stage_ev = nullptr;
  if (c->stage_timing) {
    L.stage_ev = c->stage_ev;
    HIPCHK(c, sigops::launch_batch(L, s));
    c->stage_recorded = true;
    return SIGOPS_OK;
  }
  if (!c->use_graph) {
    HIPCHK(c, sigops::launch_batch(L, s));
    return SIGOPS_OK;
  }
  if (!c->graph) {
    L.aux_grid = c->grid;  // the graph does not depend on n
    hipGraph_t g;
    HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    hipError_t le = sigops::launch_batch(L, c->stream);
    hipError_t ce = hipStreamEndCapture(c->stream, &g);
    if (le != hipSuccess) return hip_fail(c, le, "launch_batch under capture");
    if (ce != hipSuccess) return hip_fail(c, ce, "hipStreamEndCapture");
    hipError_t ie = hipGraphInstantiate(&c->graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ie != hipSuccess) {
      c->graph = nullptr;
      return hip_fail(c, ie, "hipGraphInstantiate");
    }
  }
  HIPCHK(c, hipGraphLaunch(c->graph, s));
  return SIGOPS_OK;
}

int sigops_secp256k1_recover_address(sigops_ctx* c, const uint8_t* sigs, const uint8_t* digests, size_t n,
                                     uint8_t* out_addr, uint8_t* status) {
  if (!out_addr && n) return fail(c, SIGOPS_E_ARG, "NULL buffer");
  return secp_recover_host(c, 0, sigs, digests, n, nullptr, status, out_addr);
}

int sigops_secp256r1_recover_address(sigops_ctx* c, const uint8_t* sigs, const uint8_t* digests, size_t n,
                                     uint8_t* out_addr, uint8_t* status) {
  if (!out_addr && n) return fail(c, SIGOPS_E_ARG, "NULL buffer");
  return secp_recover_host(c, 1, sigs, digests, n, nullptr, status, out_addr);
}

int sigops_message_new_dev(sigops_ctx* c, const void* d_data, const uint64_t* d_off, size_t n, void* d_digests,
                           void* stream) {
  if (!c) return SIGOPS_E_ARG;
  if (n == 0) return SIGOPS_OK;
  if (n > c->max_batch) return fail(c, SIGOPS_E_ARG, "n exceeds max_batch");
  if (!d_data || !d_off || !d_digests) return fail(c, SIGOPS_E_ARG, "NULL buffer");
  if (((uintptr_t)d_data & 3u) != 0) return fail(c, SIGOPS_E_ARG, "d_data must be 4-byte aligned");
  HIPCHK(c, sigops::launch_message_new(d_data, d_off, (uint32_t)n, d_digests, c->d_hq, c->d_horder, 2 * c->grid,
                                       (hipStream_t)stream));
  return SIGOPS_OK;
}

int sigops_message_new_noalloc(sigops_ctx* c, const void* d_data, const uint64_t* d_off, size_t n, void* d_digests,
                               void* stream) {
  if (!c) return SIGOPS_E_ARG;
  if (n == 0) return SIGOPS_OK;
  if (n > SIGOPS_NOALLOC_MAX_MESSAGES)
    return fail(c, SIGOPS_E_ARG, "n exceeds SIGOPS_NOALLOC_MAX_MESSAGES (" + std::to_string(SIGOPS_NOALLOC_MAX_MESSAGES) +
                                     ")");
  if (!d_data || !d_off || !d_digests) return fail(c, SIGOPS_E_ARG, "NULL buffer");
  if (((uintptr_t)d_data & 3u) != 0) return fail(c, SIGOPS_E_ARG, "d_data must be 4-byte aligned");
  if (!c->noalloc_scratch) return fail(c, SIGOPS_E_DEVICE, "no fixed scratch on this device");
  uint32_t* work = (uint32_t*)c->noalloc_scratch;
  uint32_t* order = (uint32_t*)((uint8_t*)c->noalloc_scratch + NOALLOC_WORK_BYTES);
  // the launch clears the work words with a kernel and writes the order before reading it, so
  // what an earlier call left in the array does not matter
  HIPCHK(c, sigops::launch_message_new(d_data, d_off, (uint32_t)n, d_digests, work, order, 2 * c->grid,
                                       (hipStream_t)stream));
  return SIGOPS_OK;
}

int sigops_message_new(sigops_ctx* c, const uint8_t* data, const uint64_t* off, size_t n, uint8_t* digests) {
  if (!c) return SIGOPS_E_ARG;
  c->err.clear();
  if (n == 0) return SIGOPS_OK;
  if (!off || !digests) return fail(c, SIGOPS_E_ARG, "NULL buffer");
  if (n > c->max_batch) return fail(c, SIGOPS_E_ARG, "n exceeds max_batch");
  if (!offsets_ok(c, off, n)) return fail(c, SIGOPS_E_ARG, "off must be non-decreasing with lengths < 2^32");
  const uint64_t base = off[0];
  const uint64_t total = off[n] - base;
  if (total && !data) return fail(c, SIGOPS_E_ARG, "NULL data");
  int rc;
  if ((rc = ensure(c, c->msg, total + 16)) || (rc = ensure(c, c->off, (n + 1) * 8)) || (rc = ensure(c, c->aux, n * 32)))
    return rc;
  hipStream_t s = c->stream;
  if ((rc = put(c, {{c->msg.p, total ? data + base : nullptr, total}, {c->off.p, off, (n + 1) * 8, base}}))) return rc;
  k_fill_bytes<<<1, 64, 0, s>>>((uint8_t*)c->msg.p + total, 0, 16);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, sigops::launch_message_new(c->msg.p, (const uint64_t*)c->off.p, (uint32_t)n, c->aux.p, c->d_hq, c->d_horder,
                                       2 * c->grid, s));
  return get(c, {{digests, c->aux.p, n * 32}});
}

int sigops_address_from_pubkeys_dev(sigops_ctx* c, const void* d_pubs, const void* d_status, size_t n, void* d_addr,
                                    void* stream) {
  if (!c) return SIGOPS_E_ARG;
  if (n == 0) return SIGOPS_OK;
  if (n > 0xFFFFFFFFull) return fail(c, SIGOPS_E_ARG, "n too large");
  if (!d_pubs || !d_addr) return fail(c, SIGOPS_E_ARG, "NULL buffer");
  HIPCHK(c, sigops::launch_fuel_address(d_pubs, d_status, (uint32_t)n, d_addr, c->grid, (hipStream_t)stream));
  return SIGOPS_OK;
}

int sigops_set_fault(sigops_ctx* c, uint32_t flags) {
  if (!c) return SIGOPS_E_ARG;
  if (flags & ~(uint32_t)SIGOPS_FAULT_SKIP_COMPUTE) return fail(c, SIGOPS_E_ARG, "unknown fault flag");
  c->fault = flags;
  return SIGOPS_OK;
}

int sigops_set_stage_timing(sigops_ctx* c, int on) {
  if (!c) return SIGOPS_E_ARG;
  HIPCHK(c, hipSetDevice(c->device));
  if (on)
    for (hipEvent_t& e : c->stage_ev)
      if (!e) HIPCHK(c, hipEventCreate(&e));
  c->stage_timing = on != 0;
  c->stage_recorded = false;
  return SIGOPS_OK;
}

int sigops_batch_stage_ms(sigops_ctx* c, float* ms) {
  if (!c || !ms) return SIGOPS_E_ARG;
  if (!c->stage_recorded) return fail(c, SIGOPS_E_ARG, "no timed batch since stage timing was switched on");
  HIPCHK(c, hipEventSynchronize(c->stage_ev[SIGOPS_STAGES]));
  for (int i = 0; i < SIGOPS_STAGES; i++) HIPCHK(c, hipEventElapsedTime(&ms[i], c->stage_ev[i], c->stage_ev[i + 1]));
  return SIGOPS_OK;
}

}  // extern "C"

namespace {

// regions of the streamed layout in the pinned block, each 256-byte aligned
struct StreamRegions {
  size_t desc, in, mrel, mlen, out, st, done, msgs, total;
};

StreamRegions stream_regions(const sigops_host::StreamGeom& g) {
  auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
  StreamRegions r;
  size_t at = 0;
  r.desc = at;
  at = al(at + g.slots * 16);
  r.in = at;
  at = al(at + g.slots * sigops_host::ST_IN_CHUNKS * 64 * 16);
  r.mrel = at;
  at = al(at + g.slots * 64 * 4);
  r.mlen = at;
  at = al(at + g.slots * 64 * 4);
  r.out = at;
  at = al(at + g.slots * sigops_host::ST_OUT_CHUNKS * 64 * 16);
  r.st = at;
  at = al(at + g.slots * 64);
  r.done = at;
  at = al(at + g.slots * 8);
  r.msgs = at;
  at = al(at + g.msg_base[g.nchunks] + 16);
  r.total = at;
  return r;
}

// device address of a pinned, device-mapped host address
void* dev_addr(void* host) {
  void* d = nullptr;
  return hipHostGetDevicePointer(&d, host, 0) == hipSuccess ? d : nullptr;
}

// Streamed mixed batch (caller buffers validated). Returns SIGOPS_OK with every result written;
// 1 when the launch could not take every chunk (the caller falls back to the copy-then-launch
// path); or an error code.
int batch_streamed(sigops_ctx* c, const uint8_t* schemes, const uint8_t* sigs, const uint8_t* aux,
                   const uint8_t* msgs, const uint64_t* msg_off, size_t n, uint8_t* out, uint8_t* status) {
  using namespace sigops_host;
  typedef std::chrono::steady_clock clk;
  const auto T0 = clk::now();
  auto since = [&](clk::time_point t) { return std::chrono::duration<float, std::milli>(t - T0).count(); };
  std::atomic<int64_t> t_packed{0};  // when the last chunk was published (ns after T0)
  // chunks: about 1/16 of the batch, 1024..16384 items, cut from the batch sorted by bucket
  const size_t chunk = std::min<size_t>(16384, std::max<size_t>(1024, ((n / 16) + 63) & ~(size_t)63));
  StreamGeom g = stream_geometry(n, chunk);
  Pool& pool = host_pool(c);
  static_assert(ST_BUCKETS == sigops::NBUCKETS && ST_ED0 == sigops::BUCKET_ED0, "one bucket numbering");
  c->s_bkt.resize(n);
  c->s_sorted.resize(n);
  const uint32_t* sorted = c->s_sorted.data();
  const uint8_t* bkt = c->s_bkt.data();
  stream_sort([&](const std::function<void(unsigned)>& fn) { pool.run(fn); }, pool.size(), n, schemes, msg_off,
              c->bucket_order, c->s_bkt.data(), c->s_sorted.data());
  {
    std::atomic<size_t> next_k{0};
    pool.run([&](unsigned) {
      for (size_t k; (k = next_k.fetch_add(1)) < g.nchunks;) stream_message_bytes(g, k, sorted, msg_off);
    });
  }
  for (size_t k = 0; k < g.nchunks; k++)
    if (g.msg_base[k + 1] >= (1ull << 32)) return 1;  // descriptors hold 32-bit tile sizes
  stream_message_layout(g);
  const StreamRegions R = stream_regions(g);
  hipError_t e;
  if (!c->s_ctl) {
    if ((e = hipHostMalloc((void**)&c->s_ctl, 256, hipHostMallocMapped | hipHostMallocCoherent)) != hipSuccess)
      return xfer_fail(c, e, "streamed batch control");
    memset(c->s_ctl, 0, 256);
    if ((e = hipMalloc((void**)&c->d_sctr, sigops::STREAM_CTR_BYTES)) != hipSuccess) return hip_fail(c, e, "hipMalloc");
    if ((e = hipEventCreateWithFlags(&c->s_ev, hipEventDisableTiming)) != hipSuccess) return hip_fail(c, e, "event");
    for (hipEvent_t& te : c->s_tev)
      if ((e = hipEventCreate(&te)) != hipSuccess) return hip_fail(c, e, "event");
  }
  if (c->s_cap < R.total) {
    if (c->s_pin) (void)hipHostFree(c->s_pin);
    c->s_pin = nullptr;
    c->s_cap = 0;
    const size_t want = (R.total + (2u << 20) - 1) & ~(size_t)((2u << 20) - 1);
    if ((e = hipHostMalloc((void**)&c->s_pin, want, hipHostMallocMapped | hipHostMallocCoherent)) != hipSuccess) {
      c->s_pin = nullptr;
      return xfer_fail(c, e, "streamed batch staging");
    }
    memset(c->s_pin, 0, want);  // the tile done words start below any batch tag
    c->s_cap = want;
  }
  int rc;
  if ((rc = ensure(c, c->msg, g.msg_base[g.nchunks] + 16))) return rc;  // HBM copies of the tiles' messages
  c->s_perm.resize(g.slots * 64);
  uint8_t* P = c->s_pin;
  StreamBufs B{(StreamDesc*)(P + R.desc), P + R.in, (uint32_t*)(P + R.mrel), (uint32_t*)(P + R.mlen), P + R.msgs,
               P + R.out, P + R.st, c->s_perm.data()};
  uint64_t* done = (uint64_t*)(P + R.done);
  uint64_t* ready = (uint64_t*)c->s_ctl;
  uint32_t* stop = (uint32_t*)(c->s_ctl + 8);
  uint32_t* err = (uint32_t*)(c->s_ctl + 12);
  const uint32_t seq = ++c->s_seq ? c->s_seq : ++c->s_seq;  // never 0, the cleared done words
  __atomic_store_n(ready, (uint64_t)seq << 32, __ATOMIC_RELEASE);
  __atomic_store_n(stop, 0u, __ATOMIC_RELEASE);
  __atomic_store_n(err, 0u, __ATOMIC_RELEASE);
  std::unique_ptr<std::atomic<uint32_t>[]> packed(new std::atomic<uint32_t>[g.nchunks]());
  std::vector<uint32_t> tiles(g.nchunks, 0);
  std::mutex pub_mu;
  size_t published = 0;
  std::atomic<bool> lost{false};
  auto pack = [&](size_t k) {
    tiles[k] = stream_pack_chunk(g, k, sorted, bkt, sigs, aux, msgs, msg_off, B);
    packed[k].store(1, std::memory_order_release);
    std::lock_guard<std::mutex> lk(pub_mu);
    while (published < g.nchunks && packed[published].load(std::memory_order_acquire)) published++;
    // the last chunk closes the count: the launch's wavefronts end once every tile is taken
    __atomic_store_n(ready,
                     ((uint64_t)seq << 32) | (published * g.tpc) | (published == g.nchunks ? sigops::STREAM_CLOSED : 0),
                     __ATOMIC_RELEASE);
    if (published == g.nchunks)
      t_packed.store(std::chrono::duration_cast<std::chrono::nanoseconds>(clk::now() - T0).count());
  };
  // chunks packed before the launch: enough tiles for half the resident wavefronts, so that the
  // waves that start first find work instead of polling host memory
  const size_t waves = (size_t)c->grid * 4;
  const size_t head = std::min(g.nchunks, (waves / 2 + chunk / 64 - 1) / (chunk / 64) + 1);
  std::atomic<size_t> next{0};
  pool.run([&](unsigned) {
    for (size_t k; (k = next.fetch_add(1)) < head;) pack(k);
  });
  hipStream_t s = c->stream;
  sigops::StreamLaunch L{};
  // device addresses of the two pinned blocks; the regions sit at the same offsets
  uint8_t* DP = (uint8_t*)dev_addr(P);
  uint8_t* DC = (uint8_t*)dev_addr(c->s_ctl);
  if (!DP || !DC) return fail(c, SIGOPS_E_DEVICE, "hipHostGetDevicePointer failed for the streamed batch");
  L.desc = DP + R.desc;
  L.in = DP + R.in;
  L.mrel = (const uint32_t*)(DP + R.mrel);
  L.mlen = (const uint32_t*)(DP + R.mlen);
  L.hmsgs = DP + R.msgs;
  L.dmsgs = c->msg.p;
  L.msg_cap = g.msg_base[g.nchunks] + 16;
  L.out = DP + R.out;
  L.st = DP + R.st;
  L.tile_done = (uint64_t*)(DP + R.done);
  L.ready = (const uint64_t*)DC;
  L.stop = (const uint32_t*)(DC + 8);
  L.err = (uint32_t*)(DC + 12);
  L.ctr = c->d_sctr;
  L.pend_list = c->d_pend + 1;
  L.pend_cap = (uint32_t)n;
  L.gk1 = c->d_gtab_k1;
  L.gr1 = c->d_gtab_r1;
  L.gb = c->d_btab;
  L.lane_tab = c->d_lane_tab;
  L.slots = sigops::STREAM_CLOSED - 1;
  L.base = 0;
  L.mask = ~0u;
  L.nq = 1;
  L.seq = seq;
  L.lane_slots = c->slots;
  L.grid = std::min<uint32_t>(c->grid, (uint32_t)((g.slots + 3) / 4));
  L.g_window = c->g_window;
  sigops::stream_poll_ticks(c->device, L);
  HIPCHK(c, hipEventRecord(c->s_tev[0], s));
  if (!(c->fault & SIGOPS_FAULT_SKIP_COMPUTE)) HIPCHK(c, sigops::launch_stream(L, s));
  HIPCHK(c, hipEventRecord(c->s_tev[1], s));
  HIPCHK(c, hipEventRecord(c->s_ev, s));
  const auto T1 = clk::now();
  // the rest of the chunks, then the results of every chunk as its tiles finish
  std::atomic<size_t> next_unpack{0};
  next.store(head);
  pool.run([&](unsigned) {
    for (size_t k; (k = next.fetch_add(1)) < g.nchunks;) pack(k);
    for (size_t k; (k = next_unpack.fetch_add(1)) < g.nchunks;) {
      while (!packed[k].load(std::memory_order_acquire)) std::this_thread::yield();
      const size_t t0 = k * g.tpc;
      bool ok = true;
      for (uint32_t j = 0; j < tiles[k] && ok; j++) {
        const uint64_t tag = ((uint64_t)seq << 32) | (t0 + j);
        for (unsigned spin = 1; __atomic_load_n(&done[t0 + j], __ATOMIC_ACQUIRE) != tag; spin++) {
          if (spin % 256 == 0) {
            // the launch has ended: a tile still not done will not be
            if (hipEventQuery(c->s_ev) == hipSuccess && __atomic_load_n(&done[t0 + j], __ATOMIC_ACQUIRE) != tag) {
              ok = false;
              break;
            }
            std::this_thread::yield();
          }
        }
      }
      if (!ok) {
        lost.store(true);
        __atomic_store_n(stop, 1u, __ATOMIC_RELEASE);
        continue;
      }
      stream_unpack_tiles(t0, t0 + tiles[k], B, out, status);
    }
  });
  const auto T3 = clk::now();
  HIPCHK(c, hipEventSynchronize(c->s_ev));
  const auto T4 = clk::now();
  c->s_ms[0] = since(T1);
  c->s_ms[1] = (float)(t_packed.load() / 1e6);
  c->s_ms[2] = since(T3);
  c->s_ms[3] = since(T4);
  (void)hipEventElapsedTime(&c->s_ms[5], c->s_tev[0], c->s_tev[1]);
  if (c->fault & SIGOPS_FAULT_SKIP_COMPUTE) return fail(c, SIGOPS_E_CANARY, "a batch result was not written");
  if (lost.load() || __atomic_load_n(err, __ATOMIC_ACQUIRE)) return 1;
  // items the careful pass redid
  uint32_t ctr[sigops::STREAM_CTR_PENDING + 1];
  HIPCHK(c, hipMemcpy(ctr, c->d_sctr, sizeof ctr, hipMemcpyDeviceToHost));
  if (ctr[sigops::STREAM_CTR_PENDING]) {
    std::vector<uint32_t> pend(std::min<size_t>(ctr[sigops::STREAM_CTR_PENDING], n));
    HIPCHK(c, hipMemcpy(pend.data(), c->d_pend + 1, pend.size() * sizeof(uint32_t), hipMemcpyDeviceToHost));
    for (uint32_t p : pend) stream_unpack_lane(p, B, out, status);
  }
  c->s_ms[4] = since(clk::now());
  return SIGOPS_OK;
}

}  // namespace

int sigops_ring_resources(sigops_ctx* c, sigops::RingResources* r) {
  if (!c || !r) return SIGOPS_E_ARG;
  r->gk1 = c->d_gtab_k1;
  r->gr1 = c->d_gtab_r1;
  r->gb = c->d_btab;
  r->lane_tab = c->d_lane_tab;
  r->lane_slots = c->slots;
  r->grid = c->grid;
  r->g_window = c->g_window;
  r->bucket_order = c->bucket_order;
  r->stream = c->stream;
  return SIGOPS_OK;
}

extern "C" {

int sigops_batch_host_ms(sigops_ctx* c, float* ms) {
  if (!c || !ms) return SIGOPS_E_ARG;
  for (int i = 0; i < SIGOPS_HOST_PHASES; i++) ms[i] = c->s_ms[i];
  return SIGOPS_OK;
}

int sigops_set_streaming(sigops_ctx* c, int on) {
  if (!c) return SIGOPS_E_ARG;
  c->streaming = on != 0;
  return SIGOPS_OK;
}

int sigops_batch(sigops_ctx* c, const uint8_t* schemes, const uint8_t* sigs, const uint8_t* aux, const uint8_t* msgs,
                 const uint64_t* msg_off, size_t n, uint8_t* out, uint8_t* status) {
  if (!c) return SIGOPS_E_ARG;
  c->err.clear();
  if (n == 0) return SIGOPS_OK;
  if (!schemes || !sigs || !aux || !msg_off || !out) return fail(c, SIGOPS_E_ARG, "NULL buffer");
  if (n > c->max_batch) return fail(c, SIGOPS_E_ARG, "n exceeds max_batch");
  if (!all_of(c, n, [&](size_t i) { return schemes[i] <= SIGOPS_SCHEME_ED25519; }))
    return fail(c, SIGOPS_E_ARG, "unknown scheme byte");
  if (!offsets_ok(c, msg_off, n))
    return fail(c, SIGOPS_E_ARG, "msg_off must be non-decreasing with lengths < 2^32");
  const uint64_t base = msg_off[0];
  const uint64_t total = msg_off[n] - base;
  if (total && !msgs) return fail(c, SIGOPS_E_ARG, "NULL msgs");
  uint8_t* st = status;
  if (!st) {
    c->host_status.resize(n);
    st = c->host_status.data();
  }
  int rc;
  if (c->streaming && n >= 4096) {
    rc = batch_streamed(c, schemes, sigs, aux, msgs, msg_off, n, out, st);
    if (rc < 0) return rc;
    if (rc == SIGOPS_OK) {
      if (!all_of(c, n, [&](size_t i) {
            return st[i] <= (schemes[i] == SIGOPS_SCHEME_ED25519 ? SIGOPS_ED_MISMATCH : SIGOPS_ST_INFINITY);
          }))
        return fail(c, SIGOPS_E_CANARY, "a batch result was not written");
      return SIGOPS_OK;
    }
    // rc == 1: the streamed launch did not take every chunk; redo the batch the copy-then-launch way
  }
  if ((rc = ensure(c, c->in, n * 64)) || (rc = ensure(c, c->aux, n * 32)) || (rc = ensure(c, c->msg, total + 16)) ||
      (rc = ensure(c, c->off, (n + 1) * 8)) || (rc = ensure(c, c->out, n * 64)) || (rc = ensure(c, c->st, n)) ||
      (rc = ensure(c, c->sch, n)))
    return rc;
  hipStream_t s = c->stream;
  if ((rc = put(c, {{c->sch.p, schemes, n},
                    {c->in.p, sigs, n * 64},
                    {c->aux.p, aux, n * 32},
                    {c->msg.p, total ? msgs + base : nullptr, total},
                    {c->off.p, msg_off, (n + 1) * 8, base}})))
    return rc;
  k_fill_bytes<<<1, 64, 0, s>>>((uint8_t*)c->msg.p + total, 0, 16);
  k_fill_bytes<<<64, 256, 0, s>>>((uint8_t*)c->st.p, CANARY, n);
  HIPCHK(c, hipGetLastError());
  if ((rc = sigops_batch_dev(c, c->sch.p, c->in.p, c->aux.p, c->msg.p, (const uint64_t*)c->off.p, n, c->out.p,
                             c->st.p, s)))
    return rc;
  if ((rc = get(c, {{out, c->out.p, n * 64}, {st, c->st.p, n}}))) return rc;
  if (!all_of(c, n, [&](size_t i) {
        return st[i] <= (schemes[i] == SIGOPS_SCHEME_ED25519 ? SIGOPS_ED_MISMATCH : SIGOPS_ST_INFINITY);
      }))
    return fail(c, SIGOPS_E_CANARY, "a batch result was not written");
  return SIGOPS_OK;
}

}  // extern "C"
