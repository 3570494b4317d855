// serve.cpp -- request servers over the mixed-batch engine.
//
// The reference has no serving layer: every call builds a device, compiles, runs one batch and
// destroys the device (src/gpu.rs:5-35,167). Here requests of any size arrive while others run.
// submit() copies a request into a host queue and returns a ticket at once; requests leave the
// queue in arrival order while the bytes of the requests in flight stay within the HBM budget.
// Two ways to run them (sigops_server_config.mode):
//   * slot batches: the dispatcher packs queued requests FIFO into the next free slot (one
//     sigops_ctx with its own stream, scratch, device buffers and captured graph), as long as the
//     batch stays within max_batch items; it launches as soon as a slot is free, so arrivals during
//     a launch go out with the next one and consecutive batches overlap on the device. The
//     completion thread polls the slots' events, scatters results back to their requests and
//     frees the slot.
//   * the device-drained ring: one persistent launch (an epoch) drains a ring of tile positions
//     in pinned, device-mapped memory. The dispatcher sorts the items of admitted requests into
//     per-bucket lists and packs them into ring positions as tiles of 64 items of one bucket, so
//     a tile mixes items of several requests; full tiles go out at once, partial ones when some of
//     the launch's wavefronts would otherwise idle. Each publication is one store of the tile
//     count; the kernel's wavefronts claim tiles from a device counter, so a request that arrives
//     while the GPU is busy joins the running launch at the next free wavefront instead of the
//     next batch. The completion thread polls the tiles' done words, copies results to their
//     requests, and frees positions in order. An epoch ends (its count is closed) when the ring
//     has been idle for idle_close_us, or when a tile holds a lane that needs the careful pass,
//     which runs after the epoch's persistent kernel; the next epoch is launched behind it.
// Either way every item is processed by the same per-item code as a standalone batch, so a
// request's results do not depend on what it was batched with.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdint>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/sigops.h"
#include "host_xfer.hpp"
#include "kernels.hpp"
#include "stream_host.hpp"

namespace {

uint64_t now_ns() {
  return (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

struct Request {
  int64_t ticket = 0;
  size_t n = 0;
  std::vector<uint8_t> schemes, sigs, aux, msgs, out, status;
  std::vector<uint64_t> off;  // n + 1, starting at 0
  uint64_t t_submit = 0, t_done = 0;
  uint64_t bytes = 0;
  std::atomic<size_t> remaining{0};  // ring: items whose results are not in yet
  // ring: the items in bucket order (order[bstart[b] .. bstart[b + 1]) are bucket b's), computed
  // by the submitting thread
  std::vector<uint32_t> order;
  uint32_t bstart[sigops::NBUCKETS + 1] = {};
  bool done = false;
  std::atomic<int> err{0};
};

uint64_t request_bytes(size_t n, uint64_t msg_bytes) {
  // device footprint: scheme + sig + aux + offsets + out + status + challenge, plus messages
  return (uint64_t)n * (1 + 64 + 32 + 8 + 64 + 1 + 32) + msg_bytes + 16;
}

template <class T>
struct PinnedBuf {
  T* p = nullptr;
  size_t cap = 0;
  bool ensure(size_t count) {
    if (count <= cap) return true;
    const size_t want = std::max<size_t>(std::max<size_t>(count, 2 * cap), 1024);
    if (p) (void)hipHostFree(p);
    p = nullptr;
    cap = 0;
    if (hipHostMalloc((void**)&p, want * sizeof(T), hipHostMallocDefault) != hipSuccess) return false;
    cap = want;
    return true;
  }
  ~PinnedBuf() {
    if (p) (void)hipHostFree(p);
  }
};

template <class T>
struct DevArr {
  T* p = nullptr;
  size_t cap = 0;
  bool ensure(size_t count) {
    if (count <= cap) return true;
    const size_t want = std::max<size_t>(std::max<size_t>(count, 2 * cap), 1024);
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    if (hipMalloc((void**)&p, want * sizeof(T)) != hipSuccess) return false;
    cap = want;
    return true;
  }
  ~DevArr() {
    if (p) (void)hipFree(p);
  }
};

struct Slot {
  sigops_ctx* ctx = nullptr;
  hipStream_t stream = nullptr;
  hipEvent_t done_ev = nullptr;
  PinnedBuf<uint8_t> h_sch, h_sig, h_aux, h_msg, h_out, h_st;
  PinnedBuf<uint64_t> h_off;
  DevArr<uint8_t> d_sch, d_sig, d_aux, d_msg, d_out, d_st;
  DevArr<uint64_t> d_off;
  std::vector<std::shared_ptr<Request>> reqs;
  uint64_t bytes = 0;
  size_t items = 0;
};

struct RingEngine;

}  // namespace

struct sigops_server {
  sigops_server_config cfg{};
  std::vector<std::unique_ptr<Slot>> slots;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<std::shared_ptr<Request>> queue;
  std::unordered_map<int64_t, std::shared_ptr<Request>> tickets;
  std::vector<Slot*> free_slots;
  std::vector<Slot*> inflight;
  uint64_t budget_used = 0;
  int64_t next_ticket = 0;
  bool stopping = false;
  bool dispatcher_done = false;
  std::thread dispatcher, completer;
  std::unique_ptr<sigops_host::Pool> pack_pool;    // used by the dispatcher only
  std::unique_ptr<sigops_host::Pool> unpack_pool;  // used by the completion thread only
  std::atomic<uint64_t> n_batches{0}, n_items{0};
  int fatal = 0;
  std::unique_ptr<RingEngine> ring;  // mode SIGOPS_SERVE_RING
  bool ring_wake = false;            // the ring made progress: the dispatcher looks again

  void dispatch_loop();
  void complete_loop();
  int launch(Slot* s);
};

int sigops_server::launch(Slot* s) {
  // pack the slot's requests into pinned staging, then enqueue H2D, the batch and D2H on its stream
  size_t n = 0, msg = 0;
  for (auto& r : s->reqs) {
    n += r->n;
    msg += r->msgs.size();
  }
  if (!s->h_sch.ensure(n) || !s->h_sig.ensure(n * 64) || !s->h_aux.ensure(n * 32) || !s->h_msg.ensure(msg + 16) ||
      !s->h_off.ensure(n + 1) || !s->h_out.ensure(n * 64) || !s->h_st.ensure(n))
    return SIGOPS_E_NOMEM;
  if (!s->d_sch.ensure(n) || !s->d_sig.ensure(n * 64) || !s->d_aux.ensure(n * 32) || !s->d_msg.ensure(msg + 16) ||
      !s->d_off.ensure(n + 1) || !s->d_out.ensure(n * 64) || !s->d_st.ensure(n))
    return SIGOPS_E_NOMEM;
  // item and message positions of each request, then the copies, spread over the pool by items
  const size_t nr = s->reqs.size();
  std::vector<size_t> ri(nr + 1), rm(nr + 1);
  ri[0] = rm[0] = 0;
  for (size_t k = 0; k < nr; k++) {
    ri[k + 1] = ri[k] + s->reqs[k]->n;
    rm[k + 1] = rm[k] + s->reqs[k]->msgs.size();
  }
  auto pack = [&](size_t ka, size_t kb) {
    for (size_t k = ka; k < kb; k++) {
      const Request& r = *s->reqs[k];
      const size_t i0 = ri[k], m0 = rm[k];
      memcpy(s->h_sch.p + i0, r.schemes.data(), r.n);
      memcpy(s->h_sig.p + i0 * 64, r.sigs.data(), r.n * 64);
      memcpy(s->h_aux.p + i0 * 32, r.aux.data(), r.n * 32);
      if (!r.msgs.empty()) memcpy(s->h_msg.p + m0, r.msgs.data(), r.msgs.size());
      for (size_t j = 0; j < r.n; j++) s->h_off.p[i0 + j] = m0 + r.off[j];
    }
  };
  const unsigned nt = pack_pool->size();
  if (nt > 1 && n >= 8192 && nr > 1) {
    // thread t packs the requests whose first item lies in [t n / nt, (t + 1) n / nt)
    pack_pool->run([&](unsigned t) {
      const size_t ka = std::lower_bound(ri.begin(), ri.begin() + nr, n * t / nt) - ri.begin();
      const size_t kb = std::lower_bound(ri.begin(), ri.begin() + nr, n * (t + 1) / nt) - ri.begin();
      pack(ka, kb);
    });
  } else {
    pack(0, nr);
  }
  const size_t m0 = rm[nr];
  s->h_off.p[n] = m0;
  memset(s->h_msg.p + m0, 0, 16);
  s->items = n;
  hipStream_t st = s->stream;
  if (hipMemcpyAsync(s->d_sch.p, s->h_sch.p, n, hipMemcpyHostToDevice, st) != hipSuccess ||
      hipMemcpyAsync(s->d_sig.p, s->h_sig.p, n * 64, hipMemcpyHostToDevice, st) != hipSuccess ||
      hipMemcpyAsync(s->d_aux.p, s->h_aux.p, n * 32, hipMemcpyHostToDevice, st) != hipSuccess ||
      hipMemcpyAsync(s->d_msg.p, s->h_msg.p, m0 + 16, hipMemcpyHostToDevice, st) != hipSuccess ||
      hipMemcpyAsync(s->d_off.p, s->h_off.p, (n + 1) * 8, hipMemcpyHostToDevice, st) != hipSuccess)
    return SIGOPS_E_DEVICE;
  int rc = sigops_batch_dev(s->ctx, s->d_sch.p, s->d_sig.p, s->d_aux.p, s->d_msg.p, s->d_off.p, n, s->d_out.p,
                            s->d_st.p, st);
  if (rc) return rc;
  if (hipMemcpyAsync(s->h_out.p, s->d_out.p, n * 64, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipMemcpyAsync(s->h_st.p, s->d_st.p, n, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipEventRecord(s->done_ev, st) != hipSuccess)
    return SIGOPS_E_DEVICE;
  n_batches++;
  n_items += n;
  return SIGOPS_OK;
}

void sigops_server::dispatch_loop() {
  (void)hipSetDevice(cfg.device);
  std::unique_lock<std::mutex> lk(mu);
  for (;;) {
    cv_work.wait(lk, [&] {
      if (stopping && queue.empty()) return true;
      return !queue.empty() && !free_slots.empty() && budget_used + queue.front()->bytes <= cfg.hbm_budget;
    });
    if (queue.empty()) {  // stopping with nothing left to launch
      dispatcher_done = true;
      cv_done.notify_all();
      return;
    }
    Slot* s = free_slots.back();
    free_slots.pop_back();
    s->reqs.clear();
    s->bytes = 0;
    size_t items = 0;
    while (!queue.empty()) {
      auto& r = queue.front();
      if (!s->reqs.empty() && (items + r->n > cfg.max_batch || budget_used + r->bytes > cfg.hbm_budget)) break;
      items += r->n;
      budget_used += r->bytes;
      s->bytes += r->bytes;
      s->reqs.push_back(r);
      queue.pop_front();
    }
    lk.unlock();
    const int rc = launch(s);
    lk.lock();
    if (rc) {
      // the requests fail with the error; the slot returns to the pool
      const uint64_t t = now_ns();
      for (auto& r : s->reqs) {
        r->err = rc;
        r->done = true;
        r->t_done = t;
      }
      budget_used -= s->bytes;
      free_slots.push_back(s);
      cv_done.notify_all();
      continue;
    }
    inflight.push_back(s);
    cv_done.notify_all();  // wakes the completion thread
  }
}

void sigops_server::complete_loop() {
  (void)hipSetDevice(cfg.device);
  std::unique_lock<std::mutex> lk(mu);
  for (;;) {
    cv_done.wait_for(lk, std::chrono::microseconds(200), [&] { return dispatcher_done || !inflight.empty(); });
    if (inflight.empty()) {
      if (dispatcher_done) return;
      continue;
    }
    // poll every in-flight slot; batches may complete out of launch order
    std::vector<Slot*> ready;
    for (size_t k = 0; k < inflight.size();) {
      if (hipEventQuery(inflight[k]->done_ev) != hipErrorNotReady) {
        ready.push_back(inflight[k]);
        inflight.erase(inflight.begin() + (long)k);
      } else {
        k++;
      }
    }
    if (ready.empty()) {
      lk.unlock();
      std::this_thread::sleep_for(std::chrono::microseconds(20));
      lk.lock();
      continue;
    }
    lk.unlock();
    for (Slot* s : ready) {
      const hipError_t e = hipEventSynchronize(s->done_ev);
      const uint64_t t = now_ns();
      // each request's results out of the slot's staging, spread over the pool by items
      const size_t nr = s->reqs.size();
      std::vector<size_t> ri(nr + 1);
      ri[0] = 0;
      for (size_t k = 0; k < nr; k++) ri[k + 1] = ri[k] + s->reqs[k]->n;
      auto scatter = [&](size_t ka, size_t kb) {
        for (size_t k = ka; k < kb; k++) {
          Request& r = *s->reqs[k];
          const size_t i0 = ri[k];
          if (e != hipSuccess) {
            r.err = SIGOPS_E_DEVICE;
          } else {
            r.out.assign(s->h_out.p + i0 * 64, s->h_out.p + (i0 + r.n) * 64);
            r.status.assign(s->h_st.p + i0, s->h_st.p + i0 + r.n);
            for (size_t j = 0; j < r.n; j++) {
              const uint8_t lim = r.schemes[j] == SIGOPS_SCHEME_ED25519 ? SIGOPS_ED_MISMATCH : SIGOPS_ST_INFINITY;
              if (r.status[j] > lim) r.err = SIGOPS_E_CANARY;
            }
          }
          r.t_done = t;
        }
      };
      const size_t n = ri[nr];
      const unsigned nt = unpack_pool->size();
      if (nt > 1 && n >= 8192 && nr > 1) {
        unpack_pool->run([&](unsigned q) {
          const size_t ka = std::lower_bound(ri.begin(), ri.begin() + nr, n * q / nt) - ri.begin();
          const size_t kb = std::lower_bound(ri.begin(), ri.begin() + nr, n * (q + 1) / nt) - ri.begin();
          scatter(ka, kb);
        });
      } else {
        scatter(0, nr);
      }
    }
    lk.lock();
    for (Slot* s : ready) {
      for (auto& r : s->reqs) r->done = true;
      s->reqs.clear();
      budget_used -= s->bytes;
      s->bytes = 0;
      free_slots.push_back(s);
    }
    cv_done.notify_all();
    cv_work.notify_all();
  }
}

namespace {

using sigops::NBUCKETS;
using sigops::STREAM_QUEUES;
using sigops_host::StreamBufs;
using sigops_host::StreamDesc;
using sigops_host::StreamItem;
using sigops_host::ST_LANES;

constexpr uint8_t ST_PENDING = 0xFF;  // a lane whose item waits for the careful pass
constexpr uint32_t Q_K1 = 0, Q_R1 = 1, Q_ED = 2;  // the ring's queues
// relative cost of a tile of each queue (per-scheme kernel times), for the home split
constexpr double Q_COST[STREAM_QUEUES] = {1.0, 1.34, 1.3};

uint32_t queue_of(uint32_t bucket) {
  return bucket == sigops::BUCKET_K1 ? Q_K1 : bucket == sigops::BUCKET_R1 ? Q_R1 : Q_ED;
}

struct RingItem {
  Request* r;
  uint32_t idx;
};

// admitted items not yet in a tile, of one bucket: runs of a request's items (its order[begin ..
// end)), taken from the front. Admission appends one run per request and bucket, not one entry per
// item.
struct Run {
  Request* r;
  uint32_t begin, end;
  uint32_t bucket;
};

struct OpenList {
  std::deque<Run> runs;
  size_t count = 0;
  size_t size() const { return count; }
  bool empty() const { return count == 0; }
  void push(const Run& x) {
    if (x.end > x.begin) {
      runs.push_back(x);
      count += x.end - x.begin;
    }
  }
  // the first (up to) c items, not removed
  uint32_t peek(RingItem* out, uint32_t c) const {
    uint32_t k = 0;
    for (auto it = runs.begin(); it != runs.end() && k < c; ++it)
      for (uint32_t j = it->begin; j < it->end && k < c; j++) out[k++] = RingItem{it->r, it->r->order[j]};
    return k;
  }
  void take(uint32_t c) {
    count -= c;
    while (c) {
      Run& f = runs.front();
      const uint32_t k = std::min(c, f.end - f.begin);
      f.begin += k;
      c -= k;
      if (f.begin == f.end) runs.pop_front();
    }
  }
  void clear() {
    runs.clear();
    count = 0;
  }
};

// host bookkeeping of one ring position
struct RingPos {
  uint64_t expect = 0;     // the done word the kernel writes for the tile held here
  uint32_t seq = 0;        // its launch
  uint32_t cnt = 0;
  bool ed = false;
  bool seen = false;       // its results were read
  bool held = false;       // lanes wait for the launch's careful pass; the position stays taken
  uint64_t pend_mask = 0;  // those lanes
  RingItem items[ST_LANES];
};

// one persistent launch
struct Epoch {
  uint32_t seq = 0;
  bool closed = false;
  bool failed = false;  // the launch was not made: its tiles are lost
  hipEvent_t ev = nullptr;
};

// regions of the ring's pinned block, each 256-byte aligned (positions' layout as in
// stream_host.hpp; every position owns a fixed message slot of `ms` bytes)
struct RingRegions {
  size_t desc, in, mrel, mlen, out, st, done, qpos, msgs, total;
};

RingRegions ring_regions(size_t R, uint64_t ms) {
  auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
  RingRegions r;
  size_t at = 0;
  r.desc = at;
  at = al(at + R * 16);
  r.in = at;
  at = al(at + R * sigops_host::ST_IN_CHUNKS * ST_LANES * 16);
  r.mrel = at;
  at = al(at + R * ST_LANES * 4);
  r.mlen = at;
  at = al(at + R * ST_LANES * 4);
  r.out = at;
  at = al(at + R * sigops_host::ST_OUT_CHUNKS * ST_LANES * 16);
  r.st = at;
  at = al(at + R * ST_LANES);
  r.done = at;
  at = al(at + R * 8);
  r.qpos = at;
  at = al(at + STREAM_QUEUES * R * 4);
  r.msgs = at;
  at = al(at + R * ms + 256);
  r.total = at;
  return r;
}

void* dev_addr(void* host) {
  void* d = nullptr;
  return hipHostGetDevicePointer(&d, host, 0) == hipSuccess ? d : nullptr;
}

struct RingEngine {
  // Each launch publishes through its own words (seq % READY_SLOTS, one per queue): a closed
  // launch's wavefronts may still read its counts while the next launch publishes; at most
  // READY_SLOTS - 1 launches are outstanding at once
  static constexpr uint32_t READY_SLOTS = 16;
  static constexpr size_t CTL_BYTES = READY_SLOTS * 4 * 8 + 64;  // then u32 stop, u32 err

  sigops_server* S = nullptr;
  sigops_ctx* ctx = nullptr;  // the resident tables, the lane tables and the stream
  sigops::RingResources res{};
  uint32_t R = 0, mask = 0, W = 0;  // positions, positions - 1, wavefronts of a launch
  uint64_t ms = 0;                  // message bytes per position
  uint64_t idle_ns = 0;
  RingRegions reg{};
  uint8_t* pin = nullptr;  // positions' regions, queue tables and message slots
  uint8_t* ctl = nullptr;  // ready words, stop, err
  uint8_t* dpin = nullptr;
  uint8_t* dctl = nullptr;
  uint8_t* dmsgs = nullptr;  // HBM copies of the tiles' messages (same offsets)
  uint32_t* dctr = nullptr;
  uint32_t* dpend = nullptr;
  StreamBufs B{};
  uint64_t* done = nullptr;
  uint32_t* qpos = nullptr;
  std::vector<RingPos> pos;
  std::vector<hipEvent_t> free_ev;
  // dispatcher only
  OpenList open[NBUCKETS];  // admitted items not yet in a tile, per bucket
  uint32_t seq = 0;
  bool ep_open = false, ep_launched = false;
  uint32_t qcount[STREAM_QUEUES] = {};  // tiles published per queue in the open launch
  uint32_t tag_next = 0;
  double work[STREAM_QUEUES] = {0.43, 0.35 / 1.34, 0.22 / 1.3};  // recent tiles per queue (decaying)
  uint64_t last_busy = 0;
  // shared between the dispatcher and the completion thread
  std::mutex pmu;                  // free positions and newly published ones
  std::vector<uint32_t> free_pos;  // positions no tile holds
  std::vector<uint32_t> published; // positions published since the completion thread last looked
  std::atomic<uint64_t> n_pub{0}, n_seen{0};
  std::atomic<uint32_t> close_seq{0};  // a launch with lanes waiting for its careful pass
  std::atomic<bool> disp_done{false};
  std::mutex emu;  // epochs, free_ev
  std::deque<Epoch> epochs;
  std::thread disp, comp;
  // packing and unpacking are spread over pools once there are a few tiles at a time
  std::unique_ptr<sigops_host::Pool> pack_pool, unpack_pool;
  struct Planned {
    uint32_t p, bucket;
  };
  std::vector<Planned> planned;  // tiles placed in this publication round, packed together
  std::vector<uint32_t> qbuckets[STREAM_QUEUES];  // each queue's buckets in processing order

  int init(sigops_server* srv);
  void release();
  void admit(Request* r);
  bool plan(uint32_t q, const RingItem* items, uint32_t c);
  void pack_planned();
  void publish();
  void open_epoch();
  void publish_counts();
  void close_epoch();
  void dispatch_loop();
  void read_tile(RingPos& P, uint32_t p, std::vector<Request*>& finished);
  void finish_lane(RingPos& P, uint32_t p, uint32_t l, std::vector<Request*>& finished, int err);
  void complete_loop();
  uint64_t* ready(uint32_t q) { return (uint64_t*)ctl + (seq % READY_SLOTS) * 4 + q; }
};

int RingEngine::init(sigops_server* srv) {
  S = srv;
  const sigops_server_config& cfg = S->cfg;
  R = cfg.ring_tiles;
  mask = R - 1;
  ms = (cfg.ring_msg_bytes / R) & ~(uint64_t)127;
  idle_ns = (uint64_t)cfg.idle_close_us * 1000;
  if (ms < 1024) return SIGOPS_E_ARG;
  sigops_config cc{};
  cc.device = cfg.device;
  cc.max_batch = 1u << 16;  // the context's own batch buffers are not used by the ring
  ctx = sigops_create(&cc);
  if (!ctx) return SIGOPS_E_DEVICE;
  if (sigops_ring_resources(ctx, &res)) return SIGOPS_E_DEVICE;
  W = res.grid * 4;
  reg = ring_regions(R, ms);
  if (hipHostMalloc((void**)&pin, reg.total, hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess) {
    pin = nullptr;
    return SIGOPS_E_NOMEM;
  }
  if (hipHostMalloc((void**)&ctl, CTL_BYTES, hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess) {
    ctl = nullptr;
    return SIGOPS_E_NOMEM;
  }
  memset(pin, 0, reg.total);
  memset(ctl, 0, CTL_BYTES);
  dpin = (uint8_t*)dev_addr(pin);
  dctl = (uint8_t*)dev_addr(ctl);
  if (!dpin || !dctl) return SIGOPS_E_DEVICE;
  if (hipMalloc((void**)&dmsgs, R * ms + 256) != hipSuccess ||
      hipMalloc((void**)&dctr, sigops::STREAM_CTR_BYTES) != hipSuccess ||
      hipMalloc((void**)&dpend, (size_t)R * ST_LANES * sizeof(uint32_t)) != hipSuccess)
    return SIGOPS_E_NOMEM;
  B = StreamBufs{(StreamDesc*)(pin + reg.desc), pin + reg.in, (uint32_t*)(pin + reg.mrel), (uint32_t*)(pin + reg.mlen),
                 pin + reg.msgs, pin + reg.out, pin + reg.st, nullptr};
  done = (uint64_t*)(pin + reg.done);
  qpos = (uint32_t*)(pin + reg.qpos);
  pos.resize(R);
  for (uint32_t p = R; p-- > 0;) free_pos.push_back(p);
  pack_pool = std::make_unique<sigops_host::Pool>(std::min(8u, sigops_host::host_threads()));
  for (uint32_t r = 0; r < NBUCKETS; r++) qbuckets[queue_of(res.bucket_order[r])].push_back(res.bucket_order[r]);
  unpack_pool = std::make_unique<sigops_host::Pool>(std::min(4u, sigops_host::host_threads()));
  for (int k = 0; k < 4; k++) {
    hipEvent_t e;
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return SIGOPS_E_DEVICE;
    free_ev.push_back(e);
  }
  disp = std::thread([this] { dispatch_loop(); });
  comp = std::thread([this] { complete_loop(); });
  return SIGOPS_OK;
}

void RingEngine::release() {
  if (disp.joinable()) disp.join();
  if (comp.joinable()) comp.join();
  if (res.stream) (void)hipStreamSynchronize(res.stream);
  for (auto& e : epochs)
    if (e.ev) free_ev.push_back(e.ev);
  for (hipEvent_t e : free_ev) (void)hipEventDestroy(e);
  free_ev.clear();
  if (dmsgs) (void)hipFree(dmsgs);
  if (dctr) (void)hipFree(dctr);
  if (dpend) (void)hipFree(dpend);
  if (pin) (void)hipHostFree(pin);
  if (ctl) (void)hipHostFree(ctl);
  if (ctx) sigops_destroy(ctx);
  dmsgs = nullptr;
  dctr = dpend = nullptr;
  pin = ctl = nullptr;
  ctx = nullptr;
}

// a request's items into the bucket lists: one run per bucket it has items of
void RingEngine::admit(Request* r) {
  r->remaining = r->n;
  for (uint32_t b = 0; b < NBUCKETS; b++) open[b].push(Run{r, r->bstart[b], r->bstart[b + 1], b});
}

// places c items of bucket q in a free position of its queue (packed by pack_planned); false
// when no position is free (the items stay where they are)
bool RingEngine::plan(uint32_t q, const RingItem* items, uint32_t c) {
  if (!ep_open) {
    std::lock_guard<std::mutex> lk(emu);
    if (epochs.size() >= READY_SLOTS - 1) return false;  // the completion thread retires them first
  }
  uint32_t p;
  {
    std::lock_guard<std::mutex> lk(pmu);
    if (free_pos.empty()) return false;
    p = free_pos.back();
    free_pos.pop_back();
  }
  if (!ep_open) open_epoch();
  RingPos& P = pos[p];
  for (uint32_t l = 0; l < c; l++) P.items[l] = items[l];
  P.cnt = c;
  P.ed = queue_of(q) == Q_ED;
  P.seen = P.held = false;
  P.pend_mask = 0;
  P.seq = seq;
  planned.push_back(Planned{p, q});
  return true;
}

// packs the planned tiles into their positions (on the pool when there are several), then enters
// them in their queues' position tables in order; publish_counts makes them visible
void RingEngine::pack_planned() {
  auto pack_one = [&](const Planned& t, uint32_t tag) {
    RingPos& P = pos[t.p];
    sigops_host::stream_pack_tile(B, t.p, t.bucket, P.cnt, (uint64_t)t.p * ms, tag, [&](uint32_t l) {
      const Request& r = *P.items[l].r;
      const uint32_t i = P.items[l].idx;
      return StreamItem{r.sigs.data() + (size_t)i * 64, r.aux.data() + (size_t)i * 32,
                        r.msgs.data() + r.off[i], r.off[i + 1] - r.off[i]};
    });
  };
  const uint32_t tag0 = tag_next;
  tag_next += (uint32_t)planned.size();
  if (planned.size() >= 8 && pack_pool->size() > 1) {
    std::atomic<size_t> next{0};
    pack_pool->run([&](unsigned) {
      for (size_t k; (k = next.fetch_add(1)) < planned.size();) pack_one(planned[k], tag0 + (uint32_t)k);
    });
  } else {
    for (size_t k = 0; k < planned.size(); k++) pack_one(planned[k], tag0 + (uint32_t)k);
  }
  {
    std::lock_guard<std::mutex> lk(pmu);
    for (size_t k = 0; k < planned.size(); k++) {
      const Planned& t = planned[k];
      RingPos& P = pos[t.p];
      const uint32_t qu = queue_of(t.bucket);
      P.expect = ((uint64_t)seq << 32) | (tag0 + (uint32_t)k);
      qpos[(size_t)qu * R + (qcount[qu] & mask)] = t.p;
      qcount[qu]++;
      work[qu] += 1.0;
      S->n_items += P.cnt;
      published.push_back(t.p);
    }
  }
  n_pub.fetch_add(planned.size(), std::memory_order_release);
  planned.clear();
}

void RingEngine::open_epoch() {
  seq = seq + 1 ? seq + 1 : 1;  // never 0, the cleared done words
  ep_open = true;
  ep_launched = false;
  for (auto& k : qcount) k = 0;
  for (uint32_t q = 0; q < STREAM_QUEUES; q++) __atomic_store_n(ready(q), (uint64_t)seq << 32, __ATOMIC_RELEASE);
}

// publishes the open launch's tiles, launching it with its first ones
void RingEngine::publish_counts() {
  for (uint32_t q = 0; q < STREAM_QUEUES; q++)
    __atomic_store_n(ready(q), ((uint64_t)seq << 32) | qcount[q], __ATOMIC_RELEASE);
  if (ep_launched) return;
  ep_launched = true;
  Epoch e;
  e.seq = seq;
  {
    std::lock_guard<std::mutex> g(emu);
    if (!free_ev.empty()) {
      e.ev = free_ev.back();
      free_ev.pop_back();
    }
  }
  if (!e.ev && hipEventCreateWithFlags(&e.ev, hipEventDisableTiming) != hipSuccess) e.ev = nullptr;
  // home queues of the launch's blocks in proportion to the recent work per queue, by tile cost
  double share[STREAM_QUEUES], tot = 0;
  for (uint32_t q = 0; q < STREAM_QUEUES; q++) tot += share[q] = work[q] * Q_COST[q];
  for (auto& w : work) w *= 0.5;
  const uint32_t s0 = (uint32_t)(256.0 * share[0] / tot + 0.5);
  const uint32_t s1 = (uint32_t)(256.0 * (share[0] + share[1]) / tot + 0.5);
  sigops::StreamLaunch L{};
  L.desc = dpin + reg.desc;
  L.in = dpin + reg.in;
  L.mrel = (const uint32_t*)(dpin + reg.mrel);
  L.mlen = (const uint32_t*)(dpin + reg.mlen);
  L.hmsgs = dpin + reg.msgs;
  L.dmsgs = dmsgs;
  L.msg_cap = R * ms + 256;  // both message buffers hold every position's slot plus 256 bytes
  L.out = dpin + reg.out;
  L.st = dpin + reg.st;
  L.tile_done = (uint64_t*)(dpin + reg.done);
  L.ready = (const uint64_t*)dctl + (seq % READY_SLOTS) * 4;
  L.qpos = (const uint32_t*)(dpin + reg.qpos);
  L.nq = STREAM_QUEUES;
  L.split0 = s0;
  L.split1 = std::max(s0, s1);
  L.stop = (const uint32_t*)(dctl + READY_SLOTS * 4 * 8);
  L.err = (uint32_t*)(dctl + READY_SLOTS * 4 * 8 + 4);
  L.ctr = dctr;
  L.pend_list = dpend;
  L.pend_cap = R * ST_LANES;
  L.gk1 = res.gk1;
  L.gr1 = res.gr1;
  L.gb = res.gb;
  L.lane_tab = res.lane_tab;
  L.slots = sigops::STREAM_CLOSED - 1;
  L.base = 0;
  L.mask = mask;
  L.seq = seq;
  L.lane_slots = res.lane_slots;
  L.grid = res.grid;
  L.g_window = res.g_window;
  sigops::stream_poll_ticks(S->cfg.device, L);
  const bool ok = e.ev && sigops::launch_stream(L, res.stream) == hipSuccess &&
                  hipEventRecord(e.ev, res.stream) == hipSuccess;
  if (!ok) {
    // the tiles of this launch will not run: the completion thread fails them
    e.failed = true;
    e.closed = true;
    ep_open = false;
  }
  S->n_batches++;
  std::lock_guard<std::mutex> g(emu);
  epochs.push_back(e);
}

void RingEngine::close_epoch() {
  if (!ep_open) return;
  ep_open = false;
  if (!ep_launched) return;  // nothing was published
  for (uint32_t q = 0; q < STREAM_QUEUES; q++)
    __atomic_store_n(ready(q), ((uint64_t)seq << 32) | sigops::STREAM_CLOSED | qcount[q], __ATOMIC_RELEASE);
  std::lock_guard<std::mutex> g(emu);
  for (auto& e : epochs)
    if (e.seq == seq) e.closed = true;
}

// Full tiles of every bucket go out at once (see below for the order); when fewer tiles are in
// flight than the launch has wavefronts, the remainders go out too (one tile per secp bucket, the Ed25519 bins merged in
// increasing length order, so a tile's lanes still hash similar lengths). A tile's messages fit
// its position's message slot.
void RingEngine::publish() {
  const uint64_t before = n_pub.load(std::memory_order_relaxed);
  bool full = false;  // out of positions
  RingItem buf[ST_LANES];
  // one tile of up to 64 items from the front of v (bucket q), cut where its messages would
  // overflow a position's message slot
  auto one = [&](OpenList& v, uint32_t q) {
    uint32_t c = v.peek(buf, ST_LANES);
    uint64_t bytes = 0;
    for (uint32_t l = 0; l < c; l++) {
      bytes += buf[l].r->off[buf[l].idx + 1] - buf[l].r->off[buf[l].idx];
      if (l > 0 && sigops_host::stream_region_bytes(bytes) > ms) {
        c = l;
        break;
      }
    }
    if (!plan(q, buf, c)) {
      full = true;
      return;
    }
    v.take(c);
  };
  // full tiles one at a time from the queue with the most of them (within a queue, its buckets in
  // processing order), so that when positions run short every queue still gets its share and the
  // wavefronts at home in it stay busy
  auto full_tiles = [&](uint32_t qu, uint32_t& first) {
    size_t t = 0;
    first = NBUCKETS;
    for (uint32_t b : qbuckets[qu]) {
      if (open[b].size() < ST_LANES) continue;
      t += open[b].size() / ST_LANES;
      if (first == NBUCKETS) first = b;
    }
    return t;
  };
  size_t tiles[STREAM_QUEUES];
  uint32_t first[STREAM_QUEUES];
  for (uint32_t qu = 0; qu < STREAM_QUEUES; qu++) tiles[qu] = full_tiles(qu, first[qu]);
  while (!full) {
    uint32_t qu = 0;
    for (uint32_t k = 1; k < STREAM_QUEUES; k++)
      if (tiles[k] > tiles[qu]) qu = k;
    if (!tiles[qu]) break;
    one(open[first[qu]], first[qu]);
    tiles[qu] = full_tiles(qu, first[qu]);
  }
  const uint64_t inflight =
      n_pub.load(std::memory_order_relaxed) + planned.size() - n_seen.load(std::memory_order_acquire);
  if (!full && inflight < W) {
    // the remainders: one tile per secp bucket, the Ed25519 bins merged in increasing length order
    while (!full && !open[sigops::BUCKET_R1].empty()) one(open[sigops::BUCKET_R1], sigops::BUCKET_R1);
    while (!full && !open[sigops::BUCKET_K1].empty()) one(open[sigops::BUCKET_K1], sigops::BUCKET_K1);
    OpenList ed;
    uint32_t last = 0;
    for (uint32_t b = sigops::BUCKET_ED0; b < NBUCKETS; b++) {
      if (open[b].empty()) continue;
      last = b;
      for (const Run& x : open[b].runs) ed.push(x);
      open[b].clear();
    }
    while (!full && !ed.empty()) one(ed, last);
    for (const Run& x : ed.runs) open[x.bucket].push(x);  // what did not fit goes back to its bin
  }
  if (!planned.empty()) pack_planned();
  if (n_pub.load(std::memory_order_relaxed) != before) publish_counts();
}

void RingEngine::dispatch_loop() {
  (void)hipSetDevice(S->cfg.device);
  std::vector<std::shared_ptr<Request>> taken;
  for (;;) {
    bool stopping, queue_empty;
    {
      std::unique_lock<std::mutex> lk(S->mu);
      auto admissible = [&] {
        return !S->queue.empty() && S->budget_used + S->queue.front()->bytes <= S->cfg.hbm_budget;
      };
      S->cv_work.wait_for(lk, std::chrono::microseconds(100),
                          [&] { return S->stopping || S->ring_wake || admissible(); });
      S->ring_wake = false;
      while (admissible()) {
        S->budget_used += S->queue.front()->bytes;
        taken.push_back(S->queue.front());
        S->queue.pop_front();
      }
      stopping = S->stopping;
      queue_empty = S->queue.empty();
    }
    for (auto& r : taken) admit(r.get());
    taken.clear();
    publish();
    bool open_items = false;
    for (auto& v : open) open_items = open_items || !v.empty();
    const bool idle = !open_items && queue_empty &&
                      n_seen.load(std::memory_order_acquire) == n_pub.load(std::memory_order_relaxed);
    const uint64_t t = now_ns();
    if (!idle) last_busy = t;
    if (ep_open && (close_seq.load() == seq || (idle && t - last_busy >= idle_ns) || (idle && stopping)))
      close_epoch();
    if (stopping && idle && !ep_open) break;
  }
  disp_done.store(true, std::memory_order_release);
}

void RingEngine::finish_lane(RingPos& P, uint32_t p, uint32_t l, std::vector<Request*>& finished, int err) {
  Request& r = *P.items[l].r;
  const uint32_t i = P.items[l].idx;
  if (err) {
    r.err = err;
  } else {
    const uint8_t st = B.st[(size_t)p * ST_LANES + l];
    if (!P.ed) {
      const uint8_t* o = B.out + (size_t)p * (sigops_host::ST_OUT_CHUNKS * ST_LANES * 16);
      for (uint32_t k = 0; k < sigops_host::ST_OUT_CHUNKS; k++)
        memcpy(r.out.data() + (size_t)i * 64 + k * 16, o + (k * ST_LANES + l) * 16, 16);
    }
    r.status[i] = st;
    if (st > (P.ed ? SIGOPS_ED_MISMATCH : SIGOPS_ST_INFINITY)) r.err = SIGOPS_E_CANARY;
  }
  if (r.remaining.fetch_sub(1) == 1) finished.push_back(&r);
}

void RingEngine::read_tile(RingPos& P, uint32_t p, std::vector<Request*>& finished) {
  for (uint32_t l = 0; l < P.cnt; l++) {
    if (!P.ed && B.st[(size_t)p * ST_LANES + l] == ST_PENDING) {
      P.pend_mask |= 1ull << l;
      continue;
    }
    finish_lane(P, p, l, finished, 0);
  }
  if (P.pend_mask) {
    P.held = true;
    close_seq.store(P.seq);
  }
  P.seen = true;
}

void RingEngine::complete_loop() {
  (void)hipSetDevice(S->cfg.device);
  std::vector<uint32_t> live;  // published positions not yet given back
  std::vector<uint32_t> freed, ready_tiles;
  std::vector<Request*> finished;
  std::vector<std::vector<Request*>> fin(unpack_pool->size());  // per pool thread
  for (;;) {
    bool progress = false;
    {
      std::lock_guard<std::mutex> lk(pmu);
      live.insert(live.end(), published.begin(), published.end());
      published.clear();
    }
    ready_tiles.clear();
    for (uint32_t p : live) {
      const RingPos& P = pos[p];
      if (!P.seen && __atomic_load_n(&done[p], __ATOMIC_ACQUIRE) == P.expect) ready_tiles.push_back(p);
    }
    if (ready_tiles.size() >= 8 && unpack_pool->size() > 1) {
      std::atomic<size_t> next{0};
      unpack_pool->run([&](unsigned t) {
        for (size_t k; (k = next.fetch_add(1)) < ready_tiles.size();)
          read_tile(pos[ready_tiles[k]], ready_tiles[k], fin[t]);
      });
      for (auto& f : fin) {
        finished.insert(finished.end(), f.begin(), f.end());
        f.clear();
      }
    } else {
      for (uint32_t p : ready_tiles) read_tile(pos[p], p, finished);
    }
    if (!ready_tiles.empty()) {
      n_seen.fetch_add(ready_tiles.size(), std::memory_order_release);
      progress = true;
    }
    // launches that have ended: the careful pass has rewritten the held lanes; a tile the launch
    // did not finish is lost (its requests fail)
    for (;;) {
      Epoch e;
      {
        std::lock_guard<std::mutex> g(emu);
        if (epochs.empty() || !epochs.front().closed) break;
        if (!epochs.front().failed && hipEventQuery(epochs.front().ev) == hipErrorNotReady) break;
        e = epochs.front();
        epochs.pop_front();
      }
      const bool dev_ok = !e.failed && hipEventQuery(e.ev) == hipSuccess;
      {
        // every tile of the launch was published before it was closed
        std::lock_guard<std::mutex> lk(pmu);
        live.insert(live.end(), published.begin(), published.end());
        published.clear();
      }
      for (uint32_t p : live) {
        RingPos& P = pos[p];
        if (P.seq != e.seq) continue;
        if (!P.seen && dev_ok && __atomic_load_n(&done[p], __ATOMIC_ACQUIRE) == P.expect) {
          read_tile(P, p, finished);
          n_seen.fetch_add(1, std::memory_order_release);
        }
        if (!P.seen) {
          for (uint32_t l = 0; l < P.cnt; l++) finish_lane(P, p, l, finished, SIGOPS_E_DEVICE);
          P.seen = true;
          n_seen.fetch_add(1, std::memory_order_release);
        }
        if (P.held) {
          for (uint32_t l = 0; l < P.cnt; l++)
            if (P.pend_mask >> l & 1) finish_lane(P, p, l, finished, dev_ok ? 0 : SIGOPS_E_DEVICE);
          P.held = false;
        }
      }
      if (e.ev) {
        std::lock_guard<std::mutex> g(emu);
        free_ev.push_back(e.ev);
      }
      progress = true;
    }
    // positions whose results are read go back to the free list (held ones after their launch)
    size_t k = 0;
    for (uint32_t p : live) {
      if (pos[p].seen && !pos[p].held)
        freed.push_back(p);
      else
        live[k++] = p;
    }
    live.resize(k);
    if (!freed.empty()) {
      std::lock_guard<std::mutex> lk(pmu);
      free_pos.insert(free_pos.end(), freed.begin(), freed.end());
      freed.clear();
      progress = true;
    }
    if (progress) {
      const uint64_t t = now_ns();
      std::lock_guard<std::mutex> lk(S->mu);
      for (Request* r : finished) {
        r->t_done = t;
        r->done = true;
        S->budget_used -= r->bytes;
      }
      finished.clear();
      S->ring_wake = true;
      S->cv_work.notify_all();
      S->cv_done.notify_all();
      continue;
    }
    if (disp_done.load(std::memory_order_acquire) && live.empty()) {
      std::lock_guard<std::mutex> g(emu);
      std::lock_guard<std::mutex> lk(pmu);
      if (epochs.empty() && published.empty()) break;
    }
    std::this_thread::sleep_for(std::chrono::microseconds(live.empty() ? 50 : 10));
  }
}

}  // namespace

extern "C" {

sigops_server* sigops_server_create(const sigops_server_config* cfg_in) {
  sigops_server_config cfg{};
  if (cfg_in) cfg = *cfg_in;
  if (!cfg.slots) cfg.slots = 2;
  if (!cfg.hbm_budget) cfg.hbm_budget = 1ull << 30;
  if (!cfg.max_batch) cfg.max_batch = 1u << 20;
  if (!cfg.ring_tiles) cfg.ring_tiles = 4096;
  if (!cfg.ring_msg_bytes) cfg.ring_msg_bytes = 256ull << 20;
  if (!cfg.idle_close_us) cfg.idle_close_us = 1000;
  if (cfg.mode > SIGOPS_SERVE_RING) return nullptr;
  if (cfg.ring_tiles < 64 || cfg.ring_tiles > (1u << 20) || (cfg.ring_tiles & (cfg.ring_tiles - 1))) return nullptr;
  if (cfg.ring_msg_bytes / cfg.ring_tiles < 1024 || cfg.ring_msg_bytes >= (1ull << 32)) return nullptr;
  auto* S = new (std::nothrow) sigops_server();
  if (!S) return nullptr;
  S->cfg = cfg;
  if (cfg.mode == SIGOPS_SERVE_RING) {
    S->ring = std::make_unique<RingEngine>();
    if (S->ring->init(S) != SIGOPS_OK) {
      {
        std::lock_guard<std::mutex> g(S->mu);
        S->stopping = true;
      }
      S->cv_work.notify_all();
      S->ring->release();
      delete S;
      return nullptr;
    }
    return S;
  }
  S->pack_pool = std::make_unique<sigops_host::Pool>(std::min(4u, sigops_host::host_threads()));
  S->unpack_pool = std::make_unique<sigops_host::Pool>(std::min(4u, sigops_host::host_threads()));
  for (uint32_t k = 0; k < cfg.slots; k++) {
    auto slot = std::make_unique<Slot>();
    sigops_config cc{};
    cc.device = cfg.device;
    cc.blocks_per_cu = cfg.blocks_per_cu;
    cc.max_batch = cfg.max_batch;
    slot->ctx = sigops_create(&cc);
    if (!slot->ctx || hipStreamCreateWithFlags(&slot->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&slot->done_ev, hipEventDisableTiming) != hipSuccess) {
      if (slot->ctx) sigops_destroy(slot->ctx);
      S->slots.clear();
      delete S;
      return nullptr;
    }
    S->free_slots.push_back(slot.get());
    S->slots.push_back(std::move(slot));
  }
  S->dispatcher = std::thread([S] { S->dispatch_loop(); });
  S->completer = std::thread([S] { S->complete_loop(); });
  return S;
}

void sigops_server_destroy(sigops_server* S) {
  if (!S) return;
  {
    std::lock_guard<std::mutex> g(S->mu);
    S->stopping = true;
  }
  S->cv_work.notify_all();
  S->cv_done.notify_all();
  if (S->ring) {
    S->ring->release();  // the dispatcher finishes the queued work first
    delete S;
    return;
  }
  S->dispatcher.join();
  S->completer.join();
  for (auto& s : S->slots) {
    (void)hipStreamSynchronize(s->stream);
    (void)hipEventDestroy(s->done_ev);
    (void)hipStreamDestroy(s->stream);
    sigops_destroy(s->ctx);
  }
  S->slots.clear();
  delete S;
}

int64_t sigops_server_submit(sigops_server* S, const uint8_t* schemes, const uint8_t* sigs, const uint8_t* aux,
                             const uint8_t* msgs, const uint64_t* msg_off, size_t n) {
  if (!S || (n && (!schemes || !sigs || !aux || !msg_off))) return SIGOPS_E_ARG;
  if (n > S->cfg.max_batch) return SIGOPS_E_ARG;
  auto r = std::make_shared<Request>();
  r->n = n;
  r->t_submit = now_ns();
  if (n) {
    for (size_t i = 0; i < n; i++) {
      if (schemes[i] > SIGOPS_SCHEME_ED25519) return SIGOPS_E_ARG;
      if (msg_off[i + 1] < msg_off[i] || msg_off[i + 1] - msg_off[i] > 0xFFFFFFFFull) return SIGOPS_E_ARG;
    }
    const uint64_t base = msg_off[0], total = msg_off[n] - base;
    if (total && !msgs) return SIGOPS_E_ARG;
    if (S->ring)
      for (size_t i = 0; i < n; i++)
        if (sigops_host::stream_region_bytes(msg_off[i + 1] - msg_off[i]) > S->ring->ms) return SIGOPS_E_ARG;
    r->schemes.assign(schemes, schemes + n);
    r->sigs.assign(sigs, sigs + n * 64);
    r->aux.assign(aux, aux + n * 32);
    if (total) r->msgs.assign(msgs + base, msgs + base + total);
    r->off.resize(n + 1);
    for (size_t i = 0; i <= n; i++) r->off[i] = msg_off[i] - base;
  }
  r->bytes = request_bytes(n, r->msgs.size());
  if (r->bytes > S->cfg.hbm_budget) return SIGOPS_E_ARG;
  if (S->ring && n) {
    // the ring's admission takes a request's items bucket by bucket: a counting sort here, on the
    // submitting thread, and the result buffers
    std::vector<uint8_t> bk(n);
    uint32_t cnt[sigops::NBUCKETS] = {};
    for (size_t i = 0; i < n; i++) {
      bk[i] = (uint8_t)sigops_host::stream_bucket(r->schemes[i], r->off[i + 1] - r->off[i]);
      cnt[bk[i]]++;
    }
    r->bstart[0] = 0;
    for (uint32_t b = 0; b < sigops::NBUCKETS; b++) r->bstart[b + 1] = r->bstart[b] + cnt[b];
    uint32_t at[sigops::NBUCKETS];
    for (uint32_t b = 0; b < sigops::NBUCKETS; b++) at[b] = r->bstart[b];
    r->order.resize(n);
    for (size_t i = 0; i < n; i++) r->order[at[bk[i]]++] = (uint32_t)i;
    r->out.assign(n * 64, 0);
    r->status.assign(n, 0);
  }
  std::lock_guard<std::mutex> g(S->mu);
  r->ticket = S->next_ticket++;
  S->tickets[r->ticket] = r;
  if (n == 0) {
    r->done = true;
    r->t_done = r->t_submit;
  } else {
    S->queue.push_back(r);
    S->cv_work.notify_one();
  }
  return r->ticket;
}

int sigops_server_wait(sigops_server* S, int64_t ticket, uint8_t* out, uint8_t* status, int timeout_ms,
                       uint64_t* t_submit_ns, uint64_t* t_done_ns) {
  if (!S) return SIGOPS_E_ARG;
  std::unique_lock<std::mutex> lk(S->mu);
  auto it = S->tickets.find(ticket);
  if (it == S->tickets.end()) return SIGOPS_E_ARG;
  std::shared_ptr<Request> r = it->second;
  auto pred = [&] { return r->done; };
  if (timeout_ms < 0) {
    S->cv_done.wait(lk, pred);
  } else if (!S->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), pred)) {
    return 1;
  }
  S->tickets.erase(ticket);
  lk.unlock();
  if (t_submit_ns) *t_submit_ns = r->t_submit;
  if (t_done_ns) *t_done_ns = r->t_done;
  if (const int e = r->err.load()) return e;
  if (out && r->n) memcpy(out, r->out.data(), r->n * 64);
  if (status && r->n) memcpy(status, r->status.data(), r->n);
  return 0;
}

int sigops_server_stats(sigops_server* S, uint64_t* batches, uint64_t* items) {
  if (!S) return SIGOPS_E_ARG;
  if (batches) *batches = S->n_batches.load();
  if (items) *items = S->n_items.load();
  return 0;
}

int sigops_server_replay(sigops_server* S, const uint8_t* schemes, const uint8_t* sigs, const uint8_t* aux,
                         const uint8_t* msgs, const uint64_t* msg_off, const uint64_t* req_off,
                         const uint64_t* arrival_ns, size_t n_requests, uint64_t* latency_ns, uint8_t* out,
                         uint8_t* status) {
  if (!S || !req_off || !arrival_ns || !latency_ns) return SIGOPS_E_ARG;
  constexpr int64_t PENDING = INT64_MIN;
  std::vector<std::atomic<int64_t>> tickets(n_requests);
  for (auto& t : tickets) t.store(PENDING, std::memory_order_relaxed);
  std::atomic<int> err{0};
  const uint64_t t0 = now_ns();
  // collector: waits for the requests in arrival order; the server stamps each completion time
  // itself, so waiting in order does not inflate the latencies, which run from the scheduled
  // arrival (a client that falls behind its schedule shows up in them)
  std::thread collector([&] {
    for (size_t r = 0; r < n_requests; r++) {
      int64_t tk;
      while ((tk = tickets[r].load(std::memory_order_acquire)) == PENDING)
        std::this_thread::sleep_for(std::chrono::microseconds(5));
      uint64_t ts = 0, td = 0;
      const size_t a = req_off[r];
      const int rc =
          sigops_server_wait(S, tk, out ? out + a * 64 : nullptr, status ? status + a : nullptr, -1, &ts, &td);
      if (rc) err = rc;
      const uint64_t due = t0 + arrival_ns[r];
      latency_ns[r] = td > due ? td - due : 0;
    }
  });
  // clients: request r is submitted by client r mod K at its arrival time, as requests from
  // several independent callers would be
  const unsigned K = (unsigned)std::min<size_t>(4, std::max<size_t>(1, n_requests));
  std::vector<std::thread> clients;
  for (unsigned c = 0; c < K; c++)
    clients.emplace_back([&, c] {
      for (size_t r = c; r < n_requests; r += K) {
        const uint64_t due = t0 + arrival_ns[r];
        for (;;) {
          const uint64_t t = now_ns();
          if (t >= due) break;
          if (due - t > 200000) std::this_thread::sleep_for(std::chrono::nanoseconds(due - t - 100000));
        }
        const size_t a = req_off[r], b = req_off[r + 1];
        int64_t tk = sigops_server_submit(S, schemes + a, sigs + a * 64, aux + a * 32, msgs, msg_off + a, b - a);
        if (tk < 0) {
          err = (int)tk;
          tk = sigops_server_submit(S, schemes, sigs, aux, msgs, msg_off, 0);  // placeholder, completes at once
        }
        tickets[r].store(tk, std::memory_order_release);
      }
    });
  for (auto& t : clients) t.join();
  collector.join();
  return err.load();
}

}  // extern "C"
