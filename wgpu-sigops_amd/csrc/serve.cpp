// serve.cpp -- continuous-batching request server over the mixed-batch engine.
//
// The reference has no serving layer: every call builds a device, compiles, runs one batch and
// destroys the device (src/gpu.rs:5-35,167). Here requests of any size arrive while others run:
//   * submit() copies a request into a host queue and returns a ticket at once;
//   * the dispatcher packs queued requests FIFO into the next free slot (one sigops_ctx with its own
//     stream, scratch, device buffers and captured graph), as long as the batch stays within
//     max_batch items and the device bytes of all batches in flight stay within the HBM budget;
//     it launches as soon as a slot is free, so arrivals during a launch go out with the next one
//     and consecutive batches overlap on the device;
//   * the completion thread polls the slots' events, so batches may finish out of order, scatters
//     results back to their requests, timestamps them and frees the slot and the budget.
// Every item is processed by the same kernels as a standalone batch, so a request's results do not
// depend on what it was batched with.
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
  bool done = false;
  int err = 0;
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

extern "C" {

sigops_server* sigops_server_create(const sigops_server_config* cfg_in) {
  sigops_server_config cfg{};
  if (cfg_in) cfg = *cfg_in;
  if (!cfg.slots) cfg.slots = 2;
  if (!cfg.hbm_budget) cfg.hbm_budget = 1ull << 30;
  if (!cfg.max_batch) cfg.max_batch = 1u << 20;
  auto* S = new (std::nothrow) sigops_server();
  if (!S) return nullptr;
  S->cfg = cfg;
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
    r->schemes.assign(schemes, schemes + n);
    r->sigs.assign(sigs, sigs + n * 64);
    r->aux.assign(aux, aux + n * 32);
    if (total) r->msgs.assign(msgs + base, msgs + base + total);
    r->off.resize(n + 1);
    for (size_t i = 0; i <= n; i++) r->off[i] = msg_off[i] - base;
  }
  r->bytes = request_bytes(n, r->msgs.size());
  if (r->bytes > S->cfg.hbm_budget) return SIGOPS_E_ARG;
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
  if (r->err) return r->err;
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
