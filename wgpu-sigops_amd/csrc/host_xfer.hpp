// host_xfer.hpp -- copies between caller buffers in host memory and device buffers, through pinned
// staging memory filled and drained by a pool of host threads.
//
// The host entry points take ordinary (pageable) caller memory, as the reference's API takes
// Rust vectors (src/secp256k1_ecdsa.rs:61-213 copies them into wgpu buffers on every call). DMA
// engines read and write pinned memory only, so a pageable copy always passes through a pinned
// buffer; the runtime's own staging does that with one CPU thread. Here the inputs are packed into
// a pinned buffer owned by the context by every pool thread at once, in parts, and each part's
// DMA is queued as soon as it is packed, so packing overlaps the transfer. Results come back the
// same way: each part is unpacked by the pool as soon as its DMA has completed.
// Host code only; no GPU code here.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

namespace sigops_host {

// Fork-join pool: run(fn) calls fn(t) once for every t in [0, size()), t = 0 on the calling
// thread, and returns when all calls have returned. One caller at a time.
class Pool {
 public:
  explicit Pool(unsigned n) : n_(n ? n : 1) {
    for (unsigned t = 1; t < n_; t++) workers_.emplace_back([this, t] { loop(t); });
  }
  ~Pool() {
    {
      std::lock_guard<std::mutex> l(m_);
      stop_ = true;
      gen_++;
    }
    cv_.notify_all();
    for (auto& w : workers_) w.join();
  }
  Pool(const Pool&) = delete;
  Pool& operator=(const Pool&) = delete;
  unsigned size() const { return n_; }

  void run(const std::function<void(unsigned)>& fn) {
    if (n_ == 1) {
      fn(0);
      return;
    }
    {
      std::lock_guard<std::mutex> l(m_);
      fn_ = &fn;
      pending_ = n_ - 1;
      gen_++;
    }
    cv_.notify_all();
    fn(0);
    std::unique_lock<std::mutex> l(m_);
    done_.wait(l, [&] { return pending_ == 0; });
    fn_ = nullptr;
  }

  // fn(begin, end) over [0, count) in one contiguous range per thread (fewer threads for small
  // counts: at least `grain` elements each)
  void for_range(size_t count, const std::function<void(size_t, size_t)>& fn, size_t grain = 1 << 16) {
    if (count == 0) return;
    const size_t parts = std::min<size_t>(n_, std::max<size_t>(1, count / grain));
    if (parts == 1) {
      fn(0, count);
      return;
    }
    run([&](unsigned t) {
      if (t >= parts) return;
      fn(count * t / parts, count * (t + 1) / parts);
    });
  }

 private:
  void loop(unsigned t) {
    uint64_t seen = 0;
    for (;;) {
      const std::function<void(unsigned)>* f;
      {
        std::unique_lock<std::mutex> l(m_);
        cv_.wait(l, [&] { return gen_ != seen; });
        seen = gen_;
        if (stop_) return;
        f = fn_;
      }
      (*f)(t);
      {
        std::lock_guard<std::mutex> l(m_);
        if (--pending_ == 0) done_.notify_one();
      }
    }
  }

  unsigned n_;
  std::vector<std::thread> workers_;
  std::mutex m_;
  std::condition_variable cv_, done_;
  const std::function<void(unsigned)>* fn_ = nullptr;
  unsigned pending_ = 0;
  uint64_t gen_ = 0;
  bool stop_ = false;
};

// threads for host-side copies: SIGOPS_HOST_THREADS, else the hardware threads, at most 8. On a
// 16-thread host share, 4 and 8 threads pack a 2^20 batch equally fast and 16 are slower (they
// compete with the runtime's own threads).
inline unsigned host_threads() {
  if (const char* e = getenv("SIGOPS_HOST_THREADS")) {
    const long v = strtol(e, nullptr, 10);
    if (v >= 1 && v <= 256) return (unsigned)v;
  }
  const unsigned hw = std::thread::hardware_concurrency();
  return std::max(1u, std::min(8u, hw));
}

struct Pinned {
  uint8_t* p = nullptr;
  size_t cap = 0;
};

inline void pinned_free(Pinned& b) {
  if (b.p) (void)hipHostFree(b.p);
  b.p = nullptr;
  b.cap = 0;
}

inline hipError_t pinned_ensure(Pinned& b, size_t bytes) {
  if (b.cap >= bytes) return hipSuccess;
  pinned_free(b);
  const size_t want = (std::max<size_t>(bytes, 1) + (2u << 20) - 1) & ~(size_t)((2u << 20) - 1);
  hipError_t e = hipHostMalloc((void**)&b.p, want, hipHostMallocDefault);
  if (e != hipSuccess) {
    b.p = nullptr;
    return e;
  }
  b.cap = want;
  return hipSuccess;
}

// an upload: `bytes` from host to dev; with rebase != 0 the bytes are u64 values and rebase is
// subtracted from each as it is packed (message offsets of a sub-range of a caller's array)
struct Up {
  void* dev;
  const void* host;
  size_t bytes;
  uint64_t rebase = 0;
};

struct Down {
  void* host;
  const void* dev;
  size_t bytes;
};

constexpr size_t XFER_PART = 8u << 20;  // bytes per part
constexpr size_t XFER_ALIGN = 4096;     // segment alignment inside the pinned buffer

inline size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// positions of the segments in the staging buffer; returns the total
template <class S>
size_t layout(const std::vector<S>& segs, std::vector<size_t>& pos) {
  pos.resize(segs.size());
  size_t at = 0;
  for (size_t k = 0; k < segs.size(); k++) {
    pos[k] = at;
    at = align_up(at + segs[k].bytes, XFER_ALIGN);
  }
  return at;
}

// packs [lo, hi) of the staging image (a multiple of 8 apart, lo a multiple of 8)
inline void pack_range(const std::vector<Up>& segs, const std::vector<size_t>& pos, uint8_t* stage, size_t lo,
                       size_t hi) {
  for (size_t k = 0; k < segs.size(); k++) {
    const size_t a = std::max(lo, pos[k]), b = std::min(hi, pos[k] + segs[k].bytes);
    if (a >= b) continue;
    const uint8_t* src = (const uint8_t*)segs[k].host + (a - pos[k]);
    if (segs[k].rebase == 0) {
      memcpy(stage + a, src, b - a);
    } else {
      // u64 elements; a - pos[k] and b - pos[k] are multiples of 8 (segment sizes are)
      const uint64_t* s = (const uint64_t*)src;
      uint64_t* d = (uint64_t*)(stage + a);
      const size_t m = (b - a) / 8;
      for (size_t i = 0; i < m; i++) d[i] = s[i] - segs[k].rebase;
    }
  }
}

// queues the uploads on stream s. The staging buffer must stay untouched until s has passed them.
inline hipError_t upload(Pool& pool, Pinned& stage, const std::vector<Up>& segs, hipStream_t s) {
  std::vector<size_t> pos;
  const size_t total = layout(segs, pos);
  if (total == 0) return hipSuccess;
  // the staging buffer is free once the stream has drained (an earlier call that stopped on an
  // error may have left a transfer from it queued)
  hipError_t e = hipStreamSynchronize(s);
  if (e != hipSuccess) return e;
  e = pinned_ensure(stage, total);
  if (e != hipSuccess) return e;
  const unsigned nt = pool.size();
  for (size_t lo = 0; lo < total; lo += XFER_PART) {
    const size_t hi = std::min(total, lo + XFER_PART);
    const size_t span = hi - lo;
    if (span >= (256u << 10) && nt > 1) {
      pool.run([&](unsigned t) {
        const size_t a = lo + (span * t / nt) / 64 * 64;
        const size_t b = t + 1 == nt ? hi : lo + (span * (t + 1) / nt) / 64 * 64;
        pack_range(segs, pos, stage.p, a, b);
      });
    } else {
      pack_range(segs, pos, stage.p, lo, hi);
    }
    for (size_t k = 0; k < segs.size(); k++) {
      const size_t a = std::max(lo, pos[k]), b = std::min(hi, pos[k] + segs[k].bytes);
      if (a >= b) continue;
      e = hipMemcpyAsync((uint8_t*)segs[k].dev + (a - pos[k]), stage.p + a, b - a, hipMemcpyHostToDevice, s);
      if (e != hipSuccess) return e;
    }
  }
  return hipSuccess;
}

// queues the downloads on stream s and unpacks each part into the caller's memory once its DMA
// has completed; returns when everything is in place. evs: events reused across calls.
inline hipError_t download(Pool& pool, Pinned& stage, const std::vector<Down>& segs, hipStream_t s,
                           std::vector<hipEvent_t>& evs) {
  std::vector<size_t> pos;
  const size_t total = layout(segs, pos);
  if (total == 0) return hipStreamSynchronize(s);
  hipError_t e = pinned_ensure(stage, total);
  if (e != hipSuccess) return e;
  const size_t nparts = (total + XFER_PART - 1) / XFER_PART;
  while (evs.size() < nparts) {
    hipEvent_t ev;
    if ((e = hipEventCreateWithFlags(&ev, hipEventDisableTiming)) != hipSuccess) return e;
    evs.push_back(ev);
  }
  for (size_t p = 0; p < nparts; p++) {
    const size_t lo = p * XFER_PART, hi = std::min(total, lo + XFER_PART);
    for (size_t k = 0; k < segs.size(); k++) {
      const size_t a = std::max(lo, pos[k]), b = std::min(hi, pos[k] + segs[k].bytes);
      if (a >= b) continue;
      e = hipMemcpyAsync(stage.p + a, (const uint8_t*)segs[k].dev + (a - pos[k]), b - a, hipMemcpyDeviceToHost, s);
      if (e != hipSuccess) return e;
    }
    if ((e = hipEventRecord(evs[p], s)) != hipSuccess) return e;
  }
  const unsigned nt = pool.size();
  for (size_t p = 0; p < nparts; p++) {
    if ((e = hipEventSynchronize(evs[p])) != hipSuccess) return e;
    const size_t lo = p * XFER_PART, hi = std::min(total, lo + XFER_PART);
    auto unpack = [&](size_t a0, size_t b0) {
      for (size_t k = 0; k < segs.size(); k++) {
        const size_t a = std::max(a0, pos[k]), b = std::min(b0, pos[k] + segs[k].bytes);
        if (a < b) memcpy((uint8_t*)segs[k].host + (a - pos[k]), stage.p + a, b - a);
      }
    };
    const size_t span = hi - lo;
    if (span >= (256u << 10) && nt > 1) {
      pool.run([&](unsigned t) {
        const size_t a = lo + (span * t / nt) / 64 * 64;
        const size_t b = t + 1 == nt ? hi : lo + (span * (t + 1) / nt) / 64 * 64;
        unpack(a, b);
      });
    } else {
      unpack(lo, hi);
    }
  }
  // the stream may hold later work of the caller's; everything of ours has completed
  return hipSuccess;
}

}  // namespace sigops_host
