// Copies between the caller's host buffers and device memory for the
// host-buffer entry points (zntt_ntt, zntt_intt, the twiddled and coset
// calls, zntt_lde).
//
// A copy from pageable memory is staged by the HIP runtime through its own
// page-locked buffers with one CPU thread; on MI355X that runs at about half
// the DMA engine's rate host to device and under a third of it device to
// host.  Large buffers therefore go through a ring of page-locked slots that
// the context owns: a few threads copy one slot between the caller's memory
// and the ring while the DMA engine moves the previous slot, so the transfer
// runs at close to the DMA rate both ways.  Small buffers, and any context
// whose ring cannot be allocated, use the plain copy.
#include <string.h>

#include <algorithm>
#include <condition_variable>
#include <mutex>
#include <thread>
#include <vector>

#include "ctx_internal.hpp"

namespace zntt_detail {

namespace {

constexpr size_t kChunk = (size_t)16 << 20;  // bytes a slot holds
constexpr int kSlots = 4;
constexpr int kWorkers = 3;  // copy threads besides the caller's

// memcpy split over the workers and the calling thread.
class CopyPool {
 public:
  CopyPool() {
    for (int i = 0; i < kWorkers; i++) th_.emplace_back([this, i] { run(i + 1); });
  }
  ~CopyPool() {
    {
      std::lock_guard<std::mutex> lk(mu_);
      stop_ = true;
    }
    cv_.notify_all();
    for (auto &t : th_) t.join();
  }
  void copy(void *dst, const void *src, size_t n) {
    {
      std::lock_guard<std::mutex> lk(mu_);
      dst_ = (uint8_t *)dst;
      src_ = (const uint8_t *)src;
      n_ = n;
      pending_ = kWorkers;
      gen_++;
    }
    cv_.notify_all();
    part(0);
    std::unique_lock<std::mutex> lk(mu_);
    done_.wait(lk, [this] { return pending_ == 0; });
  }

 private:
  // Part i of kWorkers + 1, on 64-byte boundaries.
  void part(int i) {
    const size_t parts = kWorkers + 1;
    const size_t lo = std::min(n_, (n_ * i / parts) & ~(size_t)63);
    const size_t hi = i + 1 == (int)parts ? n_ : std::min(n_, (n_ * (i + 1) / parts) & ~(size_t)63);
    if (hi > lo) memcpy(dst_ + lo, src_ + lo, hi - lo);
  }
  void run(int id) {
    uint64_t seen = 0;
    for (;;) {
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return stop_ || gen_ != seen; });
        if (stop_) return;
        seen = gen_;
      }
      part(id);
      std::lock_guard<std::mutex> lk(mu_);
      if (--pending_ == 0) done_.notify_one();
    }
  }
  std::vector<std::thread> th_;
  std::mutex mu_;
  std::condition_variable cv_, done_;
  uint64_t gen_ = 0;
  int pending_ = 0;
  bool stop_ = false;
  uint8_t *dst_ = nullptr;
  const uint8_t *src_ = nullptr;
  size_t n_ = 0;
};

}  // namespace

struct HostRing {
  uint8_t *buf = nullptr;  // kSlots * kChunk bytes, page-locked
  hipEvent_t ev[kSlots] = {};
  bool busy[kSlots] = {};  // ev[s] follows a transfer of slot s not yet waited for
  CopyPool pool;
  uint8_t *slot(int s) { return buf + (size_t)s * kChunk; }
  // Waits until no transfer uses slot s, from this copy or an earlier one
  // (a call may run several copies before it synchronises its stream).
  hipError_t claim(int s) {
    if (!busy[s]) return hipSuccess;
    const hipError_t e = hipEventSynchronize(ev[s]);
    if (e == hipSuccess) busy[s] = false;
    return e;
  }
  hipError_t mark(int s, hipStream_t stream) {
    const hipError_t e = hipEventRecord(ev[s], stream);
    busy[s] = e == hipSuccess;
    return e;
  }
};

bool ring_wanted(size_t bytes) { return bytes > kChunk; }

HostRing *ring_create() {
  HostRing *r = nullptr;
  try {
    r = new HostRing();  // starts the copy threads
  } catch (...) {
    return nullptr;  // no thread or memory for them
  }
  hipError_t e = hipHostMalloc((void **)&r->buf, (size_t)kSlots * kChunk, hipHostMallocDefault);
  for (int s = 0; s < kSlots && e == hipSuccess; s++) e = hipEventCreateWithFlags(&r->ev[s], hipEventDisableTiming);
  if (e != hipSuccess) {
    hipGetLastError();
    ring_destroy(r);
    return nullptr;
  }
  return r;
}

// No copy may be using the ring.
void ring_destroy(HostRing *r) {
  if (!r) return;
  for (auto &ev : r->ev)
    if (ev) hipEventDestroy(ev);
  if (r->buf) hipHostFree(r->buf);
  delete r;
}

int ring_ensure(zntt_ctx *c) {
  if (c->ring) return ZNTT_OK;
  if (c->ring_failed) return ZNTT_OUT_OF_MEMORY;
  c->ring = ring_create();
  c->ring_failed = c->ring == nullptr;  // the plain copy serves until the next trim
  return c->ring ? ZNTT_OK : ZNTT_OUT_OF_MEMORY;
}

// The caller has synchronised the context's stream (no copy uses the ring).
void ring_free(zntt_ctx *c) {
  c->ring_failed = false;
  ring_destroy(c->ring);
  c->ring = nullptr;
}

int ring_copy_to_device(HostRing *r, void *dev, const void *host, size_t bytes, hipStream_t stream) {
  const uint8_t *h = (const uint8_t *)host;
  uint8_t *d = (uint8_t *)dev;
  hipError_t e = hipSuccess;
  for (size_t off = 0, i = 0; off < bytes && e == hipSuccess; off += kChunk, i++) {
    const int s = (int)(i % kSlots);
    const size_t len = std::min(kChunk, bytes - off);
    e = r->claim(s);
    if (e != hipSuccess) break;
    r->pool.copy(r->slot(s), h + off, len);
    e = hipMemcpyAsync(d + off, r->slot(s), len, hipMemcpyHostToDevice, stream);
    if (e == hipSuccess) e = r->mark(s, stream);
  }
  return hip_status(e);
}

int ring_copy_to_host(HostRing *r, void *host, const void *dev, size_t bytes, hipStream_t stream) {
  uint8_t *h = (uint8_t *)host;
  const uint8_t *d = (const uint8_t *)dev;
  const size_t chunks = (bytes + kChunk - 1) / kChunk;
  // Download of chunk i into its slot; the stream orders it after the work
  // and uploads enqueued before.
  auto issue = [&](size_t i) {
    const int s = (int)(i % kSlots);
    const size_t off = i * kChunk, len = std::min(kChunk, bytes - off);
    hipError_t e = r->claim(s);
    if (e == hipSuccess) e = hipMemcpyAsync(r->slot(s), d + off, len, hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = r->mark(s, stream);
    return e;
  };
  hipError_t e = hipSuccess;
  for (size_t i = 0; i < std::min(chunks, (size_t)kSlots) && e == hipSuccess; i++) e = issue(i);
  for (size_t i = 0; i < chunks && e == hipSuccess; i++) {
    const int s = (int)(i % kSlots);
    const size_t off = i * kChunk, len = std::min(kChunk, bytes - off);
    e = r->claim(s);  // the slot's download has landed
    if (e != hipSuccess) break;
    r->pool.copy(h + off, r->slot(s), len);
    if (i + kSlots < chunks) e = issue(i + kSlots);
  }
  hipError_t e2 = hipStreamSynchronize(stream);
  return hip_status(e != hipSuccess ? e : e2);
}

int copy_to_device(zntt_ctx *c, void *dev, const void *host, size_t bytes, hipStream_t stream) {
  if (!ring_wanted(bytes) || ring_ensure(c) != ZNTT_OK)
    return hip_status(hipMemcpyAsync(dev, host, bytes, hipMemcpyHostToDevice, stream));
  return ring_copy_to_device(c->ring, dev, host, bytes, stream);
}

int copy_to_host(zntt_ctx *c, void *host, const void *dev, size_t bytes, hipStream_t stream) {
  if (!ring_wanted(bytes) || ring_ensure(c) != ZNTT_OK) {
    hipError_t e = hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    return hip_status(e);
  }
  return ring_copy_to_host(c->ring, host, dev, bytes, stream);
}

}  // namespace zntt_detail
