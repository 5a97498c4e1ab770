// Host-only stand-in for the HIP runtime and the library's kernel launchers,
// for CPU tests of the C ABI's host logic (context, twiddle cache, batches,
// serving loop) under AddressSanitizer on a machine without a GPU.
//
// "Device memory" is host heap memory.  Every launcher reads (and where the
// kernel writes, writes) the first and last byte of every buffer the real
// kernel would touch, computed from its arguments, so that a launch over a
// buffer the host side has already freed is an ASan heap-use-after-free at
// the launch.  The kernels' arithmetic is not emulated: values are
// meaningless, lifetimes and extents are what is checked.  Stream capture
// records launches into the graph and hipGraphLaunch replays them, so a
// replayed batch touches its buffers at replay time.
#include <hip/hip_runtime_api.h>
#include <stdlib.h>
#include <string.h>

#include <atomic>
#include <chrono>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

#include "batch.hpp"
#include "kernels.hpp"
#include "zntt_gpu.h"

namespace {

std::mutex g_mu;
// While set, events report "not ready": work a test has enqueued stays in
// flight (its launches have touched their buffers; completion is withheld).
std::atomic<int> g_hold{0};
// Calls to the runtime's allocators and frees, and stream waits on events.
std::atomic<long> g_allocs{0}, g_waits{0};
hipStream_t g_capturing = nullptr;
std::vector<std::function<void()>> *g_recording = nullptr;

struct FakeGraph {
  std::vector<std::function<void()>> ops;
};

// Runs a launch now, or records it when its stream is being captured.
hipError_t run_or_record(hipStream_t s, std::function<void()> fn) {
  {
    std::lock_guard<std::mutex> lk(g_mu);
    (void)s;  // branch streams forked from the captured stream are captured too
    if (g_capturing) {
      g_recording->push_back(std::move(fn));
      return hipSuccess;
    }
  }
  fn();
  return hipSuccess;
}

volatile uint8_t g_sink;
void rd(const void *p, size_t bytes) {
  if (!p || !bytes) return;
  const volatile uint8_t *b = (const volatile uint8_t *)p;
  g_sink = b[0];
  g_sink = b[bytes - 1];
}
void wr(void *p, size_t bytes) {
  if (!p || !bytes) return;
  volatile uint8_t *b = (volatile uint8_t *)p;
  b[0] = b[0];
  b[bytes - 1] = b[bytes - 1];
}
size_t ub(int field) { return 4 * (size_t)zntt::user_words(field); }
size_t db(int field) { return 4 * (size_t)zntt::dev_words(field); }

// What one pass of a batch transform reads and writes (BatchXform fields as
// the level kernel uses them).
void touch_pass(int field, const zntt::BatchXform &X, int p) {
  const size_t n = (size_t)1 << X.L;
  const bool first = p == 0, last = p == X.npass - 1;
  if (first)
    rd(X.src, (size_t)X.src_len * ub(field));
  else
    rd(X.scratch, n * db(field));
  if (last)
    wr(X.data, n * ub(field));
  else
    wr(X.scratch, n * db(field));
  rd(X.tw, (n - 1) * db(field));
  if (first && X.coset) rd(X.coset, (size_t)X.src_len * db(field));
}

}  // namespace

// ---------------------------------------------------------------------------
// HIP runtime
// ---------------------------------------------------------------------------
extern "C" {

hipError_t hipGetDeviceCount(int *count) {
  *count = 1;
  return hipSuccess;
}
hipError_t hipSetDevice(int) { return hipSuccess; }
hipError_t hipGetDevice(int *d) {
  *d = 0;
  return hipSuccess;
}
hipError_t hipDeviceGetAttribute(int *v, hipDeviceAttribute_t, int) {
  *v = 256;
  return hipSuccess;
}
hipError_t hipGetLastError(void) { return hipSuccess; }
hipError_t hipDeviceSynchronize(void) { return hipSuccess; }

hipError_t hipMalloc(void **p, size_t bytes) {
  g_allocs++;
  *p = aligned_alloc(256, (bytes + 255) & ~(size_t)255);
  return *p ? hipSuccess : hipErrorOutOfMemory;
}
hipError_t hipFree(void *p) {
  g_allocs++;
  free(p);
  return hipSuccess;
}
hipError_t hipHostMalloc(void **p, size_t bytes, unsigned int) { return hipMalloc(p, bytes); }
hipError_t hipMallocAsync(void **p, size_t bytes, hipStream_t) { return hipMalloc(p, bytes); }
hipError_t hipHostFree(void *p) { return hipFree(p); }

hipError_t hipMemcpy(void *dst, const void *src, size_t bytes, hipMemcpyKind) {
  memcpy(dst, src, bytes);
  return hipSuccess;
}
hipError_t hipMemcpyAsync(void *dst, const void *src, size_t bytes, hipMemcpyKind, hipStream_t s) {
  return run_or_record(s, [=] { memcpy(dst, src, bytes); });
}
hipError_t hipMemsetAsync(void *dst, int v, size_t bytes, hipStream_t s) {
  return run_or_record(s, [=] { memset(dst, v, bytes); });
}

hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned int) {
  *s = (hipStream_t)malloc(64);
  return hipSuccess;
}
hipError_t hipStreamDestroy(hipStream_t s) {
  free((void *)s);
  return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }

hipError_t hipEventCreate(hipEvent_t *e) {
  *e = (hipEvent_t)malloc(16);
  return hipSuccess;
}
hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned int) { return hipEventCreate(e); }
hipError_t hipEventDestroy(hipEvent_t e) {
  free((void *)e);
  return hipSuccess;
}
hipError_t hipEventRecord(hipEvent_t, hipStream_t) { return hipSuccess; }
hipError_t hipEventQuery(hipEvent_t) { return g_hold ? hipErrorNotReady : hipSuccess; }
hipError_t hipEventSynchronize(hipEvent_t) {
  while (g_hold) std::this_thread::sleep_for(std::chrono::milliseconds(1));
  return hipSuccess;
}
hipError_t hipEventElapsedTime(float *ms, hipEvent_t, hipEvent_t) {
  *ms = 0.01f;
  return hipSuccess;
}
hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned int) {
  g_waits++;
  return hipSuccess;
}

hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode) {
  std::lock_guard<std::mutex> lk(g_mu);
  g_capturing = s;
  g_recording = new std::vector<std::function<void()>>();
  return hipSuccess;
}
hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t *g) {
  std::lock_guard<std::mutex> lk(g_mu);
  auto *fg = new FakeGraph();
  fg->ops = std::move(*g_recording);
  delete g_recording;
  g_recording = nullptr;
  g_capturing = nullptr;
  *g = (hipGraph_t)fg;
  return hipSuccess;
}
hipError_t hipGraphInstantiate(hipGraphExec_t *e, hipGraph_t g, hipGraphNode_t *, char *, size_t) {
  auto *fg = new FakeGraph(*(FakeGraph *)g);
  *e = (hipGraphExec_t)fg;
  return hipSuccess;
}
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t) {
  for (auto &op : ((FakeGraph *)e)->ops) op();
  return hipSuccess;
}
hipError_t hipGraphDestroy(hipGraph_t g) {
  delete (FakeGraph *)g;
  return hipSuccess;
}
hipError_t hipGraphExecDestroy(hipGraphExec_t e) {
  delete (FakeGraph *)e;
  return hipSuccess;
}
void fake_hip_hold(int on) { g_hold = on; }
long fake_hip_alloc_calls() { return g_allocs; }
long fake_hip_stream_waits() { return g_waits; }

hipError_t hipGetSymbolAddress(void **p, const void *sym) {
  *p = (void *)sym;
  return hipSuccess;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Kernel launchers
// ---------------------------------------------------------------------------
namespace zntt {

hipError_t launch_transform(int field, int L, uint32_t *data, uint32_t *scratch, const uint32_t *tw,
                            const uint32_t *coset, const uint32_t *scale_dev, bool, hipStream_t s) {
  return run_or_record(s, [=] {
    const size_t n = (size_t)1 << L;
    wr(data, n * ub(field));
    rd(tw, (n - 1) * db(field));
    if (coset) rd(coset, n * db(field));
    if (scale_dev) rd(scale_dev, db(field));
    if (plan_passes(field, L).passes > 1) wr(scratch, n * db(field));
  });
}

hipError_t launch_build_table(int field, int L, uint32_t *flat, const uint32_t *pows, const uint32_t *one_dev,
                              hipStream_t s) {
  return run_or_record(s, [=] {
    memset(flat, 0, ((size_t)1 << L) * db(field));
    rd(pows, (size_t)(L > 0 ? L : 1) * db(field));
    rd(one_dev, db(field));
  });
}
hipError_t launch_powers(int field, size_t count, uint32_t *out, const uint32_t *pows, int npows,
                         const uint32_t *one_dev, hipStream_t s) {
  return run_or_record(s, [=] {
    memset(out, 0, count * db(field));
    rd(pows, (size_t)npows * db(field));
    rd(one_dev, db(field));
  });
}
hipError_t launch_table_from_user(int field, size_t count, uint32_t *dev, const uint32_t *user, hipStream_t s) {
  return run_or_record(s, [=] {
    rd(user, count * ub(field));
    memset(dev, 0, count * db(field));
  });
}
hipError_t launch_table_to_user(int field, size_t count, uint32_t *user, const uint32_t *dev, hipStream_t s) {
  return run_or_record(s, [=] {
    rd(dev, count * db(field));
    memset(user, 0, count * ub(field));
  });
}
hipError_t launch_inverse_table(int field, int L, uint32_t *inv, const uint32_t *fwd, hipStream_t s) {
  return run_or_record(s, [=] {
    rd(fwd, ((size_t)1 << L) * db(field));
    memset(inv, 0, ((size_t)1 << L) * db(field));
  });
}
hipError_t launch_bit_reverse(int field, int L, uint32_t *data, hipStream_t s) {
  return run_or_record(s, [=] { wr(data, ((size_t)1 << L) * ub(field)); });
}
hipError_t launch_pointwise_mul(int field, uint32_t *a, const uint32_t *b, size_t n, hipStream_t s) {
  return run_or_record(s, [=] {
    wr(a, n * ub(field));
    rd(b, n * ub(field));
  });
}
hipError_t launch_field_op(int field, int, size_t n, const uint32_t *a, const uint32_t *b, uint32_t *out,
                           hipStream_t s) {
  return run_or_record(s, [=] {
    rd(a, n * ub(field));
    rd(b, n * ub(field));
    memset(out, 0, n * ub(field));
  });
}
hipError_t probe_peaks(double *mad, double *copy, hipStream_t) {
  *mad = 1;
  *copy = 1;
  return hipSuccess;
}
hipError_t configure_batch_lds() { return hipSuccess; }
hipError_t noalloc_scratch_init(int) { return hipSuccess; }
void *noalloc_scratch_ptr(int) {
  static std::vector<uint8_t> buf(ZNTT_NOALLOC_SCRATCH_BYTES);
  return buf.data();
}

hipError_t launch_batch_level(int field, const BatchXform *xf, const BatchEntry *ent, uint32_t nent,
                              uint32_t nitems, uint32_t, uint32_t *counter, int, int, hipStream_t s) {
  return run_or_record(s, [=] {
    uint32_t items = 0;
    for (uint32_t i = 0; i < nent; i++) {
      const BatchXform &X = xf[ent[i].xf];
      touch_pass(field, X, (int)ent[i].pass);
      items += X.tiles[ent[i].pass];
    }
    if (items != nitems) abort();  // the work list must cover exactly the launch's items
    wr(counter, 2 * sizeof(uint32_t));
  });
}

hipError_t launch_batch_level_r8(int field, int, bool, bool, int, const BatchXform *xf, const BatchEntry *ent,
                                 uint32_t nent, uint32_t nitems, uint32_t uniform, uint32_t *counter, int grid,
                                 hipStream_t s) {
  return launch_batch_level(field, xf, ent, nent, nitems, uniform, counter, grid, 0, s);
}

static hipError_t whole(int field, const BatchXform *xf, const BatchEntry *ent, uint32_t n, uint32_t *counter,
                        const uint32_t *tw_all, hipStream_t s) {
  return run_or_record(s, [=] {
    for (uint32_t i = 0; i < n; i++) touch_pass(field, xf[ent[i].xf], 0);
    if (tw_all) rd(tw_all, db(field));
    wr(counter, 2 * sizeof(uint32_t));
  });
}
hipError_t launch_whole_fr(int, const BatchXform *xf, const BatchEntry *ent, uint32_t n, uint32_t *counter,
                           const uint32_t *tw_all, hipStream_t s) {
  return whole(kFr, xf, ent, n, counter, tw_all, s);
}
hipError_t launch_whole_gl(int, const BatchXform *xf, const BatchEntry *ent, uint32_t n, uint32_t *counter,
                           const uint32_t *tw_all, hipStream_t s) {
  return whole(kGl, xf, ent, n, counter, tw_all, s);
}
hipError_t launch_whole_f31(int field, int, const BatchXform *xf, const BatchEntry *ent, uint32_t n,
                            uint32_t *counter, const uint32_t *tw_all, hipStream_t s) {
  return whole(field, xf, ent, n, counter, tw_all, s);
}
hipError_t launch_fill_counter(int field, uint32_t *out, size_t n, uint64_t, uint64_t, hipStream_t s) {
  return run_or_record(s, [=] { memset(out, 0, n * ub(field)); });
}

}  // namespace zntt
