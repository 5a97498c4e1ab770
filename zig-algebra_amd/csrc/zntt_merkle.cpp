// C ABI of the Merkle commitment (BLAKE3 trees over FRI layers): argument
// checks, the node heap (the caller's buffer or context scratch) and the root
// read-back.  Kernels in merkle_kernels.hip.
#include <hip/hip_runtime.h>

#include "ctx_internal.hpp"

using namespace zntt_detail;

namespace {

// Node and hash buffers are read and written 16 bytes at a time.
bool aligned16(const void *p) { return ((uintptr_t)p & 15) == 0; }

// Builds with `launch` into the caller's nodes or a scratch heap, then copies
// node 1 (the root) to host memory.  Caller holds the context lock.
template <class Launch>
int commit(zntt_ctx *c, int cap_log, void *nodes, void *root, hipStream_t s, Launch &&launch) {
  uint32_t *heap = static_cast<uint32_t *>(nodes);
  if (!heap) {
    const size_t bytes = (size_t)32 << zntt::merkle_heap_log(cap_log, false);
    int st = ensure(c, &c->aux, &c->aux_bytes, bytes);
    if (st) return st;
    heap = static_cast<uint32_t *>(c->aux);
  }
  hipError_t e = launch(heap, nodes != nullptr);
  if (e == hipSuccess) e = hipMemcpyAsync(root, heap + 8, 32, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  return hip_status(e);
}

}  // namespace

extern "C" {

int zntt_merkle_commit(zntt_ctx *c, int field, const void *evals, size_t log_n, void *nodes, void *root,
                       void *stream) {
  if (!c || !field_ok(field) || !evals || !root || !aligned16(evals) || !aligned16(nodes))
    return ZNTT_INVALID_ARGUMENT;
  if (log_n >= 64) return ZNTT_LOG_TOO_LARGE;
  if (log_n == 0) return ZNTT_INVALID_LENGTH;  // a layer needs an antipodal pair
  if (log_n > ZNTT_MAX_LOG_N) return ZNTT_UNSUPPORTED;
  const int leaf_log = (int)log_n - 1;
  CallGuard g(c, (hipStream_t)stream);
  return commit(c, leaf_log, nodes, root, (hipStream_t)stream, [&](uint32_t *heap, bool full) {
    return zntt::launch_merkle_layer(field, evals, leaf_log, heap, full, (hipStream_t)stream);
  });
}

int zntt_merkle_from_hashes(zntt_ctx *c, const void *hashes, size_t count, void *nodes, void *root, void *stream) {
  if (!c || !hashes || !root || !aligned16(hashes) || !aligned16(nodes)) return ZNTT_INVALID_ARGUMENT;
  if (count == 0) return ZNTT_INVALID_LENGTH;  // merkle_tree.zig:148, EmptyLeaves
  int cap_log = 0;
  while (((size_t)1 << cap_log) < count) {
    if (++cap_log > ZNTT_MAX_LOG_N) return ZNTT_UNSUPPORTED;
  }
  CallGuard g(c, (hipStream_t)stream);
  return commit(c, cap_log, nodes, root, (hipStream_t)stream, [&](uint32_t *heap, bool full) {
    return zntt::launch_merkle_from_hashes(hashes, count, cap_log, heap, full, (hipStream_t)stream);
  });
}

}  // extern "C"
