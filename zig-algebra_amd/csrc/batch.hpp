// Device-side descriptors of the batched persistent launches
// (batch_kernels.hip) and their host launchers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace zntt {

constexpr int kBatchMaxPasses = 4;

// One transform of a batch, as the persistent kernel sees it.
struct BatchXform {
  const uint32_t *src;    // first pass input, user layout; elements at or past src_len read as zero
  uint32_t *data;         // user layout, 2^L elements: the result (may equal src)
  uint32_t *scratch;      // dev layout, 2^L elements (multi-pass transforms)
  const uint32_t *tw;     // flat twiddle table (dev layout) covering 2^L
  const uint32_t *coset;  // shift^i for i < src_len (dev layout) or null
  uint64_t src_len;
  uint32_t scale[9];      // n^-1 (dev form) for inverse transforms
  int has_scale;
  int L, npass;
  int K[kBatchMaxPasses], lg[kBatchMaxPasses], s0[kBatchMaxPasses];
  uint32_t tiles[kBatchMaxPasses];
};

// One (transform, pass) of a launch.  Its tiles are the launch's work items
// first .. first + tiles[pass] - 1; entries are listed with `first` strictly
// increasing from 0.
struct BatchEntry {
  uint32_t xf;
  uint32_t pass;
  uint32_t first;
  uint32_t pad;
};

// One persistent launch over `nitems` tiles described by `ent[0..nent)`.
// `uniform_tiles`: the tile count every entry has, when they are all equal
// (entry i then starts at item i * uniform_tiles), else 0.  counter[0..2)
// (next item, finished workgroups) must be zero when the launch starts; the
// launch's last workgroup leaves them zero again.  Tiles hold 2^tile_log elements; the
// workgroup has 2^tile_log / 8 threads; the grid is min(grid, nitems)
// workgroups.
hipError_t launch_batch_level(int field, const BatchXform *xf, const BatchEntry *ent, uint32_t nent,
                              uint32_t nitems, uint32_t uniform_tiles, uint32_t *counter, int grid, int tile_log,
                              hipStream_t s);
// One persistent launch over n transforms that each fit one tile of 2^K
// elements (K = their log size, 9..11 for BN254 Fr, 9..12 for the others):
// transform ent[i].xf is item i.  counter[0..2) as for launch_batch_level.
// `tw_all`: the twiddle table every one of them uses, or null; given it, the
// small fields keep it in LDS for the whole launch.
hipError_t launch_whole_fr(int K, const BatchXform *xf, const BatchEntry *ent, uint32_t n, uint32_t *counter,
                           const uint32_t *tw_all, hipStream_t s);
hipError_t launch_whole_gl(int K, const BatchXform *xf, const BatchEntry *ent, uint32_t n, uint32_t *counter,
                           const uint32_t *tw_all, hipStream_t s);
hipError_t launch_whole_f31(int field, int K, const BatchXform *xf, const BatchEntry *ent, uint32_t n,
                            uint32_t *counter, const uint32_t *tw_all, hipStream_t s);
inline int whole_max_log(int field) { return field == 1 ? 11 : 12; }
constexpr int kWholeMinLog = 9;

// The register-round level kernel (level_r8_kernel): one persistent launch
// over tiles of 2^tile_log elements whose passes all have K stages and first
// stages of one parity.  zero_pad: some first pass of the launch reads a
// source shorter than its transform (src_len < 2^L, zero padding); the
// launch then uses level_r8_pad_kernel, which skips the products of the
// padding.  level_r8_ok: whether a pass of K stages with that tile can take
// it (K in 5..kLevelMaxK[field], 2^tile_log / N threads in 64..512, N = 4
// values a thread for BN254 Fr, 8 for the others).
hipError_t launch_batch_level_r8(int field, int K, bool odd, bool zero_pad, int tile_log, const BatchXform *xf,
                                 const BatchEntry *ent, uint32_t nent, uint32_t nitems, uint32_t uniform_tiles,
                                 uint32_t *counter, int grid, hipStream_t s);
hipError_t launch_level_r8_fr(int K, bool odd, bool zero_pad, int tile_log, const BatchXform *xf,
                              const BatchEntry *ent, uint32_t nent, uint32_t nitems, uint32_t uniform_tiles,
                              uint32_t *counter, int grid, hipStream_t s);
hipError_t launch_level_r8_gl(int K, bool odd, bool zero_pad, int tile_log, const BatchXform *xf,
                              const BatchEntry *ent, uint32_t nent, uint32_t nitems, uint32_t uniform_tiles,
                              uint32_t *counter, int grid, hipStream_t s);
hipError_t launch_level_r8_f31(int field, int K, bool odd, bool zero_pad, int tile_log, const BatchXform *xf,
                               const BatchEntry *ent, uint32_t nent, uint32_t nitems, uint32_t uniform_tiles,
                               uint32_t *counter, int grid, hipStream_t s);
inline bool level_r8_ok(int field, int K, int tile_log) {
  const int kmax = field == 1 ? 9 : 10, n = field == 1 ? 4 : 8;
  const int threads = (1 << tile_log) / n;
  return K >= 5 && K <= kmax && K <= tile_log && threads >= 64 && threads <= 512;
}

hipError_t configure_batch_lds();

// out[i] for i < n: element i of the counter-based stream (seed, stream_id),
// user layout.
hipError_t launch_fill_counter(int field, uint32_t *out, size_t n, uint64_t seed, uint64_t stream_id,
                               hipStream_t s);

}  // namespace zntt
