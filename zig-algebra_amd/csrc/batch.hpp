// Device-side descriptors of the batched persistent launch (batch_kernels.hip)
// and its host launcher.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace zntt {

constexpr int kBatchMaxPasses = 4;

// One transform of a batch, as the persistent kernel sees it.
struct BatchXform {
  uint32_t *data;         // user layout, 2^L elements, transformed in place
  uint32_t *scratch;      // dev layout, 2^L elements (multi-pass transforms)
  const uint32_t *tw;     // flat twiddle table (dev layout) covering 2^L
  const uint32_t *coset;  // shift^i for i < 2^L (dev layout) or null
  uint32_t scale[9];      // n^-1 (dev form) for inverse transforms
  int has_scale;
  int L, npass;
  int K[kBatchMaxPasses], lg[kBatchMaxPasses], s0[kBatchMaxPasses];
  uint32_t tiles[kBatchMaxPasses];
};

// Work item: x = transform index, y = pass << 24 | tile.  Items of pass level
// lv are items[level_start[lv] .. level_start[lv+1]) (level_start is a host
// array of levels + 1 offsets); counters[lv] is that level's next-item
// counter (device memory, cleared by the first launch).  Tiles hold
// 2^tile_log elements; the workgroup has 2^tile_log / 8 threads; each level's
// grid is min(grid, items in the level).
hipError_t launch_batch(int field, const BatchXform *xf, const uint2 *items, const uint32_t *level_start,
                        int levels, uint32_t *counters, int grid, int tile_log, hipStream_t s);
hipError_t configure_batch_lds();

}  // namespace zntt
