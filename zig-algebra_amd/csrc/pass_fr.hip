// BN254 Fr instantiation of the pass kernels.
#include <mutex>

#include "pass_kernels.hpp"

namespace zntt {

hipError_t launch_transform_fr(int L, uint32_t *data, uint32_t *scratch, const uint32_t *tw, const uint32_t *coset,
                               const uint32_t *scale_dev, bool tw_one, hipStream_t stream) {
  static std::once_flag once;
  static hipError_t configured = hipSuccess;
  std::call_once(once, [] { configured = configure_pass_lds<FrT>(9 * 4 * (1 << 11)); });
  if (configured != hipSuccess) return configured;
  return run_passes_t<FrT>(kFr, L, data, scratch, tw, coset, scale_dev, tw_one, stream);
}

hipError_t launch_whole_fr(int K, const BatchXform *xf, const BatchEntry *ent, uint32_t n, uint32_t *counter,
                           const uint32_t *tw_all, hipStream_t s) {
  return launch_whole_t<FrT>(K, xf, ent, n, counter, tw_all, s);
}

hipError_t launch_level_r8_fr(int K, bool odd, bool zero_pad, int tile_log, const BatchXform *xf,
                              const BatchEntry *ent, uint32_t nent, uint32_t nitems, uint32_t uniform,
                              uint32_t *counter, int grid, hipStream_t s) {
  return launch_level_t<FrT, 9>(K, odd, zero_pad, tile_log, xf, ent, nent, nitems, uniform, counter, grid, s);
}

}  // namespace zntt
