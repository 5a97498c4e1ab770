// BabyBear and KoalaBear instantiations of the pass kernels.
#include <mutex>

#include "pass_kernels.hpp"

namespace zntt {

template <class F>
static hipError_t launch_f31_t(int field, int L, uint32_t *data, uint32_t *scratch, const uint32_t *tw,
                               const uint32_t *coset, const uint32_t *scale_dev, bool tw_one, hipStream_t stream) {
  static std::once_flag once;
  static hipError_t configured = hipSuccess;
  std::call_once(once, [] { configured = configure_pass_lds<F>(4 * (1 << 12)); });
  if (configured != hipSuccess) return configured;
  return run_passes_t<F>(field, L, data, scratch, tw, coset, scale_dev, tw_one, stream);
}

hipError_t launch_transform_f31(int field, int L, uint32_t *data, uint32_t *scratch, const uint32_t *tw,
                                const uint32_t *coset, const uint32_t *scale_dev, bool tw_one, hipStream_t stream) {
  if (field == kBb) return launch_f31_t<BbT>(field, L, data, scratch, tw, coset, scale_dev, tw_one, stream);
  if (field == kKb) return launch_f31_t<KbT>(field, L, data, scratch, tw, coset, scale_dev, tw_one, stream);
  return hipErrorInvalidValue;
}

hipError_t launch_whole_f31(int field, int K, const BatchXform *xf, const BatchEntry *ent, uint32_t n,
                            uint32_t *counter, const uint32_t *tw_all, hipStream_t s) {
  if (field == kBb) return launch_whole_t<BbT>(K, xf, ent, n, counter, tw_all, s);
  if (field == kKb) return launch_whole_t<KbT>(K, xf, ent, n, counter, tw_all, s);
  return hipErrorInvalidValue;
}

hipError_t launch_level_r8_f31(int field, int K, bool odd, bool zero_pad, int tile_log, const BatchXform *xf,
                               const BatchEntry *ent, uint32_t nent, uint32_t nitems, uint32_t uniform,
                               uint32_t *counter, int grid, hipStream_t s) {
  if (field == kBb)
    return launch_level_t<BbT, 10>(K, odd, zero_pad, tile_log, xf, ent, nent, nitems, uniform, counter, grid, s);
  if (field == kKb)
    return launch_level_t<KbT, 10>(K, odd, zero_pad, tile_log, xf, ent, nent, nitems, uniform, counter, grid, s);
  return hipErrorInvalidValue;
}

}  // namespace zntt
