// Goldilocks instantiation of the pass kernels.
#include "pass_kernels.hpp"

namespace zntt {

hipError_t launch_transform_gl(int L, uint32_t *data, uint32_t *scratch, const uint32_t *tw, const uint32_t *coset,
                               const uint32_t *scale_dev, hipStream_t stream) {
  static bool configured = false;
  if (!configured) {
    hipError_t e = configure_pass_lds<GlT>(2 * 4 * (1 << 12));
    if (e != hipSuccess) return e;
    configured = true;
  }
  return run_passes_t<GlT>(kGl, L, data, scratch, tw, coset, scale_dev, stream);
}

}  // namespace zntt
