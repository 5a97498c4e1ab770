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

hipError_t launch_pass_fr_padded(const PassArgs &a, hipStream_t s) {
  constexpr int N = round_elems<FrT>();
  constexpr int kLds = 9 * 4 * (1 << 11);
  static std::once_flag once;
  static hipError_t configured = hipSuccess;
  std::call_once(once, [] {
#define ZNTT_FRPA(k, o)                                                                                       \
  if (configured == hipSuccess)                                                                             \
    configured = hipFuncSetAttribute(reinterpret_cast<const void *>(&ntt_pass_r8_kernel<FrT, k, o, N, 0, true>), \
                                     hipFuncAttributeMaxDynamicSharedMemorySize, kLds);
#define ZNTT_FRPB(k) ZNTT_FRPA(k, false) ZNTT_FRPA(k, true)
    ZNTT_FRPB(5) ZNTT_FRPB(6) ZNTT_FRPB(7) ZNTT_FRPB(8) ZNTT_FRPB(9) ZNTT_FRPB(10) ZNTT_FRPB(11) ZNTT_FRPB(12)
#undef ZNTT_FRPB
#undef ZNTT_FRPA
  });
  if (configured != hipSuccess) return configured;
  const size_t tile = (size_t)1 << (a.K + a.log_g);
  if (a.K < 5 || a.K > 12 || tile < 512 || tile / N > 512 || a.K + a.log_g > a.L) return hipErrorInvalidValue;
  const dim3 grid((unsigned)(((size_t)1 << a.L) / tile)), block((unsigned)(tile / N));
  const size_t lds = (size_t)9 * 4 * tile;
  switch (a.K * 2 + ((a.s0 & 1) ? 1 : 0)) {
#define ZNTT_FRP(k)                                                                                        \
  case 2 * k: hipLaunchKernelGGL((ntt_pass_r8_kernel<FrT, k, false, N, 0, true>), grid, block, lds, s, a); break; \
  case 2 * k + 1: hipLaunchKernelGGL((ntt_pass_r8_kernel<FrT, k, true, N, 0, true>), grid, block, lds, s, a); break;
    ZNTT_FRP(5) ZNTT_FRP(6) ZNTT_FRP(7) ZNTT_FRP(8) ZNTT_FRP(9) ZNTT_FRP(10) ZNTT_FRP(11) ZNTT_FRP(12)
#undef ZNTT_FRP
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace zntt
