// Mersenne-31 instantiation of the pass kernels, for the eight-lane M31
// transform (Vec8NttM31, libs/field/src/lib.zig:52-137), and its helpers:
// the twiddle table and the lane (de)interleaving.
//
// The reference stores eight transforms interleaved, data[i * 8 + lane], and
// runs the radix-2 DIT DAG of zig-ntt on each lane: per-lane bit reversal,
// then stages s = 1..log_n with twiddles (root^(2^(log_n - s)))^j.  Here the
// lanes are first gathered into eight contiguous transforms (one coalesced
// pass over the data), each runs through the same pass kernels as every other
// field, and a second pass scatters them back.
#include <mutex>

#include "field_m31.hpp"
#include "pass_kernels.hpp"

namespace zntt {
namespace {

// flat[2^(s-1) - 1 + j] = (root^(2^(L-s)))^j for stage s = 1..L, j < 2^(s-1),
// and flat[2^L - 1] = 1; pows[k] = root^(2^k) for k < L.  Entry j of stage s
// is root^(j * 2^(L-s)), the product of pows[L - s + b] over the set bits b
// of j.
__global__ __launch_bounds__(256) void m31_table_kernel(uint32_t *flat, int L, const uint32_t *pows) {
  const uint64_t f = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  const uint64_t n = 1ull << L;
  if (f >= n) return;
  if (f == n - 1) {
    flat[f] = 1;
    return;
  }
  const int s = 64 - __clzll(f + 1);  // 2^(s-1) <= f + 1 < 2^s
  const uint64_t j = f + 1 - (1ull << (s - 1));
  uint32_t x = 1;
  for (int b = 0; b < s - 1; b++)
    if ((j >> b) & 1) x = M31Arith::mul(x, pows[L - s + b]);
  flat[f] = x;
}

// lanes[lane * n + i] = data[i * 8 + lane] (u64 values); one thread per u64.
__global__ __launch_bounds__(256) void m31_deinterleave_kernel(const uint64_t *data, uint64_t *lanes, uint64_t n) {
  const uint64_t t = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= 8 * n) return;
  const uint64_t i = t >> 3, lane = t & 7;
  lanes[lane * n + i] = data[t];
}

__global__ __launch_bounds__(256) void m31_interleave_kernel(const uint64_t *lanes, uint64_t *data, uint64_t n) {
  const uint64_t t = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= 8 * n) return;
  const uint64_t i = t >> 3, lane = t & 7;
  data[t] = lanes[lane * n + i];
}

}  // namespace

hipError_t launch_transform_m31(int L, uint32_t *data, uint32_t *scratch, const uint32_t *tw,
                                const uint32_t *scale_dev, hipStream_t stream) {
  static std::once_flag once;
  static hipError_t configured = hipSuccess;
  std::call_once(once, [] { configured = configure_pass_lds<M31T>(4 * (1 << 12)); });
  if (configured != hipSuccess) return configured;
  return run_passes_t<M31T>(kM31, L, data, scratch, tw, nullptr, scale_dev, true, stream);
}

// The eight lanes in place: every pass is one launch of the register-round
// pass kernel with three lane bits, so a tile holds the same 2^lg columns of
// all eight transforms (2^K * 2^lg * 8 <= 2^12 words) and reads and writes
// data[i * 8 + lane] with consecutive threads on consecutive words.  The
// intermediate layout in `scratch` is interleaved the same way (u32).
hipError_t launch_transform_m31_lanes(int L, uint32_t *data, uint32_t *scratch, const uint32_t *tw,
                                      const uint32_t *scale_dev, hipStream_t stream) {
  constexpr int LB = 3;
  if (L < kM31LanesMinLog || L > 28) return hipErrorInvalidValue;
  const PassPlan plan = plan_passes_shape(L, 12 - LB, 12 - LB);
  int s0 = 1;
  for (int i = 0; i < plan.passes; i++) {
    const int K = plan.K[i], lg = plan.log_g[i];
    const size_t tile = (size_t)1 << (K + lg + LB);
    // The kernel's shape: at least 512 words a tile, 8 elements a thread.
    if (K < 5 || K > 9 || tile < 512 || tile > 4096) return hipErrorInvalidValue;
    PassArgs a = {};
    a.first = i == 0;
    a.last = i == plan.passes - 1;
    a.src = a.first ? data : scratch;
    a.dst = a.last ? data : scratch;
    a.tw = tw;
    a.coset = nullptr;
    a.has_scale = (a.last && scale_dev) ? 1 : 0;
    if (a.has_scale) a.scale[0] = scale_dev[0];
    a.L = L;
    a.src_len = (uint64_t)1 << L;
    a.tw_one = 1;
    a.s0 = s0;
    a.K = K;
    a.log_g = lg + LB;
    const dim3 grid((unsigned)(((size_t)1 << L) >> (K + lg))), block((unsigned)(tile / 8));
    const size_t lds = tile * 4;
    const bool odd = (s0 & 1) != 0;
    switch (K * 2 + (odd ? 1 : 0)) {
#define ZNTT_M31L(k)                                                                                            \
  case 2 * k: hipLaunchKernelGGL((ntt_pass_r8_kernel<M31T, k, false, 8, LB>), grid, block, lds, stream, a); break; \
  case 2 * k + 1: hipLaunchKernelGGL((ntt_pass_r8_kernel<M31T, k, true, 8, LB>), grid, block, lds, stream, a); break;
      ZNTT_M31L(5) ZNTT_M31L(6) ZNTT_M31L(7) ZNTT_M31L(8) ZNTT_M31L(9)
#undef ZNTT_M31L
      default: return hipErrorInvalidValue;
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    s0 += K;
  }
  return hipSuccess;
}

hipError_t launch_m31_table(int L, uint32_t *flat, const uint32_t *pows, hipStream_t s) {
  const uint64_t n = 1ull << L;
  hipLaunchKernelGGL(m31_table_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, flat, L, pows);
  return hipGetLastError();
}

hipError_t launch_m31_lanes(bool gather, const void *src, void *dst, uint64_t n, hipStream_t s) {
  const unsigned grid = (unsigned)((8 * n + 255) / 256);
  if (gather)
    hipLaunchKernelGGL(m31_deinterleave_kernel, dim3(grid), dim3(256), 0, s, static_cast<const uint64_t *>(src),
                       static_cast<uint64_t *>(dst), n);
  else
    hipLaunchKernelGGL(m31_interleave_kernel, dim3(grid), dim3(256), 0, s, static_cast<const uint64_t *>(src),
                       static_cast<uint64_t *>(dst), n);
  return hipGetLastError();
}

}  // namespace zntt
