// HIP kernels of the MI355X NTT library (gfx950).
//
// The transform is the reference's radix-2 decimation-in-time DAG
// (libs/ntt/src/root.zig:65-90 and :148-175): bit-reverse, then stages
// s = 1..L, butterfly (x[k+j], x[k+j+h]) with twiddle tw[s][j],
// tw[s][j] = root^(j * 2^(L-s)), h = 2^(s-1).  The kernels execute exactly
// that DAG -- same butterflies, same twiddle values -- so the output equals the
// reference's for every root, primitive or not.  What changes is the schedule:
//
//   * Stages are grouped into passes of K consecutive stages.  For stages
//     s0..s0+K-1 the butterflies only couple positions that share every bit
//     outside [s0-1, s0-1+K), so the array splits into independent groups of
//     2^K elements at stride 2^(s0-1).  A workgroup loads a tile of G such
//     groups (G consecutive strides, so global accesses are G elements wide)
//     into LDS, runs the K stages there, and writes the tile back.
//   * The bit-reversal permutation is fused into the first pass's gather:
//     bit-reversed position hi*2^K + mid holds input rev_K(mid)*2^(L-K) +
//     rev_{L-K}(hi), so consecutive values of rev(hi) read consecutive inputs.
//   * The coset pre-scale (data[i] *= shift^i) is fused into that gather; the
//     inverse's n^-1 scale and the final canonical reduction into the last
//     pass's store.
//   * Between passes data lives in a scratch buffer in the device layout
//     (Fr: nine 29-bit limbs, lazily reduced), so only the first pass reads and
//     the last pass writes the caller's layout.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "field_traits.hpp"

#include "kernels.hpp"

namespace zntt {

hipError_t launch_transform(int field, int L, uint32_t *data, uint32_t *scratch, const uint32_t *tw,
                            const uint32_t *coset, const uint32_t *scale_dev, bool tw_one, hipStream_t stream) {
  if (L == 0) return hipSuccess;
  if (field == kFr) return launch_transform_fr(L, data, scratch, tw, coset, scale_dev, tw_one, stream);
  if (field == kGl) return launch_transform_gl(L, data, scratch, tw, coset, scale_dev, tw_one, stream);
  return launch_transform_f31(field, L, data, scratch, tw, coset, scale_dev, tw_one, stream);
}

// ---------------------------------------------------------------------------
// Twiddle tables.
// ---------------------------------------------------------------------------

// out[i] = base^i for i < count; thread t computes a run of kRun consecutive
// powers starting from base^(t*kRun) (square-and-multiply over `pows`).
constexpr int kRun = 16;

template <class F>
__global__ void powers_kernel(uint32_t *out, size_t count, const uint32_t *pows, int npows,
                              const uint32_t *one_dev) {
  using E = typename F::E;
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t start = t * kRun;
  if (start >= count) return;
  E x = F::from_words(one_dev);
  for (int k = 0; k < npows; k++)
    if ((start >> k) & 1) x = F::canon(F::mul(x, F::load_dev(pows, k)));
  const E base = F::load_dev(pows, 0);
  for (int i = 0; i < kRun && start + i < count; i++) {
    F::store_dev(out, start + i, x);
    x = F::canon(F::mul(x, base));
  }
}

// Stages 1..L-1 are strided copies of stage L (the powers table):
// stage s entry j = root^(j * 2^(L-s)) = stageL[j << (L - s)].
template <class F>
__global__ void stages_kernel(uint32_t *flat, int L) {
  using E = typename F::E;
  const size_t f = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t n_low = ((size_t)1 << (L - 1)) - 1;  // entries of stages 1..L-1
  if (f >= n_low) return;
  const int s = 64 - __clzll((unsigned long long)(f + 1));  // 1-based stage
  const size_t j = f + 1 - ((size_t)1 << (s - 1));
  const E x = F::load_dev(flat, n_low + (j << (L - s)));
  F::store_dev(flat, f, x);
}

template <class F>
static hipError_t build_table_t(int L, uint32_t *flat, const uint32_t *pows, const uint32_t *one_dev,
                                hipStream_t stream) {
  using E = typename F::E;
  const size_t half = (size_t)1 << (L - 1);
  const size_t n_low = half - 1;
  // stage L = powers of root, at offset 2^(L-1) - 1
  const size_t threads = (half + kRun - 1) / kRun;
  hipLaunchKernelGGL(powers_kernel<F>, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, stream,
                     flat + n_low * F::W, half, pows, L, one_dev);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (n_low > 0) {
    hipLaunchKernelGGL(stages_kernel<F>, dim3((unsigned)((n_low + 255) / 256)), dim3(256), 0, stream, flat, L);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  // trailing one
  (void)sizeof(E);
  return hipMemcpyAsync(flat + (((size_t)1 << L) - 1) * F::W, one_dev, sizeof(uint32_t) * F::W,
                        hipMemcpyDeviceToDevice, stream);
}

hipError_t launch_build_table(int field, int L, uint32_t *flat, const uint32_t *pows, const uint32_t *one_dev,
                              hipStream_t stream) {
  if (L == 0) return hipMemcpyAsync(flat, one_dev, 4 * dev_words(field), hipMemcpyDeviceToDevice, stream);
  return with_field(field, [&](auto f) { return build_table_t<decltype(f)>(L, flat, pows, one_dev, stream); });
}

hipError_t launch_powers(int field, size_t count, uint32_t *out, const uint32_t *pows, int npows,
                         const uint32_t *one_dev, hipStream_t stream) {
  const size_t threads = (count + kRun - 1) / kRun;
  const unsigned blocks = (unsigned)((threads + 255) / 256);
  with_field(field, [&](auto f) {
    hipLaunchKernelGGL(powers_kernel<decltype(f)>, dim3(blocks), dim3(256), 0, stream, out, count, pows, npows,
                       one_dev);
    return 0;
  });
  return hipGetLastError();
}

// Fr: dev = mont(user, 2^266 mod r) = x * 2^261 (canonical); Goldilocks: copy.
__global__ void fr_from_user_kernel(uint32_t *dev, const uint32_t *user, size_t count) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const Fr9 k266 = {{0x0fffead7u, 0x1d5444f4u, 0x04438aa5u, 0x03b4d096u, 0x134c84dau, 0x0e92d304u, 0x14cb95b3u,
                     0x041b9d3du, 0x00058003u}};
  Fr9 x = FrT::load_user(user, i);
  x = fr_mul(x, k266);
  fr_canon(x);
  FrT::store_dev(dev, i, x);
}

// Fr: user = canonical mont(dev, 2^256 mod r) = x * 2^256.
__global__ void fr_to_user_kernel(uint32_t *user, const uint32_t *dev, size_t count) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const Fr9 k256 = {{0x0ffffffbu, 0x04b1a0e2u, 0x18334a6bu, 0x18ed2b3eu, 0x1462e36fu, 0x11b7bc3cu, 0x1cbd99bau,
                     0x183340fbu, 0x000e0a77u}};
  Fr9 x = FrT::load_dev(dev, i);
  x = fr_mul(x, k256);
  FrT::store_user(user, i, x);
}

// 31-bit fields: dev = mont(x, 2^64 mod p) = x * 2^32; user = mont(dev, 1).
template <class F>
__global__ void f31_from_user_kernel(uint32_t *dev, const uint32_t *user, size_t count, uint32_t r2) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) dev[i] = F::A::mont(F::load_user(user, i), r2);
}

template <class F>
__global__ void f31_to_user_kernel(uint32_t *user, const uint32_t *dev, size_t count) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) F::store_user(user, i, F::A::mont(dev[i], 1u));
}

static uint32_t f31_r2(int field) {
  const uint64_t p = field == kBb ? kBbP : kKbP;
  const uint64_t r = ((uint64_t)1 << 32) % p;
  return (uint32_t)(r * r % p);
}

hipError_t launch_table_from_user(int field, size_t count, uint32_t *dev, const uint32_t *user,
                                  hipStream_t stream) {
  if (count == 0) return hipSuccess;
  if (field == kGl) return hipMemcpyAsync(dev, user, count * 8, hipMemcpyDeviceToDevice, stream);
  if (field == kBb || field == kKb) {
    const uint32_t r2 = f31_r2(field);
    if (field == kBb)
      hipLaunchKernelGGL(f31_from_user_kernel<BbT>, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, stream,
                         dev, user, count, r2);
    else
      hipLaunchKernelGGL(f31_from_user_kernel<KbT>, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, stream,
                         dev, user, count, r2);
    return hipGetLastError();
  }
  hipLaunchKernelGGL(fr_from_user_kernel, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, stream, dev, user,
                     count);
  return hipGetLastError();
}

hipError_t launch_table_to_user(int field, size_t count, uint32_t *user, const uint32_t *dev, hipStream_t stream) {
  if (count == 0) return hipSuccess;
  if (field == kGl) return hipMemcpyAsync(user, dev, count * 8, hipMemcpyDeviceToDevice, stream);
  if (field == kBb || field == kKb) {
    if (field == kBb)
      hipLaunchKernelGGL(f31_to_user_kernel<BbT>, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, stream, user,
                         dev, count);
    else
      hipLaunchKernelGGL(f31_to_user_kernel<KbT>, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, stream, user,
                         dev, count);
    return hipGetLastError();
  }
  hipLaunchKernelGGL(fr_to_user_kernel, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, stream, user, dev,
                     count);
  return hipGetLastError();
}

template <class F>
__global__ void inverse_table_kernel(uint32_t *inv, const uint32_t *fwd, int L, const uint32_t *one_dev) {
  using E = typename F::E;
  const size_t f = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t total = ((size_t)1 << L) - 1;
  if (f > total) return;
  if (f == total) {  // trailing one
    F::store_dev(inv, f, F::load_dev(one_dev, 0));
    return;
  }
  const int s = 64 - __clzll((unsigned long long)(f + 1));
  const size_t h = (size_t)1 << (s - 1), off = h - 1, j = f - off;
  const E x = j == 0 ? F::load_dev(one_dev, 0) : F::neg(F::load_dev(fwd, off + h - j));
  F::store_dev(inv, f, x);
}

hipError_t launch_inverse_table(int field, int L, uint32_t *inv, const uint32_t *fwd, hipStream_t stream) {
  // The forward table's trailing entry is one in dev form; reuse it.
  const size_t total = ((size_t)1 << L);
  const uint32_t *one_dev = fwd + (total - 1) * dev_words(field);
  const unsigned blocks = (unsigned)((total + 255) / 256);
  with_field(field, [&](auto f) {
    hipLaunchKernelGGL(inverse_table_kernel<decltype(f)>, dim3(blocks), dim3(256), 0, stream, inv, fwd, L,
                       one_dev);
    return 0;
  });
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// bitReverse (root.zig:38-58): swap i <-> rev(i) for rev(i) > i.
// ---------------------------------------------------------------------------
template <int UW>
__global__ void bit_reverse_kernel(uint32_t *data, int L) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >> L) return;
  const size_t j = (size_t)(__brevll((unsigned long long)i) >> (64 - L));
  if (j <= i) return;
  uint32_t *a = data + i * UW, *b = data + j * UW;
#pragma unroll
  for (int k = 0; k < UW; k++) {
    const uint32_t t = a[k];
    a[k] = b[k];
    b[k] = t;
  }
}

template <class F>
__global__ void pointwise_mul_kernel(uint32_t *a, const uint32_t *b, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    F::store_user(a, i, F::mul(F::load_user(a, i), F::load_user(b, i)));
}

hipError_t launch_pointwise_mul(int field, uint32_t *a, const uint32_t *b, size_t n, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  const unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 65536);
  with_field(field, [&](auto f) {
    hipLaunchKernelGGL(pointwise_mul_kernel<decltype(f)>, dim3(blocks), dim3(256), 0, stream, a, b, n);
    return 0;
  });
  return hipGetLastError();
}

hipError_t launch_bit_reverse(int field, int L, uint32_t *data, hipStream_t stream) {
  if (L == 0) return hipSuccess;
  const size_t n = (size_t)1 << L;
  const unsigned blocks = (unsigned)((n + 255) / 256);
  if (field == kFr)
    hipLaunchKernelGGL(bit_reverse_kernel<8>, dim3(blocks), dim3(256), 0, stream, data, L);
  else
    hipLaunchKernelGGL(bit_reverse_kernel<2>, dim3(blocks), dim3(256), 0, stream, data, L);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Diagnostics.
// ---------------------------------------------------------------------------
// Fr ops on user-layout operands.  mul: the device multiply takes its second
// operand in dev form, so b is converted first (mont(b, 2^266) = b * 2^261);
// the product of two reference-form values is then reference-form.
__global__ void fr_op_kernel(int op, size_t n, const uint32_t *a, const uint32_t *b, uint32_t *out) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  constexpr uint32_t C[9] = ZNTT_FR_C4R;
  Fr9 x = FrT::load_user(a, i), y = FrT::load_user(b, i), r;
  if (op == 0) {
    r = FrT::add_lazy(x, y);
  } else if (op == 1) {
#pragma unroll
    for (int k = 0; k < 9; k++) r.v[k] = x.v[k] + C[k] - y.v[k];
  } else {
    const Fr9 k266 = {{0x0fffead7u, 0x1d5444f4u, 0x04438aa5u, 0x03b4d096u, 0x134c84dau, 0x0e92d304u, 0x14cb95b3u,
                       0x041b9d3du, 0x00058003u}};
    Fr9 yd = fr_mul(y, k266);
    fr_canon(yd);
    r = fr_mul(x, yd);
  }
  FrT::store_user(out, i, r);
}

// 31-bit ops on user-layout operands; mul = mont(mont(x, y), 2^64 mod p).
template <class F>
__global__ void f31_op_kernel(int op, size_t n, const uint32_t *a, const uint32_t *b, uint32_t *out, uint32_t r2) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t x = F::load_user(a, i), y = F::load_user(b, i);
  const uint32_t r = op == 0 ? F::A::add(x, y) : (op == 1 ? F::A::sub(x, y) : F::A::mont(F::A::mont(x, y), r2));
  F::store_user(out, i, r);
}

__global__ void gl_op_kernel(int op, size_t n, const uint64_t *a, const uint64_t *b, uint64_t *out) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t x = a[i], y = b[i];
  out[i] = op == 0 ? gl_add(x, y) : (op == 1 ? gl_sub(x, y) : gl_mul(x, y));
}

hipError_t launch_field_op(int field, int op, size_t n, const uint32_t *a, const uint32_t *b, uint32_t *out,
                           hipStream_t stream) {
  if (n == 0) return hipSuccess;
  const unsigned blocks = (unsigned)((n + 255) / 256);
  if (field == kFr)
    hipLaunchKernelGGL(fr_op_kernel, dim3(blocks), dim3(256), 0, stream, op, n, a, b, out);
  else if (field == kGl)
    hipLaunchKernelGGL(gl_op_kernel, dim3(blocks), dim3(256), 0, stream, op, n, (const uint64_t *)a,
                       (const uint64_t *)b, (uint64_t *)out);
  else if (field == kBb)
    hipLaunchKernelGGL(f31_op_kernel<BbT>, dim3(blocks), dim3(256), 0, stream, op, n, a, b, out, f31_r2(field));
  else
    hipLaunchKernelGGL(f31_op_kernel<KbT>, dim3(blocks), dim3(256), 0, stream, op, n, a, b, out, f31_r2(field));
  return hipGetLastError();
}

// Eight independent v_mad_u64_u32 chains per lane, 16 per loop trip.
__global__ void mad_probe_kernel(uint64_t *out, uint32_t x, int iters) {
  uint64_t a0 = threadIdx.x, a1 = a0 + 1, a2 = a0 + 2, a3 = a0 + 3, a4 = a0 + 4, a5 = a0 + 5, a6 = a0 + 6,
           a7 = a0 + 7;
  const uint32_t xx = x + threadIdx.x, yy = x ^ blockIdx.x;
  for (int i = 0; i < iters; i++) {
#define ZNTT_MAD(acc) asm volatile("v_mad_u64_u32 %0, vcc, %1, %2, %0" : "+v"(acc) : "v"(xx), "v"(yy) : "vcc");
    ZNTT_MAD(a0) ZNTT_MAD(a1) ZNTT_MAD(a2) ZNTT_MAD(a3) ZNTT_MAD(a4) ZNTT_MAD(a5) ZNTT_MAD(a6) ZNTT_MAD(a7)
    ZNTT_MAD(a0) ZNTT_MAD(a1) ZNTT_MAD(a2) ZNTT_MAD(a3) ZNTT_MAD(a4) ZNTT_MAD(a5) ZNTT_MAD(a6) ZNTT_MAD(a7)
#undef ZNTT_MAD
  }
  out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = a0 ^ a1 ^ a2 ^ a3 ^ a4 ^ a5 ^ a6 ^ a7;
}

__global__ void copy_probe_kernel(uint4 *dst, const uint4 *src, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

hipError_t probe_peaks(double *mad_per_s, double *copy_bytes_per_s, hipStream_t stream) {
  int dev = 0, cus = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e != hipSuccess) return e;
  hipEvent_t e0, e1;
  if ((e = hipEventCreate(&e0)) != hipSuccess) return e;
  if ((e = hipEventCreate(&e1)) != hipSuccess) return e;
  const int blocks = cus * 8, threads = 256, iters = 8192;
  uint64_t *out = nullptr;
  e = hipMalloc(&out, sizeof(uint64_t) * blocks * threads);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(mad_probe_kernel, dim3(blocks), dim3(threads), 0, stream, out, 3u, 64);
    hipEventRecord(e0, stream);
    hipLaunchKernelGGL(mad_probe_kernel, dim3(blocks), dim3(threads), 0, stream, out, 5u, iters);
    hipEventRecord(e1, stream);
    e = hipEventSynchronize(e1);
    float ms = 0;
    hipEventElapsedTime(&ms, e0, e1);
    *mad_per_s = 16.0 * iters * (double)blocks * threads / (ms * 1e-3);
    hipFree(out);
  }
  if (e == hipSuccess) {
    const size_t bytes = (size_t)1 << 30;
    uint4 *a = nullptr, *b = nullptr;
    e = hipMalloc(&a, bytes);
    if (e == hipSuccess) e = hipMalloc(&b, bytes);
    if (e == hipSuccess) {
      const size_t n = bytes / sizeof(uint4);
      hipMemsetAsync(a, 1, bytes, stream);
      hipLaunchKernelGGL(copy_probe_kernel, dim3(cus * 16), dim3(256), 0, stream, b, a, n);
      hipEventRecord(e0, stream);
      for (int r = 0; r < 5; r++) hipLaunchKernelGGL(copy_probe_kernel, dim3(cus * 16), dim3(256), 0, stream, b, a, n);
      hipEventRecord(e1, stream);
      e = hipEventSynchronize(e1);
      float ms = 0;
      hipEventElapsedTime(&ms, e0, e1);
      *copy_bytes_per_s = 5.0 * 2.0 * bytes / (ms * 1e-3);
    }
    if (a) hipFree(a);
    if (b) hipFree(b);
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  return e;
}

}  // namespace zntt
