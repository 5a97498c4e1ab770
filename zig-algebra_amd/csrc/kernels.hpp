// Host-side launch interface of the HIP kernels (ntt_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace zntt {

// kM31 is internal: only the eight-lane M31 transform uses it.
enum FieldId { kFr = 1, kGl = 2, kBb = 3, kKb = 4, kM31 = 5 };

// Words (u32) per element in the kernels' two layouts.
//  user: the reference's storage layout (Fr 8, Goldilocks 2)
//  dev:  scratch buffers and twiddle tables (Fr 9 radix-2^29 limbs, Goldilocks 2)
inline int user_words(int field) { return field == kFr ? 8 : 2; }
inline int dev_words(int field) { return field == kFr ? 9 : (field == kGl ? 2 : 1); }

// One pass covers DIT stages s0 .. s0+K-1 (1-based) of a 2^L-point transform.
struct PassArgs {
  const uint32_t *src;     // first pass: user layout, else dev layout (AoS)
  uint32_t *dst;           // last pass: user layout, else dev layout (AoS)
  const uint32_t *tw;      // flat per-stage table, dev layout; stage s at 2^(s-1)-1
  const uint32_t *coset;   // first pass only: shift^i for i < 2^L (dev), or null
  uint32_t scale[9];       // last pass only: dev-form multiplier when has_scale
  int has_scale;
  int L, s0, K, log_g;     // tile = 2^log_g groups of 2^K elements
  int first, last;
  uint64_t src_len;        // batch kernel, first pass: elements at or past it read as zero
  int tw_one;              // entry 0 of every stage of `tw` is 1 (a table of powers)
};

struct PassPlan {
  int passes;
  int K[8];
  int log_g[8];
  int tile_log;  // log2 of elements per workgroup tile
};

// Stage split of a 2^L-point transform into LDS passes (the field's default
// tile), and the same for an explicit tile of 2^tile_log elements with passes
// of at most kmax stages (used by the batch autotuner).
PassPlan plan_passes(int field, int L);
PassPlan plan_passes_shape(int L, int tile_log, int kmax);

// Enqueue a full transform over `data` (user layout, 2^L elements, in place).
// `tw` is a flat dev-layout table holding at least 2^L - 1 entries.
// `scratch` must hold 2^L dev-layout elements when L exceeds one tile.
// `tw_one`: entry 0 of every stage of `tw` is 1 (tables of powers), which
// lets the first stages skip multiplications by it.
hipError_t launch_transform(int field, int L, uint32_t *data, uint32_t *scratch, const uint32_t *tw,
                            const uint32_t *coset, const uint32_t *scale_dev, bool tw_one, hipStream_t stream);

// Per-field pass launchers (pass_fr.hip, pass_gl.hip).
hipError_t launch_transform_fr(int L, uint32_t *data, uint32_t *scratch, const uint32_t *tw, const uint32_t *coset,
                               const uint32_t *scale_dev, bool tw_one, hipStream_t stream);
hipError_t launch_transform_gl(int L, uint32_t *data, uint32_t *scratch, const uint32_t *tw, const uint32_t *coset,
                               const uint32_t *scale_dev, bool tw_one, hipStream_t stream);
hipError_t launch_transform_f31(int field, int L, uint32_t *data, uint32_t *scratch, const uint32_t *tw,
                                const uint32_t *coset, const uint32_t *scale_dev, bool tw_one, hipStream_t stream);

// Twiddle table builders (dev layout).  `flat` holds 2^L entries: stages
// 1..L at offsets 2^(s-1)-1 and a trailing one (index 2^L - 1).
//   pows: dev-form root^(2^k), k < L (device buffer, L entries)
hipError_t launch_build_table(int field, int L, uint32_t *flat, const uint32_t *pows, const uint32_t *one_dev,
                              hipStream_t stream);
// Powers base^i for i < count (dev layout), pows = base^(2^k) for k < 64.
hipError_t launch_powers(int field, size_t count, uint32_t *out, const uint32_t *pows, int npows,
                         const uint32_t *one_dev, hipStream_t stream);
// Dev-layout table from user-layout entries (count entries).
hipError_t launch_table_from_user(int field, size_t count, uint32_t *dev, const uint32_t *user, hipStream_t stream);
// User-layout entries from a dev-layout table (canonical).
hipError_t launch_table_to_user(int field, size_t count, uint32_t *user, const uint32_t *dev, hipStream_t stream);
// Inverse table for inttWithTwiddles: inv[s][j] = j == 0 ? 1 : -fwd[s][h - j].
hipError_t launch_inverse_table(int field, int L, uint32_t *inv, const uint32_t *fwd, hipStream_t stream);
// In-place bit-reversal permutation of n = 2^L user-layout elements.
hipError_t launch_bit_reverse(int field, int L, uint32_t *data, hipStream_t stream);
// a[i] <- a[i] * b[i] for i < n, both in the caller's layout.  BN254 Fr:
// the Montgomery product of two data-form values is x*y*2^251 (R' = 2^261),
// which a later scale by c*2^5 (dev form) returns to the 2^256 form.
hipError_t launch_pointwise_mul(int field, uint32_t *a, const uint32_t *b, size_t n, hipStream_t stream);
// Merkle trees (merkle_kernels.hip).  heap is the node array (32 bytes a
// node, node s + i of the level starting at s); with `full` it receives the
// whole tree (2^(cap_log + 1) nodes), otherwise only the nodes above the
// first launch, 2^merkle_heap_log(cap_log, false) of them.  Node 1 is the root.
int merkle_heap_log(int cap_log, bool full);
// FRI layer of 2^(leaf_log + 1) user-layout values: leaf j hashes values j
// and j + 2^leaf_log.
hipError_t launch_merkle_layer(int field, const void *evals, int leaf_log, uint32_t *heap, bool full,
                               hipStream_t s);
// Fold the previous FRI layer (2^(leaf_log + 2) values, user layout) into
// `next` (2^(leaf_log + 1) values) and commit `next` in the same launch: the
// whole tree into heap, as launch_merkle_layer with `full`.  inv_pows[x] =
// g^-x for the previous layer's generator g, as launch_fri_fold.
hipError_t launch_merkle_fold(int field, const void *prev, void *next, int leaf_log, uint32_t *heap,
                              const uint32_t *inv_pows, const uint32_t alpha_dev[9], hipStream_t s);
// Tree over `count` stored hashes padded to 2^cap_log leaves.
hipError_t launch_merkle_from_hashes(const void *hashes, uint64_t count, int cap_log, uint32_t *heap, bool full,
                                     hipStream_t s);
// One FRI fold (fri_kernels.hip): next[j] = (cur[j] + cur[j+half]) / 2 +
// alpha (cur[j] - cur[j+half]) / (2 g^j), user layout, with inv_pows[j] =
// g^-j (dev layout) and alpha in dev form.
hipError_t launch_fri_fold(int field, const void *cur, void *next, uint64_t half, const uint32_t *inv_pows,
                           const uint32_t alpha_dev[9], hipStream_t s);
// FRI query openings: for each of q queries (pair index e0[k]) and each of
// the `rounds` layers, the antipodal pair (2 elements, user layout) into
// vals and the Merkle path (32-byte nodes) into paths.
struct FriOpenArgs {
  const void *layer[64];  // layer r (user layout), r < rounds
  const void *tree[64];   // layer r's node array
  uint32_t rounds, log_domain, elem_words;
};
hipError_t launch_fri_open(const FriOpenArgs &a, const uint64_t *e0, uint64_t q, uint32_t *vals, uint32_t *paths,
                           hipStream_t s);
// Mersenne-31 (pass_m31.hip), for the eight-lane transform: one transform
// of 2^L u64 values (tw: flat table from launch_m31_table; scale_dev: a plain
// multiplier of every output, or null), the table of root's stage powers
// (pows[k] = root^(2^k), k < L, plain u32), and gathering the eight
// interleaved lanes (data[i * 8 + lane]) into lane-major order or back.
hipError_t launch_transform_m31(int L, uint32_t *data, uint32_t *scratch, const uint32_t *tw,
                                const uint32_t *scale_dev, hipStream_t stream);
hipError_t launch_m31_table(int L, uint32_t *flat, const uint32_t *pows, hipStream_t s);
hipError_t launch_m31_lanes(bool gather, const void *src, void *dst, uint64_t n, hipStream_t s);
// All eight interleaved lanes of 2^L u64 values each, transformed in place
// without gathering them (data[i * 8 + lane]; scratch: 8 * 2^L u32).  For
// 6 <= L <= 28; returns hipErrorInvalidValue outside that range.
constexpr int kM31LanesMinLog = 6;
hipError_t launch_transform_m31_lanes(int L, uint32_t *data, uint32_t *scratch, const uint32_t *tw,
                                      const uint32_t *scale_dev, hipStream_t stream);
// Elementwise device field op (0 add, 1 sub, 2 mul) on user-layout arrays.
hipError_t launch_field_op(int field, int op, size_t n, const uint32_t *a, const uint32_t *b, uint32_t *out,
                           hipStream_t stream);
// The static device scratch of zntt_execute_device_noalloc
// (ZNTT_NOALLOC_SCRATCH_BYTES a GPU, noalloc_scratch.hip): init looks up its
// address on `device`, which must be current; ptr returns it (null before).
hipError_t noalloc_scratch_init(int device);
void *noalloc_scratch_ptr(int device);
// Microbenchmarks: v_mad_u64_u32 rate (mul/s) and device copy bandwidth (B/s).
hipError_t probe_peaks(double *mad_per_s, double *copy_bytes_per_s, hipStream_t stream);

}  // namespace zntt
