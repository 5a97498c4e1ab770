/*
 * zntt_gpu.h -- C ABI of the MI355X (gfx950) NTT library, libzntt_gpu.so.
 *
 * The library is a drop-in backend for zig-algebra's `libs/ntt` (zig-ntt).
 * Its seven core entry points are exactly the functions a Zig `GpuNtt` shim
 * would bind through `extern fn`, one per zig-ntt function, with the same
 * argument meaning, the same element layout and the same error sets:
 *
 *   zig-ntt (libs/ntt/src/root.zig)            C ABI
 *   ----------------------------------------   -------------------------------
 *   bitReverse          :38-58                 zntt_bit_reverse
 *   ntt                 :65-90                 zntt_ntt
 *   intt                :95-107                zntt_intt
 *   precomputeTwiddles  :112-134               zntt_precompute_twiddles
 *   freeTwiddles        :136-141               zntt_free_twiddles
 *   nttWithTwiddles     :148-175               zntt_ntt_with_twiddles
 *   inttWithTwiddles    :180-212               zntt_intt_with_twiddles
 *
 * Element layout (byte-identical to the Zig types, so the shim passes
 * `data.ptr` straight through):
 *   ZNTT_FIELD_BN254_FR   32 bytes: four little-endian u64 limbs of the
 *                         Montgomery form x * 2^256 mod r (libs/field/src/
 *                         field.zig:884, montgomery.zig:23-79), r from
 *                         libs/curve/src/bn254.zig:67.
 *   ZNTT_FIELD_GOLDILOCKS  8 bytes: the canonical u64 value
 *                         (libs/field/src/field.zig:142), p = 2^64 - 2^32 + 1.
 *   ZNTT_FIELD_BABYBEAR,   8 bytes: the canonical u64 value, as Goldilocks;
 *   ZNTT_FIELD_KOALABEAR   p = 2^31 - 2^27 + 1 and 2^31 - 2^24 + 1.
 *
 * Semantics:
 *   - Host-buffer calls are synchronous and in place, like the Zig functions:
 *     they return with `data` updated.  They are serialised per context.
 *   - `root` is not validated (libs/ntt/README.md:151-152): outputs equal the
 *     reference's radix-2 DIT DAG for every root, primitive or not.
 *   - Error order follows the reference: LogTooLarge -> LengthMismatch ->
 *     InvalidTwiddles (table count) -> InvalidTwiddles (stage lengths).
 *     Deliberate deviation: every stage length is checked before `data` is
 *     touched (the reference bit-reverses first, root.zig:154-161), so a
 *     rejected call leaves `data` unchanged.
 *   - Device-only conditions (HIP errors, device memory, the context's HBM
 *     budget) are reported as ZNTT_DEVICE_ERROR / ZNTT_OUT_OF_MEMORY; a Zig
 *     shim that must keep the reference's exact error sets maps them to a
 *     panic.
 */
#ifndef ZNTT_GPU_H
#define ZNTT_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* Status codes (0 = success).                                               */
/* ------------------------------------------------------------------------ */
enum zntt_status {
  ZNTT_OK = 0,
  ZNTT_LOG_TOO_LARGE = 1,     /* error.LogTooLarge      root.zig:33-36        */
  ZNTT_LENGTH_MISMATCH = 2,   /* error.LengthMismatch   root.zig:68,98,151,183 */
  ZNTT_INVALID_LENGTH = 3,    /* error.InvalidLength    root.zig:43           */
  ZNTT_INVALID_TWIDDLES = 4,  /* error.InvalidTwiddles  root.zig:152,161,184,193 */
  ZNTT_OUT_OF_MEMORY = 5,     /* error.OutOfMemory      root.zig:112 (and device memory / HBM budget) */
  ZNTT_DEVICE_ERROR = 6,      /* a HIP runtime call failed                       */
  ZNTT_NO_DEVICE = 7,         /* zntt_ctx_create: no such HIP device             */
  ZNTT_INVALID_ARGUMENT = 8,  /* null context/pointer, unknown field id          */
  ZNTT_UNSUPPORTED = 9,       /* a size this device path does not implement      */
  ZNTT_ORDER_TOO_LARGE = 10,  /* error.OrderTooLarge    roots.zig:150 (LDE/roots) */
  ZNTT_DEGREE_TOO_LARGE = 11, /* error.DegreeTooLarge   libs/poly/src/poly.zig:200 */
  ZNTT_INVALID_PARAMETERS = 12, /* FriError.InvalidParameters libs/fri/src/root.zig:77 */
  ZNTT_NOT_PREPARED = 13       /* zntt_execute_device_noalloc: a table it needs is not cached */
};

enum zntt_field {
  ZNTT_FIELD_BN254_FR = 1,
  ZNTT_FIELD_GOLDILOCKS = 2,
  ZNTT_FIELD_BABYBEAR = 3,   /* p = 2^31 - 2^27 + 1, libs/field/src/predef/babybear.zig  */
  ZNTT_FIELD_KOALABEAR = 4   /* p = 2^31 - 2^24 + 1, libs/field/src/predef/koalabear.zig */
};

/* Direction of a device transform (zntt_transform.direction). */
enum zntt_direction {
  ZNTT_FORWARD = 0,        /* ntt(root)                                        */
  ZNTT_INVERSE = 1,        /* intt(root): ntt(root^-1) then scale by n^-1      */
  ZNTT_COSET_FORWARD = 2   /* data[i] *= shift^i, then ntt(root)               */
};

/* Largest log2 size the device path implements (2^28 BN254 Fr = 8 GiB). */
#define ZNTT_MAX_LOG_N 28

typedef struct zntt_ctx zntt_ctx;

/* ------------------------------------------------------------------------ */
/* Context.  Owns a HIP stream, scratch memory and the device-resident       */
/* twiddle cache.  No zig-ntt counterpart: zig-ntt is stateless; the shim    */
/* creates one context per GPU and passes it to every call.                  */
/* ------------------------------------------------------------------------ */

/* hbm_budget_bytes: cap on device memory the context may hold (0 = 64 GiB). */
int zntt_ctx_create(zntt_ctx **out, int device, uint64_t hbm_budget_bytes);
/* Releases the caller's handle.  Batches and servers made on `ctx` hold a
 * reference to it: while any is alive, the context's stream, lock and the
 * twiddle tables they use stay valid (unused cache entries and buffers are
 * freed at once), and the last zntt_batch_destroy / zntt_server_destroy
 * frees the rest.  `ctx` itself must not be passed to any call afterwards. */
void zntt_ctx_destroy(zntt_ctx *ctx);
/* Device bytes currently held (scratch + twiddle cache + staging). */
uint64_t zntt_ctx_bytes_in_use(const zntt_ctx *ctx);
/* Frees every cached twiddle / coset table that no live batch or server
 * uses (those stay, pinned), and the context's scratch and staging buffers.
 * A reservation (zntt_ctx_reserve) stays.  Waits for the context's stream
 * and the device first. */
int zntt_ctx_trim(zntt_ctx *ctx);
/* Confines device-memory failures of the seven zig-ntt functions below to
 * this call, for sizes up to 2^max_log_n of every field.  zig-ntt's error
 * sets have no OutOfMemory for ntt, intt, bitReverse or *WithTwiddles
 * (root.zig:33-212); a backend that allocates device memory lazily could
 * only turn a full device into a panic there.  This call allocates, up
 * front, everything those functions use at such sizes: the staging and
 * pass-scratch buffers, a foreign table's upload buffer, and two table
 * slots of 2^max_log_n entries.  Afterwards ntt, intt, bitReverse,
 * nttWithTwiddles and inttWithTwiddles at log_n <= max_log_n allocate
 * nothing: a table the cache does not hold is built into the least recently
 * used slot (a root's powers for ntt / intt, the derived inverse of a
 * registered table for inttWithTwiddles).  precomputeTwiddles still
 * allocates its table (its error set has OutOfMemory).  The reservation
 * lasts until the context is freed; a second call with a larger max_log_n
 * grows it, a smaller one changes nothing.  Returns ZNTT_OUT_OF_MEMORY or
 * ZNTT_DEVICE_ERROR (the reservation is then unchanged; buffers it already
 * grew stay), ZNTT_UNSUPPORTED above ZNTT_MAX_LOG_N.  No zig-ntt
 * counterpart: the shim calls it once after zntt_ctx_create.
 * Device bytes, sized for BN254 Fr (the largest element: 32 B in the
 * caller's layout, 36 B on the device): 2^max_log_n x 244 B, 4.1 GB at
 * 2^24. */
int zntt_ctx_reserve(zntt_ctx *ctx, size_t max_log_n);
const char *zntt_status_string(int status);
/* Bytes per element of a field, 0 for an unknown id. */
size_t zntt_element_size(int field);

/* ------------------------------------------------------------------------ */
/* The seven zig-ntt functions (host buffers, synchronous, in place).        */
/* ------------------------------------------------------------------------ */
/* Host buffers over 16 MiB move through the context's 64 MiB of page-locked
 * host memory, copied by three threads of the context while the DMA engine
 * moves the previous part; both are made at the first such call (or by
 * zntt_ctx_reserve at a size that needs them) and freed by zntt_ctx_trim
 * (unless reserved) or zntt_ctx_destroy.  Without them the call copies from
 * the caller's pageable memory directly, more slowly. */

/* bitReverse(F, data) -- root.zig:38-58 */
int zntt_bit_reverse(zntt_ctx *ctx, int field, void *data, size_t len);

/* ntt(F, data, log_n, root) -- root.zig:65-90.  `root` points at one element. */
int zntt_ntt(zntt_ctx *ctx, int field, void *data, size_t len, size_t log_n, const void *root);

/* intt(F, data, log_n, root) -- root.zig:95-107 */
int zntt_intt(zntt_ctx *ctx, int field, void *data, size_t len, size_t log_n, const void *root);

/* precomputeTwiddles(F, log_n, root, allocator) -- root.zig:112-134.
 * The caller owns the table memory, as with the reference's allocator: it
 * passes `num_stages == log_n` host buffers, stage s holding 2^s elements,
 * and the library fills them (tw[s][j] = (root^(2^(log_n-1-s)))^j).  The
 * table is also registered, keyed by `stages`, in the context's device-
 * resident cache so that *_with_twiddles on the same table reads it from HBM.
 * Errors: LogTooLarge (log_n >= 64), OutOfMemory, InvalidArgument
 * (num_stages != log_n or a null stage pointer). */
int zntt_precompute_twiddles(zntt_ctx *ctx, int field, size_t log_n, const void *root,
                             void *const *stages, size_t num_stages);

/* freeTwiddles(F, tw, allocator) -- root.zig:136-141.  Releases the device
 * copy registered for `stages`; the caller then frees its host buffers. */
int zntt_free_twiddles(zntt_ctx *ctx, const void *const *stages, size_t num_stages);

/* nttWithTwiddles(F, data, log_n, twiddles) -- root.zig:148-175.
 * `stages`/`stage_lens`/`num_stages` are the Zig `[]const []const F`.  A table
 * registered by zntt_precompute_twiddles is read from the device cache; any
 * other (foreign) table is uploaded and used exactly as given.  A registered
 * table is found by its stage pointers and then checked against a
 * fingerprint of sampled entries taken at registration: buffers freed
 * without zntt_free_twiddles and reused for other content are detected, the
 * stale registration dropped and the table used as a foreign one.  (A
 * registered table edited in place must be registered again.) */
int zntt_ntt_with_twiddles(zntt_ctx *ctx, int field, void *data, size_t len, size_t log_n,
                           const void *const *stages, const size_t *stage_lens, size_t num_stages);

/* inttWithTwiddles(F, data, log_n, twiddles) -- root.zig:180-212
 * (w = 1 for j = 0, else -tw[s-1][h-j]; then scale by n^-1). */
int zntt_intt_with_twiddles(zntt_ctx *ctx, int field, void *data, size_t len, size_t log_n,
                            const void *const *stages, const size_t *stage_lens, size_t num_stages);

/* ------------------------------------------------------------------------ */
/* Beyond the zig-ntt surface.  No reference counterpart: the reference has  */
/* no coset transform, LDE, device-resident or batched API.                  */
/* ------------------------------------------------------------------------ */

/* Coset-forward NTT (host buffer): data[i] *= shift^i, then ntt(root).
 * Errors as zntt_ntt. */
int zntt_coset_ntt(zntt_ctx *ctx, int field, void *data, size_t len, size_t log_n,
                   const void *root, const void *shift);

/* Low-degree extension (host buffers): zero-pad 2^log_k coefficients to
 * 2^(log_k+log_blowup) and run the coset-forward NTT with the field's
 * primitive 2^(log_k+log_blowup)-th root (libs/field/src/roots.zig:149-157).
 * `out` holds 2^(log_k+log_blowup) elements. */
int zntt_lde(zntt_ctx *ctx, int field, const void *coeffs, size_t log_k, size_t log_blowup,
             const void *shift, void *out);

/* The field's primitive 2^log_size-th root of unity, as
 * libs/field/src/roots.zig:149-157 defines it (z = smallest QNR >= 2,
 * z^(odd_part << (two_adicity - log_size))).  ORDER_TOO_LARGE past the
 * field's two-adicity. */
int zntt_primitive_root(int field, size_t log_size, void *out);

/* One transform on device-resident data (HBM), enqueued on `stream`, a
 * hipStream_t with HIP's meaning (NULL = the null stream; zntt_ctx_stream()
 * gives the context's own stream).  Asynchronous: returns
 * once the work is enqueued.  Transforms above one tile (BN254 Fr 2^11, the
 * other fields 2^12) use the context's scratch buffer between passes; a
 * call on a different stream from the context's previous asynchronous call
 * makes its stream wait for that call's work first (an event), so calls on
 * several streams of one context are correct and run one after another on
 * the device.  Use one context per stream for transforms that overlap.  `root`/`shift` are host pointers to one
 * element (shift only for ZNTT_COSET_FORWARD).  The twiddle table comes from
 * the device-resident cache (built on first use, on `stream`). */
typedef struct {
  int field;           /* zntt_field                          */
  int direction;       /* zntt_direction                      */
  size_t log_n;        /* transform size 2^log_n              */
  void *data;          /* device pointer, 2^log_n elements    */
  const void *root;    /* host pointer, one element           */
  const void *shift;   /* host pointer, coset shift (or NULL) */
} zntt_transform;

int zntt_execute_device(zntt_ctx *ctx, const zntt_transform *t, void *stream);

/* No counterpart in the reference.  For callers that cannot allocate
 * device memory -- code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc: zntt_execute_device with no allocation of
 * any kind (no hipMalloc, hipFree or allocating library call).
 *   - The twiddle table of (field, root, 2^log_n) -- of root^-1 for
 *     ZNTT_INVERSE -- and, for ZNTT_COSET_FORWARD, the shift's coset table
 *     must already be in the context's cache, put there by an earlier call
 *     such as zntt_execute_device on the same transform; otherwise
 *     ZNTT_NOT_PREPARED, with nothing enqueued.
 *   - Transforms above one tile keep their intermediate passes in a static
 *     device buffer of ZNTT_NOALLOC_SCRATCH_BYTES per GPU (2^log_n * 36
 *     bytes BN254 Fr, 8 Goldilocks, 4 BabyBear / KoalaBear: up to 2^20,
 *     2^23 and 2^24 elements); larger ones fail with ZNTT_OUT_OF_MEMORY
 *     before anything is enqueued.
 *   - Every such call on one GPU shares that buffer: calls whose work
 *     overlaps on the device (on different streams, or from different
 *     contexts on the same GPU) race.  Calls on one stream are ordered.
 * Results equal zntt_execute_device's bit for bit. */
#define ZNTT_NOALLOC_SCRATCH_BYTES (64u << 20)
int zntt_execute_device_noalloc(zntt_ctx *ctx, const zntt_transform *t, void *stream);

/* nttWithTwiddles / inttWithTwiddles (root.zig:148-175, :180-212) on
 * device-resident data: `data` is a device pointer of `len` = 2^log_n
 * elements, the table is the caller's host `[]const []const F` as for
 * zntt_ntt_with_twiddles (registered: read from the device cache; foreign:
 * uploaded and used as given), and the transform is enqueued on `stream`.
 * Same checks in the same order as the host-buffer calls. */
int zntt_ntt_with_twiddles_device(zntt_ctx *ctx, int field, void *data, size_t len, size_t log_n,
                                  const void *const *stages, const size_t *stage_lens, size_t num_stages,
                                  void *stream);
int zntt_intt_with_twiddles_device(zntt_ctx *ctx, int field, void *data, size_t len, size_t log_n,
                                   const void *const *stages, const size_t *stage_lens, size_t num_stages,
                                   void *stream);

/* ------------------------------------------------------------------------ */
/* Batched transforms.  No reference counterpart.  All transforms of a batch */
/* share one field; each result equals zntt_execute_device on it alone.      */
/* ------------------------------------------------------------------------ */
typedef struct zntt_batch zntt_batch;

/* Plans a batch of device transforms (sizes and directions may differ):
 * resolves and pins their twiddle / coset tables in the device cache,
 * allocates per-transform scratch and the work list.  `ts[i].data` stay the
 * transforms' buffers for the life of the batch.  A shape chosen earlier by
 * zntt_batch_autotune for the same sizes and directions is reused.  Sizes
 * run from 2^3 to 2^ZNTT_MAX_LOG_N (ZNTT_UNSUPPORTED otherwise). */
int zntt_batch_create(zntt_ctx *ctx, const zntt_transform *ts, size_t count, zntt_batch **out);
/* Enqueues every transform of the batch on `stream`, replayed from a HIP
 * graph captured on the first call: per pass level (at most four), one
 * persistent launch per kernel family of the level (the register-round
 * level kernel per pass length and parity, the generic level kernel, the
 * whole-transform kernel per size), the launches of a level on parallel
 * branches of the graph.  Do not run two executions of one batch at once (they share
 * the scratch and work counters). */
int zntt_batch_execute(zntt_batch *b, void *stream);
/* Waits for `stream` and reports a HIP error raised by the batch's launches. */
int zntt_batch_status(zntt_batch *b, void *stream);
/* Times candidate shapes (tile size, pass length, resident workgroups per CU,
 * item order) on shadow copies of the data, keeps the fastest for this batch
 * and for later batches of the same signature.  `best_ms`: its time per
 * execution.  The caller's data is not modified. */
int zntt_batch_autotune(zntt_batch *b, int reps, double *best_ms);
/* The batch's current shape.  level_kernel: 1 when passes with a full tile
 * run in the register-round level kernel (one launch per pass length and
 * parity of the first stage), 0 when every pass runs in the generic one. */
int zntt_batch_shape(const zntt_batch *b, int *tile_log, int *kmax, int *wgs_per_cu, int *order,
                     int *level_kernel);
/* Sets the batch's shape to one the autotuner could pick (tile_log the
 * field's tile size or one less: BN254 Fr 11 or 10, the other fields 12 or
 * 11; kmax tile_log - 2 or tile_log - 3; wgs_per_cu 1..4 within 160 KiB of
 * LDS per CU; order and level_kernel 0 or 1), e.g. one a previous run's
 * autotune reported, without timing.  ZNTT_INVALID_ARGUMENT for any other;
 * on a failure the batch keeps its previous shape.  No counterpart in the
 * reference, which has no batches. */
int zntt_batch_set_shape(zntt_batch *b, int tile_log, int kmax, int wgs_per_cu, int order, int level_kernel);
void zntt_batch_destroy(zntt_batch *b);

/* ------------------------------------------------------------------------ */
/* Serving loop for coset low-degree extensions.  No reference counterpart   */
/* (the reference's STARK example evaluates its LDE by Horner's rule,        */
/* examples/stark_prover.zig:117-129).  Requests are queued as they are      */
/* submitted, admitted while the HBM budget and the active-request limit     */
/* allow, and run by continuous batching: every scheduler step holds the     */
/* next pass of every admitted request, grouped into a few persistent        */
/* launches (requests that fit one tile by size; passes of larger requests   */
/* by pass length and parity, so large requests share launches; the rest     */
/* together), run as parallel branches of the server's stream.  A request    */
/* admitted while others are on the device joins the next step.  A first     */
/* pass reads only the request's 2^log_n coefficients and skips the          */
/* products of the zero padding (stages 1..log_blowup of the bit-reversed    */
/* input copy their u, up to blowup 8).                                      */
/* A request's result equals zntt_lde on the same coefficients.              */
/* ------------------------------------------------------------------------ */
typedef struct zntt_server zntt_server;

/* Device memory provider for request buffers (inputs the server generates,
 * outputs, scratch).  `alloc` returns a zntt status; `*out` must be 256-byte
 * aligned device memory.  Called from the server's scheduler. */
typedef struct {
  void *user;
  int (*alloc)(void *user, size_t bytes, void **out);
  void (*free)(void *user, void *ptr, size_t bytes);
} zntt_allocator;

typedef struct {
  int field;                        /* field of every request                       */
  uint64_t hbm_budget;              /* request memory limit in bytes (0: 64 GiB)     */
  int max_active;                   /* admitted requests at once (0: 256)            */
  int max_log_n;                    /* largest evaluation size 2^max_log_n the server
                                       prepares twiddles for at creation (0: none)   */
  int simulate;                     /* 1: no device work, for tests of the scheduling
                                       on a host without a GPU (ctx may be NULL)     */
  const zntt_allocator *allocator;  /* NULL: an arena of hbm_budget bytes owned by
                                       the server (a counter of fake addresses when
                                       simulating)                                   */
} zntt_server_config;

typedef struct {
  size_t log_n;        /* coefficients: 2^log_n                                      */
  size_t len;          /* number of coefficients; must equal 2^log_n                 */
  size_t log_blowup;   /* evaluations: 2^(log_n + log_blowup), at least 8            */
  const void *shift;   /* host pointer, coset shift (one element)                    */
  const void *coeffs;  /* device pointer to the coefficients, read by the first pass
                          and kept valid by the caller until the request is done;
                          NULL: the server generates them into its own memory from
                          the counter-based stream (seed, stream_id)               */
  uint64_t seed, stream_id;
  int keep_result;     /* 1: the evaluations stay in server memory until
                          zntt_server_release; 0: freed when done (timing only)     */
  void *out;           /* device pointer for the 2^(log_n+log_blowup) evaluations,
                          owned by the caller and not charged to the budget; NULL:
                          server memory (see keep_result)                          */
} zntt_lde_request;

enum zntt_request_state {
  ZNTT_REQ_QUEUED = 0,
  ZNTT_REQ_RUNNING = 1,
  ZNTT_REQ_DONE = 2,
  ZNTT_REQ_FAILED = 3,
  ZNTT_REQ_CANCELLED = 4
};

typedef struct {
  int state;           /* zntt_request_state                                */
  int status;          /* zntt_status of a failed request                   */
  double queued_ms;    /* submission to admission                           */
  double latency_ms;   /* submission to completion seen by the scheduler    */
  void *data;          /* device pointer to the evaluations (kept results)  */
  size_t log_n, log_blowup;
} zntt_request_info;

typedef struct {
  uint64_t submitted, admitted, completed, failed, cancelled;
  uint64_t steps;           /* scheduler steps launched                      */
  uint64_t elements_done;   /* evaluations produced                          */
  uint64_t bytes_in_use, peak_bytes;
  uint64_t queued, active;  /* requests now waiting / admitted               */
  double busy_ms;           /* device time of completed steps                */
} zntt_server_stats;

/* Starts a server on the context's device (a scheduler thread unless
 * simulating).  Twiddles for sizes up to 2^max_log_n and the coset table of
 * the field's generator are built up front and stay resident. */
int zntt_server_create(zntt_ctx *ctx, const zntt_server_config *cfg, zntt_server **out);
/* Queues a request.  Argument errors are returned here, in the order of
 * zntt_lde: LogTooLarge, LengthMismatch (len != 2^log_n), OrderTooLarge, and
 * ZNTT_UNSUPPORTED beyond ZNTT_MAX_LOG_N or below 2^3 evaluations.  A
 * request larger than the whole budget fails with OutOfMemory when its turn
 * for admission comes. */
int zntt_server_submit(zntt_server *s, const zntt_lde_request *r, uint64_t *ticket);
int zntt_server_query(zntt_server *s, uint64_t ticket, zntt_request_info *out);
/* Blocks until the request is done, failed or cancelled (timeout_ms < 0:
 * no limit); ZNTT_DEVICE_ERROR on timeout. */
int zntt_server_wait(zntt_server *s, uint64_t ticket, double timeout_ms);
/* A queued request is dropped at once; an admitted one runs to the end of
 * its passes and its memory is freed then.  Either ends CANCELLED. */
int zntt_server_cancel(zntt_server *s, uint64_t ticket);
/* Copies a finished request's evaluations (kept or caller-owned) to host
 * memory: `len` elements, which must be 2^(log_n+log_blowup)
 * (ZNTT_LENGTH_MISMATCH otherwise).  Do not release the request meanwhile.
 * Results over 16 MiB go through 64 MiB of page-locked host memory and a
 * stream the server makes at its first such read; reads from several
 * threads take turns there. */
int zntt_server_read(zntt_server *s, uint64_t ticket, void *host, size_t len);
/* Forgets a finished request and frees a kept result. */
int zntt_server_release(zntt_server *s, uint64_t ticket);
/* Blocks until no request is queued or admitted. */
int zntt_server_drain(zntt_server *s);
int zntt_server_stats_get(zntt_server *s, zntt_server_stats *out);
/* Simulation only: one non-blocking scheduler iteration (retire, admit,
 * launch one step); the number of steps then in flight. */
int zntt_server_sim_pump(zntt_server *s, int *in_flight);
/* Simulation only: marks the oldest `n` in-flight steps as completed. */
int zntt_server_sim_complete(zntt_server *s, int n);
/* Drains, stops the scheduler and frees everything, kept results included. */
void zntt_server_destroy(zntt_server *s);

/* Device fill with the counter-based stream that generated requests use:
 * element i of (seed, stream_id) for i < n, in the field's storage layout. */
int zntt_fill_counter(zntt_ctx *ctx, int field, uint64_t seed, uint64_t stream_id, void *data, size_t n,
                      void *stream);

/* ------------------------------------------------------------------------ */
/* Polynomial multiplication by NTT.  Replaces the O(n^2) schoolbook product */
/* of libs/poly (Polynomial.mul, poly.zig:197-212) with forward transforms   */
/* of both operands zero-padded to 2^L >= deg a + deg b + 1, a pointwise     */
/* product and an inverse transform; the cyclic product of that length is   */
/* the linear one, so the coefficients are exactly the schoolbook ones.      */
/* ------------------------------------------------------------------------ */

/* Polynomial.mul on coefficient arrays (host buffers, coefficient i of x^i,
 * storage layout).  Degrees are those of the arrays with trailing zero
 * coefficients removed, as the reference's normalised Polynomial holds them.
 * A zero operand gives the zero polynomial (*out_len = 0, poly.zig:198).
 * ZNTT_DEGREE_TOO_LARGE when deg a + deg b + 1 > out_cap (the reference's
 * max_degree + 1, poly.zig:200); ZNTT_ORDER_TOO_LARGE when the padded size
 * exceeds the field's two-adicity; ZNTT_UNSUPPORTED beyond ZNTT_MAX_LOG_N.
 * On success out[0 .. *out_len) holds the product. */
int zntt_poly_mul(zntt_ctx *ctx, int field, const void *a, size_t a_len, const void *b, size_t b_len, void *out,
                  size_t out_cap, size_t *out_len);

/* The same on device buffers, enqueued on `stream`: out[0 .. a_len + b_len - 1)
 * = a * b with the operands taken at their full lengths (trailing zeros give
 * zero coefficients).  Uses context scratch of 2 * 2^L elements. */
int zntt_poly_mul_device(zntt_ctx *ctx, int field, const void *a, size_t a_len, const void *b, size_t b_len,
                         void *out, void *stream);

/* ------------------------------------------------------------------------ */
/* Merkle commitment of an LDE / FRI layer (BLAKE3).                         */
/*                                                                           */
/* The FRI prover commits a layer f of n = 2^log_n evaluations (natural      */
/* domain order) by a Merkle tree over n/2 leaves                            */
/*   leaf j = BLAKE3(toBytes(f[j]) || toBytes(f[j + n/2]))                   */
/* (libs/fri/src/root.zig:283-292, :378-384), built as                       */
/* MerkleTree.initFromHashes (libs/merkle/src/merkle_tree.zig:147-182).      */
/* toBytes is the canonical value, little-endian: 32 bytes for BN254 Fr      */
/* (out of the Montgomery form), 8 for Goldilocks, 4 for BabyBear and        */
/* KoalaBear.  The tree is the reference's node array: 2 * cap nodes of 32   */
/* bytes, node 0 zero, node 1 the root, leaf i at node cap + i, BLAKE3("")   */
/* leaves past the last real one.                                            */
/* Device buffers here are 16-byte aligned (ZNTT_INVALID_ARGUMENT if not).   */
/* ------------------------------------------------------------------------ */

/* Commits one layer: `evals` is a device buffer of 2^log_n elements in the
 * storage layout, 1 <= log_n <= ZNTT_MAX_LOG_N.  `nodes` is null or a device
 * buffer of 2^log_n * 32 bytes that receives the whole tree (cap = n/2).  The
 * root's 32 bytes are written to host memory `root` once the work enqueued
 * on `stream` has finished (the call waits for it).  Replaces the commit of
 * fri/src/root.zig:378-384 (leaves + initFromHashes + tree.root()). */
int zntt_merkle_commit(zntt_ctx *ctx, int field, const void *evals, size_t log_n, void *nodes, void *root,
                       void *stream);

/* MerkleTree.initFromHashes over `count` >= 1 leaf hashes (device buffer of
 * count * 32 bytes); `nodes` null or a device buffer of 2 * cap * 32 bytes,
 * cap = count rounded up to a power of two (at most 2^ZNTT_MAX_LOG_N).  The
 * root goes to host memory `root` as above (merkle_tree.zig:147-182, :189). */
int zntt_merkle_from_hashes(zntt_ctx *ctx, const void *hashes, size_t count, void *nodes, void *root, void *stream);

/* ------------------------------------------------------------------------ */
/* FRI prover: the LDE's consumer (libs/fri/src/root.zig:311-441, the query  */
/* openings :506-573).  Each round commits the current layer (the Merkle     */
/* commitment above), absorbs the root, squeezes alpha with                  */
/* challengeFieldChecked and folds on the device; the last layer is          */
/* interpolated by an inverse transform on the device, its first             */
/* 2^log_residual_degree coefficients absorbed; then num_queries indices are */
/* drawn with challengeU64 and the pair values and Merkle paths gathered     */
/* from the device.  The transcript stays the caller's: the library calls    */
/* back into it, so a Zig shim passes its own zig-transcript object.         */
/* ------------------------------------------------------------------------ */

/* Config (root.zig:174-203). */
typedef struct {
  size_t log_domain;
  size_t log_initial_degree;
  size_t log_final;
  size_t log_residual_degree;
  size_t num_queries;
} zntt_fri_config;

/* A transcript as the prover uses it: absorbBytes and challengeBytes
 * (libs/transcript/src/root.zig:55-60, :93-114).  challengeU64 and
 * challengeFieldChecked are built from challenge_bytes as the reference
 * builds them (:117-121, :167-180). */
typedef struct {
  void (*absorb_bytes)(void *transcript, const uint8_t *data, size_t len);
  void (*challenge_bytes)(void *transcript, uint8_t *out, size_t len);
} zntt_transcript_ops;

/* The Proof (root.zig:207-245) in caller-allocated host arrays, sized by
 * zntt_fri_proof_sizes (R = log_domain - log_final rounds, B = the field's
 * toBytes length: 32 Fr, 8 Goldilocks, 4 BabyBear / KoalaBear):
 *   roots       R x 32          layers[r].merkle_root (log_size = log_domain - r)
 *   residual    2^log_residual_degree x B   the residual's serializeElem bytes
 *   pair_index  num_queries     queries[q].pair_index
 *   values      num_queries x R x 2B        queries[q].values[r] (serializePair)
 *   paths       num_queries x sum_r (log_domain - r - 1) x 32, query by query,
 *               round by round, each path leaf-sibling first (MerkleTree.prove). */
typedef struct {
  uint8_t *roots;
  uint8_t *residual;
  uint64_t *pair_index;
  uint8_t *values;
  uint8_t *paths;
} zntt_fri_proof;

int zntt_fri_proof_sizes(int field, const zntt_fri_config *cfg, size_t *roots_bytes, size_t *residual_bytes,
                         size_t *values_bytes, size_t *paths_bytes);

/* prove(F, allocator, transcript, evals, config) -- root.zig:311-332.
 * `evals`: device buffer of evals_len values (storage layout, natural domain
 * order, 16-byte aligned).  Errors, in the reference's order:
 * ZNTT_INVALID_PARAMETERS for an invalid config, evals_len != 2^log_domain or
 * log_domain above the field's two-adicity; ZNTT_UNSUPPORTED beyond
 * ZNTT_MAX_LOG_N.  Work runs on `stream`; the call returns when the proof is
 * written.  Device memory: about 2^log_domain x (value + 32) bytes for the
 * folded layers and the trees, context scratch kept for the next call
 * (zntt_ctx_trim releases it). */
int zntt_fri_prove(zntt_ctx *ctx, int field, const zntt_fri_config *cfg, const zntt_transcript_ops *ops,
                   void *transcript, const void *evals, size_t evals_len, zntt_fri_proof *proof, void *stream);

/* The library's own transcript, byte for byte zig-transcript's
 * (Transcript.init / absorbBytes / challengeBytes).  Its two functions match
 * zntt_transcript_ops, so {zntt_transcript_absorb_bytes,
 * zntt_transcript_challenge_bytes} with a zntt_transcript* serves as the
 * prover's transcript. */
typedef struct zntt_transcript zntt_transcript;
int zntt_transcript_create(const void *label, size_t len, zntt_transcript **out);
void zntt_transcript_destroy(zntt_transcript *t);
void zntt_transcript_absorb_bytes(void *t, const uint8_t *data, size_t len);
void zntt_transcript_challenge_bytes(void *t, uint8_t *out, size_t len);

/* BLAKE3 with a 32-byte output (the hash of the transcript and the trees),
 * exposed for checks. */
void zntt_blake3(const void *data, size_t len, uint8_t *out);

/* ------------------------------------------------------------------------ */
/* Eight-lane Mersenne-31 transform (Vec8NttM31, libs/field/src/lib.zig:     */
/* 39-137).  `data` holds eight interleaved transforms of 2^log_n M31 values, */
/* data[i * 8 + lane], each value a canonical u64 (the reference's M31,      */
/* p = 2^31 - 1).  Each lane gets zig-ntt's radix-2 DIT DAG with `root`      */
/* (M31's two-adicity is 1, so for log_n >= 2 no root is primitive; the DAG  */
/* is computed for whatever root is given, as the reference does).           */
/* ------------------------------------------------------------------------ */

/* Largest log_n the device path implements (8 x 2^25 values = 2 GiB). */
#define ZNTT_VEC8_MAX_LOG_N 25

/* nttVec8M31Checked(data, log_n, root) -- lib.zig:52-106 (host buffer, in
 * place, synchronous).  ZNTT_INVALID_LENGTH when log_n >= 61 or
 * len != 8 * 2^log_n, before `data` is touched (lib.zig:53-55, 131-136);
 * ZNTT_UNSUPPORTED beyond ZNTT_VEC8_MAX_LOG_N.  The reference's unchecked
 * nttVec8M31 is the same call with the status ignored. */
int zntt_ntt_vec8_m31(zntt_ctx *ctx, void *data, size_t len, size_t log_n, const void *root);

/* inttVec8M31Checked -- lib.zig:114-129: the forward DAG with root^-1 (zero
 * for a zero root), then every value times (2^log_n)^-1. */
int zntt_intt_vec8_m31(zntt_ctx *ctx, void *data, size_t len, size_t log_n, const void *root);

/* The same on a device buffer of 8 * 2^log_n values, enqueued on `stream`;
 * returns when the transform is done.  inverse: 0 forward, 1 inverse. */
int zntt_vec8_m31_device(zntt_ctx *ctx, int inverse, void *data, size_t log_n, const void *root, void *stream);

/* Blocks until every call enqueued on the context's own stream is done. */
int zntt_ctx_synchronize(zntt_ctx *ctx);

/* The context's own HIP stream (a hipStream_t). */
void *zntt_ctx_stream(zntt_ctx *ctx);

/* ------------------------------------------------------------------------ */
/* Diagnostics for tests and the benchmark.                                  */
/* ------------------------------------------------------------------------ */

/* Elementwise field operation on device: out[i] = op(a[i], b[i]) for the
 * library's device arithmetic (op 0 add, 1 sub, 2 mul), host buffers of n
 * elements in the field's storage layout.  Used to check the device field
 * code against an independent implementation. */
int zntt_debug_field_op(zntt_ctx *ctx, int field, int op, const void *a, const void *b,
                        void *out, size_t n);

/* Seeded uniform elements in the field's storage layout (host memory): a
 * SplitMix64-seeded xoshiro256++ stream; each draw takes 32 (Fr) or 8
 * (Goldilocks) little-endian bytes and is rejected if >= the modulus, as the
 * reference's `random` does (libs/field/src/field.zig:982-990, :224-232); Fr
 * values are then put in Montgomery form.  The benchmark's input generator. */
int zntt_fill_uniform(int field, uint64_t seed, void *out, size_t n);

/* Measured rate of the device's v_mad_u64_u32 (32x32->64 multiply-add), in
 * multiplies per second, from an unrolled independent-chain microbenchmark;
 * and the measured device-to-device copy bandwidth in bytes per second. */
int zntt_probe_peaks(zntt_ctx *ctx, double *mad_per_s, double *copy_bytes_per_s);

#ifdef __cplusplus
}
#endif

#endif /* ZNTT_GPU_H */
