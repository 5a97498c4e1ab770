// C ABI of the FRI prover (libs/fri/src/root.zig:311-441, queries :506-573)
// and of the library's own Fiat-Shamir transcript
// (libs/transcript/src/root.zig:38-121).
//
// Device: each round's Merkle commitment (merkle_kernels.hip), each fold
// (fri_kernels.hip), the residual's inverse transform and the gathers of
// the opened values and paths.  Host: the transcript (through the caller's
// callbacks), the challenge decoding, and the byte layout of the proof.
#include <hip/hip_runtime.h>

#include <vector>

#include "ctx_internal.hpp"
#include "host_blake3.hpp"

using namespace zntt_detail;
using zntt::host::U256;

struct zntt_transcript {
  std::vector<uint8_t> stream;  // bytes fed since the last re-key
};

namespace {

using Fr = zntt::host::Fr;

// toBytes length (field.zig:108, :1002): 32 BN254 Fr, 8 Goldilocks, 4 the
// 31-bit fields.
size_t num_bytes(int field) {
  if (field == ZNTT_FIELD_BN254_FR) return 32;
  return field == ZNTT_FIELD_GOLDILOCKS ? 8 : 4;
}

uint64_t modulus_u64(int field) {
  if (field == ZNTT_FIELD_GOLDILOCKS) return zntt::host::Gl::P;
  return field == ZNTT_FIELD_BABYBEAR ? zntt::host::kBabyBear.p : zntt::host::kKoalaBear.p;
}

int two_adicity_of(int field) { return two_adicity(field); }

// toBytes of a storage-layout element: the canonical value, little-endian.
void to_bytes(int field, const uint8_t *elem, uint8_t *out) {
  if (field == ZNTT_FIELD_BN254_FR) {
    U256 x;
    memcpy(x.l, elem, 32);
    x = Fr::to_canonical(x);
    memcpy(out, x.l, 32);
    return;
  }
  memcpy(out, elem, num_bytes(field));  // little-endian u64, low bytes
}

void le_feed(std::vector<uint8_t> &s, uint64_t v) {
  for (int i = 0; i < 8; i++) s.push_back((uint8_t)(v >> (8 * i)));
}

// challengeU64 (transcript root.zig:117-121).
uint64_t challenge_u64(const zntt_transcript_ops *ops, void *t) {
  uint8_t b[8];
  ops->challenge_bytes(t, b, 8);
  uint64_t v = 0;
  for (int i = 0; i < 8; i++) v |= (uint64_t)b[i] << (8 * i);
  return v;
}

// challengeFieldChecked (transcript root.zig:167-180): NUM_BYTES bytes per
// draw, accepted when they encode a value below the modulus
// (fromBytesChecked, field.zig:217-221 and :975-979).
Elem challenge_field(const zntt_transcript_ops *ops, void *t, int field) {
  const size_t nb = num_bytes(field);
  for (;;) {
    uint8_t b[32];
    ops->challenge_bytes(t, b, nb);
    if (field == ZNTT_FIELD_BN254_FR) {
      U256 v;
      memcpy(v.l, b, 32);
      if (!Fr::lt(v, Fr::P)) continue;
      const U256 m = Fr::from_canonical(v);
      Elem e = {};
      memcpy(e.w, m.l, 32);
      return e;
    }
    uint64_t v = 0;
    for (size_t i = 0; i < nb; i++) v |= (uint64_t)b[i] << (8 * i);
    if (v >= modulus_u64(field)) continue;
    Elem e = {};
    memcpy(e.w, &v, 8);
    return e;
  }
}

// Config.rounds + Config.validate (root.zig:188-203).
int fri_rounds(const zntt_fri_config *c, size_t *rounds) {
  if (c->log_domain == 0 || c->log_domain > 63) return ZNTT_INVALID_PARAMETERS;
  if (c->log_final >= c->log_domain) return ZNTT_INVALID_PARAMETERS;
  if (c->log_residual_degree >= c->log_final) return ZNTT_INVALID_PARAMETERS;
  if (c->log_final > 12 || c->log_residual_degree > 12) return ZNTT_INVALID_PARAMETERS;
  if (c->num_queries == 0) return ZNTT_INVALID_PARAMETERS;
  const size_t r = c->log_domain - c->log_final;
  if (c->log_initial_degree > c->log_domain) return ZNTT_INVALID_PARAMETERS;
  if (c->log_initial_degree < r) return ZNTT_INVALID_PARAMETERS;
  if (c->log_initial_degree - r != c->log_residual_degree) return ZNTT_INVALID_PARAMETERS;
  *rounds = r;
  return ZNTT_OK;
}

size_t path_entries(const zntt_fri_config *c, size_t rounds) {
  size_t d = 0;
  for (size_t r = 0; r < rounds; r++) d += c->log_domain - r - 1;
  return d;
}

// Device buffers of one prove call, released on every exit.
struct Buffers {
  zntt_ctx *c;
  std::vector<std::pair<void *, size_t>> held;
  explicit Buffers(zntt_ctx *ctx) : c(ctx) {}
  ~Buffers() {
    for (auto &h : held) dev_free(c, h.first, h.second);
  }
  int get(size_t bytes, void **out) {
    int st = dev_alloc(c, bytes, out);
    if (st == ZNTT_OK) held.emplace_back(*out, bytes);
    return st;
  }
};

bool aligned16(const void *p) { return ((uintptr_t)p & 15) == 0; }

int prove(zntt_ctx *c, int field, const zntt_fri_config *cfg, const zntt_transcript_ops *ops, void *tr,
          const void *evals, size_t evals_len, zntt_fri_proof *proof, hipStream_t s) {
  size_t rounds;
  int st = fri_rounds(cfg, &rounds);
  if (st) return st;
  const size_t ld = cfg->log_domain;
  if (evals_len != ((size_t)1 << ld)) return ZNTT_INVALID_PARAMETERS;
  if (ld > (size_t)two_adicity_of(field)) return ZNTT_INVALID_PARAMETERS;
  if (ld > ZNTT_MAX_LOG_N) return ZNTT_UNSUPPORTED;
  const size_t n = evals_len, ub = user_bytes(field), nb = num_bytes(field);

  // Layers 1..R (layer 0 is the caller's buffer) and the R trees, whose
  // node arrays hold n_r x 32 bytes each.
  Buffers buf(c);
  std::vector<uint8_t *> layer(rounds + 1), tree(rounds);
  layer[0] = (uint8_t *)evals;
  {
    size_t lbytes = 0, tbytes = 0;
    for (size_t r = 0; r < rounds; r++) {
      lbytes += (n >> (r + 1)) * ub;
      tbytes += (n >> r) * 32;
    }
    void *lp = nullptr, *tp = nullptr;
    if ((st = buf.get(lbytes, &lp)) || (st = buf.get(tbytes, &tp))) return st;
    uint8_t *l = (uint8_t *)lp, *t = (uint8_t *)tp;
    for (size_t r = 0; r < rounds; r++) {
      layer[r + 1] = l;
      l += (n >> (r + 1)) * ub;
      tree[r] = t;
      t += (n >> r) * 32;
    }
  }

  const Elem half_inv = f_inv(field, f_from_u64(field, 2));
  const std::array<uint32_t, 9> half_dev = f_to_dev(field, half_inv);
  for (size_t r = 0; r < rounds; r++) {
    const int log_cur = (int)(ld - r);
    const size_t half = n >> (r + 1);
    // Commit (root.zig:378-384).
    hipError_t e = zntt::launch_merkle_layer(field, layer[r], log_cur - 1, (uint32_t *)tree[r], true, s);
    uint8_t *root = proof->roots + 32 * r;
    if (e == hipSuccess) e = hipMemcpyAsync(root, tree[r] + 32, 32, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_status(e);
    // Absorb the root, squeeze alpha (:387-389).
    ops->absorb_bytes(tr, root, 32);
    const Elem alpha = challenge_field(ops, tr, field);
    // Fold (:392-403) with g^-j from the table of powers of the inverse of
    // the layer's domain generator (Domain.init: omega^(2^(adicity - k)),
    // the primitive root of order 2^k).
    const uint32_t *tab = nullptr;
    const Elem ginv = f_inv(field, f_primitive_root(field, log_cur));
    if ((st = coset_table(c, field, ginv, half, s, &tab))) return st;
    const std::array<uint32_t, 9> ah_dev = f_to_dev(field, f_mul(field, alpha, half_inv));
    e = zntt::launch_fri_fold(field, layer[r], layer[r + 1], half, tab, ah_dev.data(), half_dev.data(), s);
    if (e != hipSuccess) return hip_status(e);
  }

  // Residual (:407-428): the final layer's coefficients by an inverse
  // transform over its domain -- the unique solution of the Vandermonde
  // system interpolateToCoeffs eliminates -- then the first
  // 2^log_residual_degree of them, serialised and absorbed in order.
  const int lf = (int)cfg->log_final;
  const size_t m = (size_t)1 << lf, rl = (size_t)1 << cfg->log_residual_degree;
  {
    TwiddleEntry *inv = nullptr;
    const Elem gf = f_primitive_root(field, lf);
    if ((st = power_table(c, field, f_inv(field, gf), lf, s, &inv))) return st;
    const std::array<uint32_t, 9> scale = f_to_dev(field, f_inv(field, f_from_u64(field, m)));
    if ((st = run_device(c, field, lf, layer[rounds], inv->table, nullptr, &scale, s))) return st;
    std::vector<uint8_t> coef(rl * ub);
    hipError_t e = hipMemcpyAsync(coef.data(), layer[rounds], rl * ub, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_status(e);
    for (size_t i = 0; i < rl; i++) {
      to_bytes(field, coef.data() + i * ub, proof->residual + i * nb);
      ops->absorb_bytes(tr, proof->residual + i * nb, nb);
    }
  }

  // Queries (buildQueries, :546-570): draw every index first, then gather the
  // opened pairs and the Merkle paths from the device in two launches.
  const size_t q = cfg->num_queries, depth = path_entries(cfg, rounds);
  std::vector<uint64_t> addr;
  addr.reserve(q * rounds * 2 + q * depth);
  std::vector<size_t> js(q * rounds);
  for (size_t k = 0; k < q; k++) {
    const uint64_t seed = challenge_u64(ops, tr);
    size_t idx = (size_t)(seed % (uint64_t)n);
    proof->pair_index[k] = idx;
    for (size_t r = 0; r < rounds; r++) {
      const size_t half = n >> (r + 1), j = idx % half;
      js[k * rounds + r] = j;
      addr.push_back((uint64_t)(uintptr_t)(layer[r] + j * ub));
      addr.push_back((uint64_t)(uintptr_t)(layer[r] + (j + half) * ub));
      idx = j;
    }
  }
  const size_t nvals = addr.size();
  for (size_t k = 0; k < q; k++)
    for (size_t r = 0; r < rounds; r++)
      for (size_t pos = (n >> (r + 1)) + js[k * rounds + r]; pos > 1; pos >>= 1)
        addr.push_back((uint64_t)(uintptr_t)(tree[r] + 32 * (pos ^ 1)));
  const size_t npaths = addr.size() - nvals;
  void *d_addr = nullptr, *d_out = nullptr;
  const size_t out_bytes = nvals * ub + npaths * 32;
  if ((st = buf.get(addr.size() * 8, &d_addr)) || (st = buf.get(out_bytes, &d_out))) return st;
  std::vector<uint8_t> host(out_bytes);
  hipError_t e = hipMemcpyAsync(d_addr, addr.data(), addr.size() * 8, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = zntt::launch_gather((const uint64_t *)d_addr, nvals, (int)(ub / 4), (uint32_t *)d_out, s);
  if (e == hipSuccess)
    e = zntt::launch_gather((const uint64_t *)d_addr + nvals, npaths, 8, (uint32_t *)((uint8_t *)d_out + nvals * ub),
                            s);
  if (e == hipSuccess) e = hipMemcpyAsync(host.data(), d_out, out_bytes, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return hip_status(e);
  for (size_t i = 0; i < nvals; i++) to_bytes(field, host.data() + i * ub, proof->values + i * nb);
  if (npaths) memcpy(proof->paths, host.data() + nvals * ub, npaths * 32);
  return ZNTT_OK;
}

}  // namespace

extern "C" {

int zntt_fri_proof_sizes(int field, const zntt_fri_config *cfg, size_t *roots_bytes, size_t *residual_bytes,
                         size_t *values_bytes, size_t *paths_bytes) {
  if (!field_ok(field) || !cfg || !roots_bytes || !residual_bytes || !values_bytes || !paths_bytes)
    return ZNTT_INVALID_ARGUMENT;
  size_t rounds;
  int st = fri_rounds(cfg, &rounds);
  if (st) return st;
  const size_t nb = num_bytes(field);
  *roots_bytes = rounds * 32;
  *residual_bytes = ((size_t)1 << cfg->log_residual_degree) * nb;
  *values_bytes = cfg->num_queries * rounds * 2 * nb;
  *paths_bytes = cfg->num_queries * path_entries(cfg, rounds) * 32;
  return ZNTT_OK;
}

int zntt_fri_prove(zntt_ctx *c, int field, const zntt_fri_config *cfg, const zntt_transcript_ops *ops,
                   void *transcript, const void *evals, size_t evals_len, zntt_fri_proof *proof, void *stream) {
  if (!c || !field_ok(field) || !cfg || !ops || !ops->absorb_bytes || !ops->challenge_bytes || !evals || !proof ||
      !aligned16(evals))
    return ZNTT_INVALID_ARGUMENT;
  if (!proof->roots || !proof->residual || !proof->pair_index || !proof->values || !proof->paths)
    return ZNTT_INVALID_ARGUMENT;
  CallGuard g(c);
  return prove(c, field, cfg, ops, transcript, evals, evals_len, proof, (hipStream_t)stream);
}

int zntt_transcript_create(const void *label, size_t len, zntt_transcript **out) {
  if (!out || (!label && len)) return ZNTT_INVALID_ARGUMENT;
  auto *t = new (std::nothrow) zntt_transcript;
  if (!t) return ZNTT_OUT_OF_MEMORY;
  // Transcript.init: the label's length (u64 little-endian), then the label.
  le_feed(t->stream, len);
  const uint8_t *p = (const uint8_t *)label;
  t->stream.insert(t->stream.end(), p, p + len);
  *out = t;
  return ZNTT_OK;
}

void zntt_transcript_destroy(zntt_transcript *t) { delete t; }

void zntt_transcript_absorb_bytes(void *tv, const uint8_t *data, size_t len) {
  auto *t = static_cast<zntt_transcript *>(tv);
  le_feed(t->stream, len);
  t->stream.insert(t->stream.end(), data, data + len);
}

void zntt_transcript_challenge_bytes(void *tv, uint8_t *out, size_t len) {
  auto *t = static_cast<zntt_transcript *>(tv);
  uint8_t d[32];
  zntt::host::Blake3::hash(t->stream.data(), t->stream.size(), d);
  for (size_t off = 0; off < len;) {
    const size_t k = len - off < 32 ? len - off : 32;
    memcpy(out + off, d, k);
    off += k;
    if (off < len) zntt::host::Blake3::hash(d, 32, d);
  }
  // Re-key: the stream restarts as the last digest.
  t->stream.assign(d, d + 32);
}

void zntt_blake3(const void *data, size_t len, uint8_t *out) {
  zntt::host::Blake3::hash((const uint8_t *)data, len, out);
}

}  // extern "C"
