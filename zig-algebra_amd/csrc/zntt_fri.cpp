// C ABI of the FRI prover (libs/fri/src/root.zig:311-441, queries :506-573)
// and of the library's own Fiat-Shamir transcript
// (libs/transcript/src/root.zig:38-121).
//
// Device: each round's Merkle commitment (merkle_kernels.hip), each fold
// (fri_kernels.hip), the residual's inverse transform and the query
// openings (values and paths, one launch).  Host: the transcript (through the caller's
// callbacks), the challenge decoding, and the byte layout of the proof.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <new>
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

  // Context memory (kept between calls, grown when a larger proof needs
  // it): layers 1..R (layer 0 is the caller's buffer), the R trees, whose
  // node arrays hold n_r x 32 bytes each, the query seeds and the opened
  // values and paths.  Page-locked host memory stages the roots, the
  // residual, the seeds and the openings.
  const size_t q = cfg->num_queries, depth = path_entries(cfg, rounds);
  const size_t nvals = q * rounds * 2, npaths = q * depth;
  size_t lbytes = 0, tbytes = 0;
  for (size_t r = 0; r < rounds; r++) {
    lbytes += (n >> (r + 1)) * ub;
    tbytes += (n >> r) * 32;
  }
  const size_t rl = (size_t)1 << cfg->log_residual_degree;
  const size_t ebytes = (q * 8 + 15) & ~(size_t)15, obytes = nvals * ub + npaths * 32;
  if ((st = ensure(c, &c->aux, &c->aux_bytes, lbytes + tbytes + ebytes + obytes))) return st;
  const size_t hbytes = ebytes + std::max(obytes, std::max(rl * ub, (size_t)32));
  if ((st = ensure_pinned(c, hbytes))) return st;
  uint8_t *pin = (uint8_t *)c->pinned, *pin_out = pin + ebytes;
  std::vector<uint8_t *> layer(rounds + 1), tree(rounds);
  layer[0] = (uint8_t *)evals;
  uint8_t *mem = (uint8_t *)c->aux;
  for (size_t r = 0; r < rounds; r++) {
    layer[r + 1] = mem;
    mem += (n >> (r + 1)) * ub;
  }
  for (size_t r = 0; r < rounds; r++) {
    tree[r] = mem;
    mem += (n >> r) * 32;
  }
  uint64_t *d_seeds = (uint64_t *)mem;
  uint8_t *d_out = mem + ebytes;

  // Domain generators: g_r (order 2^(ld - r)) is the primitive root of that
  // order, and g_{r+1} = g_r^2 (Domain.init, root.zig:133-138), so one root
  // and one inversion serve every round.
  Elem ginv = f_inv(field, f_primitive_root(field, (int)ld));
  // Round r commits layer r (root.zig:378-384), absorbs its root and
  // squeezes alpha (:387-389), then folds layer r into layer r + 1
  // (:392-403).  The fold runs inside the launch that hashes layer r + 1's
  // leaves, so each layer after the first is committed by the launch that
  // produces it; only the last fold, whose layer is not committed, runs
  // alone.  The fold reads g^-j from the table of powers of the inverse of
  // the layer's domain generator.
  hipError_t e = zntt::launch_merkle_layer(field, layer[0], (int)ld - 1, (uint32_t *)tree[0], true, s);
  if (e != hipSuccess) return hip_status(e);
  for (size_t r = 0; r < rounds; r++) {
    const size_t half = n >> (r + 1);
    uint8_t *root = proof->roots + 32 * r;
    e = hipMemcpyAsync(pin_out, tree[r] + 32, 32, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_status(e);
    memcpy(root, pin_out, 32);
    ops->absorb_bytes(tr, root, 32);
    const Elem alpha = challenge_field(ops, tr, field);
    const uint32_t *tab = nullptr;
    if ((st = coset_table(c, field, ginv, half, s, &tab))) return st;
    ginv = f_mul(field, ginv, ginv);
    const std::array<uint32_t, 9> alpha_dev = f_to_dev(field, alpha);
    if (r + 1 < rounds)
      e = zntt::launch_merkle_fold(field, layer[r], layer[r + 1], (int)(ld - r) - 2, (uint32_t *)tree[r + 1], tab,
                                   alpha_dev.data(), s);
    else
      e = zntt::launch_fri_fold(field, layer[r], layer[r + 1], half, tab, alpha_dev.data(), s);
    if (e != hipSuccess) return hip_status(e);
  }

  // Residual (:407-428): the final layer's coefficients by an inverse
  // transform over its domain -- the unique solution of the Vandermonde
  // system interpolateToCoeffs eliminates -- then the first
  // 2^log_residual_degree of them, serialised and absorbed in order.
  const int lf = (int)cfg->log_final;
  const size_t m = (size_t)1 << lf;
  {
    TwiddleEntry *inv = nullptr;
    if ((st = power_table(c, field, ginv, lf, s, &inv))) return st;  // g_R^-1, R = ld - lf
    const std::array<uint32_t, 9> scale = f_to_dev(field, f_inv(field, f_from_u64(field, m)));
    if ((st = run_device(c, field, lf, layer[rounds], inv->table, nullptr, &scale, s))) return st;
    e = hipMemcpyAsync(pin_out, layer[rounds], rl * ub, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_status(e);
    for (size_t i = 0; i < rl; i++) {
      to_bytes(field, pin_out + i * ub, proof->residual + i * nb);
      ops->absorb_bytes(tr, proof->residual + i * nb, nb);
    }
  }

  // Queries (buildQueries, :546-570): draw every pair index, then open all
  // of them on the device in one launch (the slot of round r is the index
  // modulo the round's half size).
  uint64_t *seeds = (uint64_t *)pin;
  for (size_t k = 0; k < q; k++) {
    const uint64_t seed = challenge_u64(ops, tr);
    proof->pair_index[k] = seeds[k] = seed % (uint64_t)n;
  }
  zntt::FriOpenArgs oa = {};
  for (size_t r = 0; r < rounds; r++) {
    oa.layer[r] = layer[r];
    oa.tree[r] = tree[r];
  }
  oa.rounds = (uint32_t)rounds;
  oa.log_domain = (uint32_t)ld;
  oa.elem_words = (uint32_t)(ub / 4);
  e = hipMemcpyAsync(d_seeds, seeds, q * 8, hipMemcpyHostToDevice, s);
  if (e == hipSuccess)
    e = zntt::launch_fri_open(oa, d_seeds, q, (uint32_t *)d_out, (uint32_t *)(d_out + nvals * ub), s);
  if (e == hipSuccess) e = hipMemcpyAsync(pin_out, d_out, obytes, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return hip_status(e);
  for (size_t i = 0; i < nvals; i++) to_bytes(field, pin_out + i * ub, proof->values + i * nb);
  if (npaths) memcpy(proof->paths, pin_out + nvals * ub, npaths * 32);
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
  CallGuard g(c, (hipStream_t)stream);
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
