// Stage split of a transform into LDS passes (host code, shared by the
// single-transform launchers, the batch planner and the serving loop).
#include "kernels.hpp"

namespace zntt {

// Elements per workgroup tile by field: BN254 Fr 2^11 (72 KiB of LDS in nine
// u32 planes), the other fields 2^12 (32 / 16 KiB).
static int tile_log_for(int field) { return field == kFr ? 11 : 12; }

PassPlan plan_passes(int field, int L) {
  const int tl = tile_log_for(field);
  // BN254 Fr: 2^10 tiles (256 threads, four workgroups a CU instead of two)
  // whenever passes of at most 8 stages need no more passes than passes of 9
  // (L = 12..16, 19..24, 28): the 2^24 forward takes 1.76 ms against 1.81 with
  // 2^11 tiles on one box, 1.85 against 1.92 on another.
  if (field == kFr && L > tl && (L + 7) / 8 == (L + 8) / 9) return plan_passes_shape(L, tl - 1, tl - 3);
  return plan_passes_shape(L, tl, tl - 2);  // keep >= 4 groups per tile for wide global accesses
}

PassPlan plan_passes_shape(int L, int tl, int kmax) {
  PassPlan p = {};
  p.tile_log = tl;
  if (L <= tl) {
    p.passes = 1;
    p.K[0] = L;
    p.log_g[0] = 0;
    return p;
  }
  const int P = (L + kmax - 1) / kmax;
  p.passes = P;
  int s0 = 1;
  for (int i = 0; i < P; i++) {
    const int K = L / P + (i < L % P ? 1 : 0);
    p.K[i] = K;
    const int room = tl - K;
    const int avail = (i == 0) ? (L - K) : (s0 - 1);
    p.log_g[i] = room < avail ? room : avail;
    s0 += K;
  }
  return p;
}

}  // namespace zntt
