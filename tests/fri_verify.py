"""The FRI verifier, restated in Python integers from verifyOn
(libs/fri/src/root.zig:603-718), for checking proofs in tests.

Proof layout: the dict of flat byte strings that the oracle's fri_prove and
the library's Context.fri_prove return (roots, residual, pair_index,
values, paths; see include/zntt_gpu.h, zntt_fri_proof).
"""
import oracle as O

_PRIM = {}


def prim(field, log_n):
    """Generator of the FRI domain of order 2^log_n (Domain.init, root.zig:133-138)."""
    if (field, log_n) not in _PRIM:
        _, r = O.primitive_root(field, log_n)
        _PRIM[(field, log_n)] = O.value(field, r)
    return _PRIM[(field, log_n)]


def verify(field, label, proof, ld, li, lf, lr, nq):
    """verifyOn restated in Python integers; True when the proof verifies."""
    p = O.MODULUS[field]
    nb = O.NUM_BYTES[field]
    rounds = ld - lf
    n = 1 << ld
    t = O.Transcript(label)
    roots = [proof["roots"][32 * r:32 * r + 32] for r in range(rounds)]
    alphas = []
    for r in range(rounds):
        t.absorb_bytes(roots[r])
        alphas.append(O.value(field, t.challenge_field(field)))
    rl = 1 << lr
    res = [int.from_bytes(proof["residual"][i * nb:(i + 1) * nb], "little") for i in range(rl)]
    for c in res:
        t.absorb_bytes(c.to_bytes(nb, "little"))
    gf = prim(field, lf)
    final = []
    for i in range(1 << lf):
        x = pow(gf, i, p)
        final.append(sum(c * pow(x, k, p) for k, c in enumerate(res)) % p)
    half_inv = pow(2, -1, p)
    vals = proof["values"]
    paths = proof["paths"]
    pp = 0
    for q in range(nq):
        seed = t.challenge_u64()
        if proof["pair_index"][q] != seed % n:
            return False
        e = proof["pair_index"][q]
        for r in range(rounds):
            layer_log = ld - r
            half_r = (1 << layer_log) // 2
            j = e % half_r
            off = (q * rounds + r) * 2 * nb
            xb, nxb = vals[off:off + nb], vals[off + nb:off + 2 * nb]
            x, negx = int.from_bytes(xb, "little"), int.from_bytes(nxb, "little")
            cur = O.blake3(xb + nxb)
            idx = j
            for _ in range(layer_log - 1):
                sib = paths[pp:pp + 32]
                pp += 32
                cur = O.blake3(cur + sib) if idx % 2 == 0 else O.blake3(sib + cur)
                idx //= 2
            if cur != roots[r]:
                return False
            xv = pow(prim(field, layer_log), j, p)
            even = (x + negx) * half_inv % p
            odd = (x - negx) * pow(xv, -1, p) * half_inv % p
            expect = (even + alphas[r] * odd) % p
            if r + 1 < rounds:
                noff = (q * rounds + r + 1) * 2 * nb
                slot = 0 if j < half_r // 2 else 1
                child = int.from_bytes(vals[noff + slot * nb:noff + (slot + 1) * nb], "little")
                if expect != child:
                    return False
            elif expect != final[j]:
                return False
            e = j
    return True
