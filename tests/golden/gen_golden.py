#!/usr/bin/env python3
"""Regenerate the committed golden fixtures under tests/golden/.

Run in the build container (not on a GPU host):  python tests/golden/gen_golden.py

Two kinds of fixtures are written:

1. Vectors the reference's own tests hold, extracted as data from the
   zig-algebra sources (read as text from /root/reference, which exists only in
   the build container):
     * ``ref_ntt_kat.json``  -- the DFT known-answer cases of
       libs/ntt/src/root.zig:438-449 (Goldilocks and BabyBear, log_n 1..4);
     * ``ref_field_vectors.json`` -- the add/mul/inv/neg vectors of
       libs/field/src/lib.zig:257-276 (BabyBear and Goldilocks rows);
     * ``ref_fri_kat.json`` -- the transcript challenge sequence of
       libs/fri/src/root.zig:1210-1214 (label "wire-compat", four Goldilocks
       draws, toBytes concatenated) and the Goldilocks FRI domain of
       :1272 (log_n 8, the 256 elements in natural order);
     * ``ref_blake3_kat.json`` -- the BLAKE3 known answers of
       libs/hash/src/root.zig:121-127 ("hello world") and :173-247 (45
       lengths 0..4097 of a repeating 16-byte pattern).
2. Python-int vectors computed from definitions (tests/golden/pyref.py), with
   no use of the oracle or the HIP library:
     * ``fr_constants.json``  -- Montgomery constants, roots, two-adicity, QNR;
     * ``fr_field_vectors.json`` -- BN254 Fr add/sub/mul/neg/inv on edge and
       LCG operands;
     * ``dft_vectors.json`` -- O(n^2) DFT and Horner-coset vectors;
     * ``textbook_digests.json`` -- SHA-256 of the recursive textbook NTT of
       the seeded workloads (BN254 Fr 2^16 seed 0, 2^20 seed 1), plus head/tail
       elements.

``gen_oracle_digests.py`` writes the digests of the larger workloads from the
oracle once it is pinned by the fixtures above.
"""
from __future__ import annotations

import json
import os
import re
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import pyref  # noqa: E402

REF = "/root/reference"


def dump(name: str, obj) -> None:
    with open(os.path.join(HERE, name), "w") as f:
        json.dump(obj, f, indent=1, sort_keys=True)
        f.write("\n")
    print("wrote", name, flush=True)


def hexs(xs):
    return [hex(x) for x in xs]


def extract_ref_kat() -> None:
    path = os.path.join(REF, "libs/ntt/src/root.zig")
    text = open(path).read().splitlines()
    cases = []
    field = None
    for line in text[437:449]:  # root.zig:438-449
        if "// Goldilocks" in line:
            field = "goldilocks"
            continue
        if "// BabyBear" in line:
            field = "babybear"
            continue
        m = re.search(r"\.log_n = (\d+), \.root = (0x[0-9a-f]+), \.input = &\.\{([^}]*)\}, "
                      r"\.expect = &\.\{([^}]*)\}", line)
        if not m:
            continue
        cases.append({
            "field": field,
            "log_n": int(m.group(1)),
            "root": m.group(2),
            "input": [s.strip() for s in m.group(3).split(",") if s.strip()],
            "expect": [s.strip() for s in m.group(4).split(",") if s.strip()],
        })
    assert len(cases) == 8, len(cases)
    dump("ref_ntt_kat.json", {"source": "libs/ntt/src/root.zig:438-449", "cases": cases})


def extract_ref_field_vectors() -> None:
    path = os.path.join(REF, "libs/field/src/lib.zig")
    text = open(path).read().splitlines()
    rows = {"babybear": text[257:266], "goldilocks": text[267:276]}  # lib.zig:258-266, 268-276
    out = {}
    for name, lines in rows.items():
        vecs = []
        for line in lines:
            m = re.search(r"\.a = (0x[0-9a-f]+), \.b = (0x[0-9a-f]+), \.sum = (0x[0-9a-f]+), "
                          r"\.prod = (0x[0-9a-f]+), \.inv_a = (0x[0-9a-f]+), \.neg_b = (0x[0-9a-f]+)", line)
            assert m, line
            vecs.append(dict(zip(["a", "b", "sum", "prod", "inv_a", "neg_b"], m.groups())))
        assert len(vecs) == 9
        out[name] = vecs
    dump("ref_field_vectors.json", {"source": "libs/field/src/lib.zig:257-276", "vectors": out})


def extract_ref_blake3() -> None:
    path = os.path.join(REF, "libs/hash/src/root.zig")
    text = open(path).read().splitlines()
    hw = []
    for line in text[121:127]:  # root.zig:122-127
        hw += re.findall(r"0x([0-9a-f]{2})", line)
    assert len(hw) == 32, len(hw)
    body = "\n".join(text[172:247])  # root.zig:173-247
    pat = re.search(r"const pattern = \[_\]u8\{([^}]*)\}", body)
    pattern = [int(x) for x in pat.group(1).split(",")]
    vecs = [{"len": int(a), "digest": b}
            for a, b in re.findall(r'\.len = (\d+), \.digest = "([0-9a-f]{64})"', body)]
    assert len(pattern) == 16 and len(vecs) == 45, (len(pattern), len(vecs))
    dump("ref_blake3_kat.json", {"source": "libs/hash/src/root.zig:121-127, 173-247",
                                 "hello_world": "".join(hw), "pattern": pattern, "vectors": vecs})


def extract_ref_fri() -> None:
    path = os.path.join(REF, "libs/fri/src/root.zig")
    text = open(path).read().splitlines()
    seq = re.search(r'"([0-9a-f]{64})"', text[1211]).group(1)  # root.zig:1212 (Goldilocks)
    assert "Goldilocks" in text[1209] and "wire-compat" in text[1215], (text[1209], text[1215])
    line = text[1271]  # root.zig:1272
    assert "the Goldilocks base-field domain" in line and ".log_n = 8" in line
    dom = [int(h, 16) for h in re.findall(r'"([0-9a-f]{64})"', line)]
    assert len(dom) == 256, len(dom)
    dump("ref_fri_kat.json", {"source": "libs/fri/src/root.zig:1210-1214, 1272",
                              "transcript": {"label": "wire-compat", "field": "goldilocks", "draws": 4,
                                             "hex": seq},
                              "domain": {"field": "goldilocks", "log_n": 8, "values": [hex(v) for v in dom]}})


def lcg(seed: int):
    x = seed
    while True:
        x = (x * 6364136223846793005 + 1442695040888963407) % (1 << 64)
        yield x


def fr_constants() -> None:
    p = pyref.FR
    s, q = pyref.two_adicity(p)
    c = pyref.mont_constants(p)
    roots = {str(l): hex(pyref.primitive_root(p, l)) for l in range(0, s + 1)}
    gl_roots = {str(l): hex(pyref.primitive_root(pyref.GL, l)) for l in range(0, 33)}
    dump("fr_constants.json", {
        "modulus": hex(p),
        "two_adicity": s,
        "odd_part": hex(q),
        "smallest_qnr": pyref.smallest_qnr(p),
        "coset_shift": 5,
        "coset_shift_legendre": pyref.legendre(5, p),
        "r_mod_p": hex(c["r_mod_p"]),
        "r2_mod_p": hex(c["r2_mod_p"]),
        "m0_inv64": hex(c["m0_inv64"]),
        "m0_inv32": hex(c["m0_inv32"]),
        "primitive_roots": roots,
        "goldilocks_primitive_roots": gl_roots,
        "goldilocks_smallest_qnr": pyref.smallest_qnr(pyref.GL),
        "goldilocks_coset_shift": 7,
    })


def fr_field_vectors() -> None:
    p = pyref.FR
    edge = [0, 1, 2, p - 1, p - 2, (p - 1) // 2, (p + 1) // 2, 1 << 64, (1 << 128) - 1, (1 << 253)]
    gen = lcg(0x5EED)
    rnd = []
    for _ in range(40):
        v = 0
        for _ in range(4):
            v = (v << 64) | next(gen)
        rnd.append(v % p)
    vecs = []
    ops = [(a, b) for a in edge for b in edge[:5]] + list(zip(rnd[0::2], rnd[1::2]))
    for a, b in ops:
        vecs.append({
            "a": hex(a), "b": hex(b),
            "sum": hex((a + b) % p), "diff": hex((a - b) % p), "prod": hex(a * b % p),
            "neg_b": hex(-b % p), "inv_a": hex(pow(a, -1, p) if a else 0),
            "a_mont": hex(pyref.to_mont(a)),
        })
    dump("fr_field_vectors.json", {"modulus": hex(p), "vectors": vecs})


def dft_vectors() -> None:
    out = []
    gen = lcg(0xD1F7)
    for fname, p, logs in (("fr", pyref.FR, range(0, 8)), ("goldilocks", pyref.GL, range(0, 9)),
                           ("babybear", pyref.BB, range(0, 7))):
        for log_n in logs:
            n = 1 << log_n
            xs = [next(gen) % p for _ in range(n)]
            w = pyref.primitive_root(p, log_n)
            shift = 5 if p == pyref.FR else (7 if p == pyref.GL else 31)
            out.append({
                "field": fname, "log_n": log_n, "root": hex(w), "input": hexs(xs),
                "dft": hexs(pyref.dft(xs, w, p)),
                "coset_shift": hex(shift),
                "coset": hexs(pyref.horner_coset(xs, n, w, shift, p)),
            })
    # LDE cases: 2^k coefficients, blowup 2^b, evaluated by Horner on shift*w^i.
    lde = []
    for fname, p, shift in (("fr", pyref.FR, 5), ("goldilocks", pyref.GL, 7)):
        for log_k, log_b in ((0, 1), (1, 2), (2, 2), (3, 3), (4, 2)):
            k, n = 1 << log_k, 1 << (log_k + log_b)
            xs = [next(gen) % p for _ in range(k)]
            w = pyref.primitive_root(p, log_k + log_b)
            lde.append({"field": fname, "log_k": log_k, "log_blowup": log_b, "coset_shift": hex(shift),
                        "coeffs": hexs(xs), "evals": hexs(pyref.horner_coset(xs, n, w, shift, p))})
    dump("dft_vectors.json", {"dft": out, "lde": lde})


def textbook_digests() -> None:
    p = pyref.FR
    res = {}
    for name, seed, log_n in (("fr_2p16_seed0", 0, 16), ("fr_2p20_seed1", 1, 20)):
        t0 = time.time()
        xs = pyref.uniform_elements(p, seed, 1 << log_n)
        w = pyref.primitive_root(p, log_n)
        fwd = pyref.textbook_ntt(xs, w, p)
        res[name] = {
            "seed": seed, "log_n": log_n,
            "input_digest": pyref.digest(p, xs),
            "input_head": hexs(xs[:4]),
            "forward_digest": pyref.digest(p, fwd),
            "forward_head": hexs(fwd[:4]), "forward_tail": hexs(fwd[-4:]),
        }
        print(f"{name}: {time.time() - t0:.1f}s", flush=True)
    dump("textbook_digests.json", {"algorithm": "recursive radix-2 textbook NTT in Python ints",
                                   "storage": "Montgomery limbs, little-endian (x*2^256 mod r)",
                                   "cases": res})


def main() -> None:
    steps = sys.argv[1:] or ["kat", "field", "blake3", "fri", "const", "frvec", "dft", "textbook"]
    if "kat" in steps:
        extract_ref_kat()
    if "blake3" in steps:
        extract_ref_blake3()
    if "fri" in steps:
        extract_ref_fri()
    if "field" in steps:
        extract_ref_field_vectors()
    if "const" in steps:
        fr_constants()
    if "frvec" in steps:
        fr_field_vectors()
    if "dft" in steps:
        dft_vectors()
    if "textbook" in steps:
        textbook_digests()


if __name__ == "__main__":
    main()
