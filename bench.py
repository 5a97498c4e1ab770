#!/usr/bin/env python3
"""Benchmark: BN254 Fr forward NTT of 2^24 elements on one MI355X.

A step is one in-place forward NTT (libs/ntt/src/root.zig:65-90 semantics,
root = primitive_root(24)) over 2^24 BN254 Fr elements already resident in
HBM, like the reference's own benchmark which transforms the same buffer in
place on every iteration (libs/pairing/src/bench.zig:16-24, 107-111).

Prints one JSON line:
  value        whole-job throughput, Melem/s (elements transformed / time)
  roofline     the dominant kernel (the LDS stage-range pass) against the
               VALU 32-bit multiply peak: 136 32-bit multiplies per 254-bit
               Montgomery product (2 s^2 + s for s = 8 limbs), one product
               per butterfly, (n/2) * log2 n butterflies per transform
  cpu_baseline the CPU oracle (a C restatement of the reference path) on the
               same workload, one thread, timed on this host
"""
import argparse
import hashlib
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "zig-algebra_amd"))

MI355X_CUS = 256
MI355X_CLOCK_HZ = 2.4e9
LANES_PER_CU_CLK = 64          # v_mad_u64_u32 issues at full VALU rate on gfx950
VALU_MUL_PEAK = MI355X_CUS * MI355X_CLOCK_HZ * LANES_PER_CU_CLK  # 3.93e13 mul/s
MULS_PER_MODMUL = 136
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "fr_ntt_2p24_traffic.json")


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--log-n", type=int, default=24)
    ap.add_argument("--cpu-log-n", type=int, default=24, help="size of the CPU-baseline sample")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-host", action="store_true", help="skip the PCIe-inclusive host-buffer timing")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(transforms shard with no communication, see DESIGN.md)", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    import zntt_amd as Z

    log_n = args.log_n
    n = 1 << log_n
    field = Z.FIELD_FR
    ctx = Z.Context(0)
    t0 = time.time()
    x = Z.fill_uniform(field, 1, n)
    root = Z.primitive_root(field, log_n)
    gen_s = time.time() - t0
    dev = torch.from_numpy(x.view(np.int64)).cuda()
    stream = torch.cuda.current_stream()
    sp = stream.cuda_stream

    def step():
        ctx.execute_device(field, Z.FORWARD, log_n, dev.data_ptr(), root, stream=sp)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record(stream)
    for _ in range(args.steps):
        step()
    end.record(stream)
    torch.cuda.synchronize()
    ms_total = start.elapsed_time(end)
    ms_step = ms_total / args.steps
    value = n / (ms_step * 1e-3) / 1e6

    # Correctness of the benchmarked path: one fresh forward transform of the
    # seeded input, compared with the committed digest of the oracle's output.
    check = torch.from_numpy(x.view(np.int64)).cuda()
    ctx.execute_device(field, Z.FORWARD, log_n, check.data_ptr(), root, stream=sp)
    torch.cuda.synchronize()
    digest = hashlib.sha256(check.cpu().numpy().tobytes()).hexdigest()
    gold = json.load(open(os.path.join(ROOT, "tests", "golden", "oracle_digests.json")))["fr"].get(str(log_n))
    bit_exact = None if gold is None else (digest == gold["forward"])
    del check

    # Dominant kernel: the stage-range pass; log_n stages over `passes` launches
    # (the library's BN254 Fr plan: one pass up to 2^11, else ceil(log_n / 9)).
    passes = 1 if log_n <= 11 else -(-log_n // 9)
    muls_per_transform = (n // 2) * log_n * MULS_PER_MODMUL
    launch_ms = ms_step / passes
    achieved = muls_per_transform / passes / (launch_ms * 1e-3)
    traffic = None
    if os.path.exists(TRAFFIC_FILE):
        try:
            traffic = json.load(open(TRAFFIC_FILE)).get("hbm_bytes_per_launch")
        except (ValueError, OSError):
            traffic = None
    roofline = {"bound": "valu", "achieved": achieved / 1e12, "peak": VALU_MUL_PEAK / 1e12,
                "unit": "Tmul32/s", "frac": achieved / VALU_MUL_PEAK, "traffic": traffic}

    extras = {"passes_per_transform": passes, "avg_pass_ms": launch_ms, "bit_exact_vs_oracle_digest": bit_exact,
              "input_gen_s": round(gen_s, 2)}
    mad, copy = ctx.probe_peaks()
    extras["measured_mad_peak_Tmul_s"] = mad / 1e12
    extras["measured_copy_GB_s"] = copy / 1e9
    extras["frac_of_measured_mad_peak"] = achieved / mad

    if not args.skip_host:
        # PCIe-inclusive rate through the synchronous host-buffer entry point.
        h = x.copy()
        ctx.ntt(field, h, log_n, root)
        t1 = time.time()
        ctx.ntt(field, h, log_n, root)
        extras["host_buffer_melem_s"] = n / (time.time() - t1) / 1e6

    cpu = None
    if not args.skip_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        cl = args.cpu_log_n
        cx = O.gen_uniform(O.FIELD_FR, 1, 1 << cl)
        croot = O.primitive_root(O.FIELD_FR, cl)[1]
        t1 = time.time()
        assert O.ntt(O.FIELD_FR, cx, cl, croot) == O.OK
        cs = time.time() - t1
        cpu = {"value": (1 << cl) / cs / 1e6, "unit": "Melem/s", "cores": 1, "kind": "port",
               "sample": f"one forward NTT of 2^{cl} BN254 Fr (seed 1), C oracle -O3, 1 thread, {cs:.1f} s"}

    line = {
        "metric": "Melem/s BN254 Fr NTT 2^24",
        "value": value,
        "unit": "Melem/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic: uniform BN254 Fr in [0, r), seed 1 (SplitMix64/xoshiro256++)",
        "config": {"workload": f"BN254 Fr forward NTT, 2^{log_n} elements, in place, resident in HBM",
                   "field": "bn254_fr", "log_n": log_n, "elements": n},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "extras": extras,
    }
    print(json.dumps(line), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
