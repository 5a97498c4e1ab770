"""CPU check of the device Goldilocks arithmetic (csrc/field_gl.hpp).

tests/hostsim/gl_reduce_main.cpp includes the same header the kernels use,
built by g++ as host functions, and compares gl_reduce128, gl_mul,
gl_butterfly, gl_butterfly_one, gl_butterfly_t and the radix-8 rounds' shift
products gl_shl24/48/72 with unsigned __int128 arithmetic mod p:
every pair of edge values (0, 1, eps, p - 1, p, 2^64 - 1, ...), random
128-bit values and products, and 128-bit values built to take each of the
reduction's three cases (carried; not carried with r >= eps; not carried with
r < eps, which random products reach about once in 2^32).  The run fails if a
case is never reached.
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SIM = os.path.join(HERE, "hostsim")
OUT = os.path.join(SIM, "_build")


def test_gl_reduction_matches_int128_on_every_case():
    if shutil.which("g++") is None or not os.path.exists("/opt/rocm/include/hip/hip_runtime_api.h"):
        pytest.skip("needs g++ and the HIP headers")
    r = subprocess.run(["make", "-C", SIM, f"OUT={OUT}", os.path.join(OUT, "gl_reduce")], capture_output=True,
                       text=True)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    r = subprocess.run([os.path.join(OUT, "gl_reduce"), "2000000"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-2000:]
    assert "ok gl_reduce" in r.stdout
