"""CPU regression tests of the C ABI's object lifetimes, under AddressSanitizer.

tests/hostsim builds the library's host code (context, twiddle cache,
batches, serving loop) with g++ against a stand-in HIP runtime whose kernel
launchers read and write the first and last byte of every buffer the real
kernel would touch.  Each scenario destroys or trims a context while a batch
or server made on it is alive and then keeps using that child:

* a batch replayed and destroyed after zntt_ctx_destroy;
* zntt_ctx_trim while a batch pins twiddle tables, then replays;
* a server submitting, reading and releasing after zntt_ctx_destroy;
* zntt_ctx_trim while served requests are in flight;
* a batch of a destroyed context destroyed after a new context has filled
  its cache (the new context's byte count must not change).

Before children held a reference on their context, every one of these was a
heap-use-after-free (reads of freed twiddle tables by replayed launches,
writes into freed cache entries and into the freed context).

Two more scenarios use the stand-in runtime's call counters:
* zntt_execute_device_noalloc makes no allocation or free once its tables
  are cached, and refuses oversized or unprepared transforms up front;
* device transforms on two streams of one context wait for each other's use
  of the shared scratch exactly when the stream changes;
* after zntt_ctx_reserve the seven zig-ntt functions allocate nothing at the
  reserved sizes (every field, new roots, registered and foreign tables,
  across a trim), so device memory can run out only at the reservation.

And one the stand-in runtime's memory makes possible: host buffers of 32 and
128 MiB and a 64 MiB LDE input go through the context's page-locked ring in
chunks and come back intact.
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SIM = os.path.join(HERE, "hostsim")
OUT = os.path.join(SIM, "_build")
SCENARIOS = ["batch_outlives_ctx", "trim_with_live_batch", "server_outlives_ctx", "trim_with_live_server",
             "late_destroy_after_new_ctx", "noalloc_allocates_nothing", "streams_share_scratch_in_order",
             "reserved_calls_allocate_nothing", "host_copies_through_the_ring"]


@pytest.fixture(scope="module")
def harness():
    if shutil.which("g++") is None or not os.path.exists("/opt/rocm/include/hip/hip_runtime_api.h"):
        pytest.skip("needs g++ and the HIP headers")
    r = subprocess.run(["make", "-C", SIM, "-j4", f"OUT={OUT}"], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return os.path.join(OUT, "lifetime")


@pytest.mark.parametrize("scenario", SCENARIOS)
def test_lifetime_scenario_is_clean_under_asan(harness, scenario):
    env = dict(os.environ, ASAN_OPTIONS="abort_on_error=0:halt_on_error=1")
    r = subprocess.run([harness, scenario], capture_output=True, text=True, timeout=120, env=env)
    assert "AddressSanitizer" not in r.stderr, r.stderr[-4000:]
    assert r.returncode == 0, r.stdout + r.stderr[-4000:]
    assert f"ok {scenario}" in r.stdout
