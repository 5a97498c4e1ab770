"""Shared pytest set-up: markers and import paths.

``gpu`` tests need an MI355X (gfx950) and the built HIP library; everything
else runs on the CPU (the oracle, the golden fixtures, host logic, and that the
C-ABI library loads and exports every declared symbol).
"""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "zig-algebra_amd"), os.path.join(ROOT, "oracle"),
          os.path.join(ROOT, "tests", "golden")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU and the built HIP library")
