"""Shared test setup: the `gpu` marker and import paths.

GPU tests (marked `gpu`) call libosc through its C ABI on a real MI355X and
compare with the CPU oracle under oracle/.  Everything else runs on the CPU.
"""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("osss-ucx_amd", "oracle"):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line(
        "markers", "gpu: needs a real MI355X (runs the HIP kernels)")
