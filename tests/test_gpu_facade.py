"""The C++ facade (include/acts_kf.hpp) on the GPU: per-track fit() calls
give the batched results bit for bit (tests/cpp/facade_check.cpp)."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "facade_check")


def test_facade_single_and_batch_agree():
    if not os.path.exists(BIN):
        pytest.fail(f"{BIN} not built (python -m acts_amd.build)")
    print("facade_check 300 tracks", flush=True)
    res = subprocess.run([BIN, "300"], capture_output=True, text=True, timeout=100)
    print(res.stdout, res.stderr, flush=True)
    assert res.returncode == 0, res.stdout + res.stderr
