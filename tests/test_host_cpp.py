"""C++ host mirror (small-field-metal-ntt_amd/host/sfntt.hpp) through its own
test program, tests/cpp/host_api_test.cpp (built by tests/cpp/build_host_test.py)."""
import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "cpp"))
import build_host_test  # noqa: E402


def _run(mode: str, timeout: int):
    exe = build_host_test.build(force=False)
    r = subprocess.run([exe, mode], capture_output=True, text=True, timeout=timeout)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "FAIL" not in r.stdout
    return r.stdout


def test_host_mirror_errors_and_sizes():
    out = _run("--no-device", 60)
    assert out.count("ok") >= 8


@pytest.mark.gpu
def test_host_mirror_parity_against_oracle():
    out = _run("--parity", 300)
    assert "CosetLdeBatch.execute" in out and "BbBatchNtt forward row-major" in out
