"""CPU: the proof server's scheduling (admission by HBM budget, per-shape
pools, eviction, requests submitted while workers run, overlap only on
backlog, the bypass limit, the cross-proof MSM batcher) built against CPU
stand-ins for the HIP runtime and the prover core, under AddressSanitizer and
ThreadSanitizer.  Regression: requests used to live in a std::vector that
zks_submit grew while workers held references into it, so a worker could mark
a stale copy done and the waiter never woke."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = [os.path.join(ROOT, "tests", "server_mock", "mock.cpp"), os.path.join(ROOT, "tests", "server_mock", "main.cpp"),
       os.path.join(ROOT, "zkwasm-prover_amd", "csrc", "server.cpp")]


@pytest.mark.parametrize("san", ["address", "thread"])
def test_server_scheduling_under_sanitizer(tmp_path, san):
    if not shutil.which("g++") or not os.path.isdir("/opt/rocm/include"):
        pytest.skip("needs g++ and the ROCm headers")
    exe = str(tmp_path / ("srv_" + san))
    cmd = ["g++", "-std=c++17", "-g", "-O1", "-fsanitize=" + san, "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
           "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(ROOT, "zkwasm-prover_amd", "csrc")] + SRC + \
        ["-o", exe, "-lpthread"]
    subprocess.run(cmd, check=True, capture_output=True)
    for _ in range(3):
        r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
        assert "Sanitizer" not in r.stderr, r.stderr[-4000:]
