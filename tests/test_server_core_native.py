"""The job server's host core (menos_amd/csrc/server_core.cpp) on the CPU, with
a stub device backend, under AddressSanitizer + UndefinedBehaviorSanitizer and
under ThreadSanitizer: queue, lanes, the arena ledger, early completion of a
round's small jobs, concurrent waiters, replay, failures and shutdown.
Built with g++ (host code only)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = [os.path.join(ROOT, "tests", "native", "server_core_stress.cpp"),
       os.path.join(ROOT, "menos_amd", "csrc", "server_core.cpp")]
INC = ["-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(ROOT, "menos_amd", "csrc")]

SANITIZERS = {
    "address": ["-fsanitize=address,undefined", "-fno-omit-frame-pointer", "-fno-sanitize-recover=undefined"],
    "thread": ["-fsanitize=thread"],
}


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ is not installed")
@pytest.mark.parametrize("sanitizer", sorted(SANITIZERS))
def test_server_core_under_sanitizer(tmp_path, sanitizer):
    exe = str(tmp_path / f"server_core_{sanitizer}")
    cmd = ["g++", "-std=c++20", "-O1", "-g", *SANITIZERS[sanitizer], *INC, *SRC, "-pthread", "-o", exe]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", TSAN_OPTIONS="halt_on_error=1")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "all scenarios passed" in r.stdout
    assert "WARNING: ThreadSanitizer" not in r.stderr and "ERROR: AddressSanitizer" not in r.stderr, r.stderr
