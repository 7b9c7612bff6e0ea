"""Native host code on the CPU under sanitizers. The job server's host core
(menos_amd/csrc/server_core.cpp) with
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


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ is not installed")
def test_generic_kernel_replay_in_range_and_exact(tmp_path):
    """The per-group gate kernel's launch replayed lane by lane on the CPU
    (tests/native/gate_index_check.cpp): every access in range, every
    amplitude read and written once, results bitwise the oracle's generic
    kernel for k = 1..6 on 10..16 qubits."""
    exe = str(tmp_path / "gate_index_check")
    cmd = ["g++", "-std=c++20", "-O2", "-ffp-contract=off", "-fsanitize=address,undefined", "-fno-omit-frame-pointer",
           "-I" + os.path.join(ROOT, "menos_amd", "csrc"), "-I" + os.path.join(ROOT, "oracle"),
           os.path.join(ROOT, "tests", "native", "gate_index_check.cpp"), os.path.join(ROOT, "oracle", "oracle.cpp"),
           "-pthread", "-o", exe]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "bit-exact" in r.stdout
