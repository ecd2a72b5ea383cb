"""The request server (sa_server.hpp) against a stub engine on the CPU under
ThreadSanitizer and AddressSanitizer + UndefinedBehaviorSanitizer: client
threads submit and wait concurrently while four workers create, run and
evict sessions under a budget that holds about two of them.  The stub
checks that no session is used after it is destroyed, that the bytes of live
sessions never pass the budget, and that every result is the function of
its request (tests/native/server_stub.cpp)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "native", "server_stub.cpp")


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ is needed")
@pytest.mark.parametrize("san", ["thread", "address,undefined"])
def test_server_under_sanitizer(san, tmp_path):
    exe = tmp_path / "stub"
    cmd = ["g++", "-std=c++17", "-O1", "-g", f"-fsanitize={san}", "-fno-omit-frame-pointer", "-I",
           os.path.join(ROOT, "include"), SRC, "-o", str(exe), "-pthread"]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1", ASAN_OPTIONS="detect_leaks=1:halt_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    res = subprocess.run([str(exe), "300"], capture_output=True, text=True, env=env, timeout=240)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "WARNING: ThreadSanitizer" not in res.stderr and "ERROR: AddressSanitizer" not in res.stderr
    assert "failures 0" in res.stdout
