"""The serving runtime's scheduling and memory lifetimes on the CPU: server.cpp
linked against a host mock of the instance/batch C-ABI (tests/cpp/server_mock.cpp)
and run under AddressSanitizer and ThreadSanitizer with the serving benchmark's
request mix.  A member used after it was destroyed, a destroyed handle left in
a batch, a data race or a request that never finishes fails the test."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = [os.path.join(HERE, "cpp", "server_mock.cpp"), os.path.join(ROOT, "zisafvm_amd", "csrc", "server.cpp"),
       os.path.join(ROOT, "zisafvm_amd", "csrc", "setup.cpp")]


@pytest.mark.parametrize("members", [64, 3])
@pytest.mark.parametrize("san", ["address,undefined", "thread"])
def test_server_runtime_under_sanitizers(san, members):
    # members 3: fewer batch slots than requests in flight, so requests that
    # find their batch full wait for a slot (none may fail)
    out = os.path.join(HERE, "cpp", "build", "server_mock_" + san.split(",")[0])
    os.makedirs(os.path.dirname(out), exist_ok=True)
    subprocess.run(["g++", "-std=c++17", "-g", "-O1", "-fopenmp", f"-fsanitize={san}", "-fno-omit-frame-pointer",
                    *SRC, "-o", out], check=True)
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1", ASAN_OPTIONS="detect_leaks=1")
    r = subprocess.run([out, "48", "400", str(members)], capture_output=True, text=True, timeout=600, env=env)
    print(r.stdout[-2000:], r.stderr[-4000:])
    assert r.returncode == 0 and "server_mock: OK" in r.stdout
