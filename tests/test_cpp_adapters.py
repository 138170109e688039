"""The C++ adapters compile against include/zfv.h and libzfv_hip.so (CPU) and
drive the GPU through the reference's RateOfChange / TimeIntegration
contracts (GPU)."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "cpp", "adapter_check.cpp")
OUT = os.path.join(HERE, "cpp", "build", "adapter_check")


def build():
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    lib_dir = os.path.join(ROOT, "zisafvm_amd", "_lib")
    cmd = ["g++", "-O2", "-std=c++17", SRC, "-o", OUT, f"-L{lib_dir}", "-lzfv_hip", f"-Wl,-rpath,{lib_dir}"]
    subprocess.run(cmd, check=True)
    return OUT


def test_adapters_compile_and_link():
    assert os.path.exists(build())


@pytest.mark.gpu
def test_adapters_drive_the_gpu(gpu):
    exe = OUT if os.path.exists(OUT) else build()
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and "adapter_check: OK" in r.stdout
