"""The solve server's host logic on the CPU: csu_serve.cpp compiled with a
host stand-in of the device library (tests/serve_host/stub_dev.cpp, a
closed-form residual per solve) and the address and undefined-behaviour
sanitizers, driven by tests/serve_host/drive.cpp: one context per solve,
several slots, batched units with solve, cell and HBM limits, CFL ramps,
iteration caps, a diverging solve and the mixed-precision switch give the
closed-form results; batches hold only solves that share kernels; every
context is freed."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="no g++")
def test_server_logic_under_sanitizers(tmp_path):
    exe = str(tmp_path / "drive")
    src = [os.path.join(ROOT, "csu_amd/csrc/csu_serve.cpp"), os.path.join(ROOT, "tests/serve_host/stub_dev.cpp"),
           os.path.join(ROOT, "tests/serve_host/drive.cpp")]
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
           "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(ROOT, "tests/serve_host")] + src + \
          ["-lpthread", "-o", exe]
    subprocess.run(cmd, check=True, timeout=300)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and r.stdout.strip() == "ok"
