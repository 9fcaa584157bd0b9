"""The grid barrier protocol of the batched persistent solver
(engine_amd/csrc/gridbar.h), run on host threads: barrier semantics,
counter resets across rounds, and the bounded wait."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_gridbar_protocol_on_threads(tmp_path):
    exe = str(tmp_path / "gridbar_protocol")
    subprocess.run(["g++", "-std=c++17", "-O2", "-pthread", "-I", os.path.join(ROOT, "engine_amd", "csrc"),
                    os.path.join(ROOT, "tests", "native", "gridbar_protocol.cpp"), "-o", exe], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
