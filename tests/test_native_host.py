"""Host-side bookkeeping of the native library, compiled with g++ and run on
the CPU: the buddy division of arena chunks."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "doki_amd", "csrc")


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_buddy_pieces(tmp_path, seed):
    exe = tmp_path / "buddy_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-Wall", f"-I{CSRC}",
                    os.path.join(HERE, "native", "buddy_check.cpp"), "-o", str(exe)],
                   check=True)
    r = subprocess.run([str(exe), str(seed)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip() == "ok"
