"""Host-compiled check of the kernels' index arithmetic (tests/cpu/geom_check.cc).

It includes szp_amd/csrc/szp_geom.h, the same header the HIP kernels use for
item/tile geometry, and walks every tile, round and thread slot of many
(n, B, T) shapes: every element a compress thread loads must lie inside the
field and inside its own item, the fast simple-tile offsets must equal the
general path's, and the octets of a round must cover its elements once.
"""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_geom_check(tmp_path):
    exe = str(tmp_path / "geom_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-Wall", "-I",
                    os.path.join(ROOT, "szp_amd", "csrc"),
                    os.path.join(ROOT, "tests", "cpu", "geom_check.cc"), "-o", exe], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failures" in r.stdout
