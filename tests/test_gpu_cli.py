"""The drop-in command-line drivers (szp_amd/bin, built from szp_amd/cli):
the reference's example/testfloat_{compress,decompress}_fastmode1.cc
argument order, the .szp file holding the stream the oracle writes, and the
.szp.out file the oracle's decompression (float and double, both fastModes)."""
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle_lib
from golden_inputs import make_input, make_input_f64

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "szp_amd", "bin")


@pytest.mark.parametrize("dtype,access", [("float", "nonrandom"), ("float", "random"),
                                          ("double", "nonrandom"), ("double", "random")])
def test_cli_roundtrip(tmp_path, dtype, access):
    n, B, T, eb = 300001, 64, 8, 1e-3
    x = make_input_f64("waves", n, 21) if dtype == "double" else make_input("waves", n, 21)
    src = str(tmp_path / ("field." + dtype))
    x.tofile(src)
    env = dict(os.environ, SZP_AMD_NUM_THREADS=str(T))
    print("cli case", dtype, access, flush=True)
    r = subprocess.run([os.path.join(BIN, "testfloat_compress_fastmode1"), src, str(B), str(eb),
                        dtype, access, "2", "1", str(tmp_path / "c.csv")],
                       capture_output=True, text=True, env=env, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "CR=" in r.stdout and "Performance Statistics" in r.stdout
    fast = 2 if access == "random" else 1
    want = oracle_lib.compress_mode(fast, x, 0, eb, 0.0, B, T)
    got = np.fromfile(src + ".szp", np.uint8)
    assert got.size == want.size and np.array_equal(got, want)
    assert open(str(tmp_path / "c.csv")).read().startswith("Run,TimeSeconds,CompressionSize")
    r = subprocess.run([os.path.join(BIN, "testfloat_decompress_fastmode1"), src + ".szp", str(n),
                        str(B), dtype, access, "1"],
                       capture_output=True, text=True, env=env, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Max absolute error" in r.stdout and "PSNR" in r.stdout
    ref = oracle_lib.decompress_mode(fast, want, n, B, T, dtype=x.dtype)
    dec = np.fromfile(src + ".szp.out", x.dtype)
    ut = np.uint64 if dtype == "double" else np.uint32
    assert np.array_equal(dec.view(ut), ref.view(ut))
