"""Builds libszp_amd.so (HIP kernels + C ABI) in-tree for gfx950 with hipcc,
and the drop-in command-line drivers (szp_amd/bin) linked against it.

    python szp_amd/build.py      # or __graft_entry__.build(); runs without importing the package
"""
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libszp_amd.so")
SOURCES = ["szp_kernels.hip", "szp_api.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def hipcc_command(out=LIB):
    return [
        HIPCC, "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
        "-Wall", "-Wno-unused-result",
        # Packed FP32 instructions stay off in all device code.
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-I", os.path.join(ROOT, "include"), "-I", CSRC,
        *[os.path.join(CSRC, s) for s in SOURCES],
        "-Wl,-Bsymbolic", "-ldl", "-o", out,
    ]


CLI = ["testfloat_compress_fastmode1", "testfloat_decompress_fastmode1"]
BIN = os.path.join(PKG, "bin")


def cli_command(name):
    """g++ line of one drop-in command-line driver (szp_amd/cli), linked
    against the in-tree library (found through an $ORIGIN rpath)."""
    return ["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"),
            os.path.join(PKG, "cli", name + ".cc"), "-L", PKG, "-lszp_amd",
            "-Wl,-rpath,$ORIGIN/..", "-o", os.path.join(BIN, name)]


def build(verbose=False):
    tmp = LIB + ".tmp"
    cmd = hipcc_command(tmp)
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libszp_amd.so")
    os.replace(tmp, LIB)
    os.makedirs(BIN, exist_ok=True)
    for name in CLI:
        c = cli_command(name)
        if verbose:
            print(" ".join(c), flush=True)
        r = subprocess.run(c, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError("g++ failed building " + name)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
