"""GPU parity: libszp_amd.so (HIP kernels, called through its C ABI) against
the golden fixtures written by the reference and against the CPU oracle.

Bar: compressed streams byte-identical, decompressed values bit-identical.
"""
import hashlib
import sys

import numpy as np
import pytest

import oracle_lib
from golden_inputs import make_input

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def szp():
    import torch
    assert torch.cuda.is_available(), "no HIP device visible"
    import szp_amd
    return szp_amd


@pytest.fixture(scope="module")
def torch_mod():
    import torch
    return torch


def _legacy_roundtrip(szp, x, mode, ab, rel, B, T):
    szp.set_threads(T)
    try:
        stream, size = szp.szp_compress(1, 0, x, mode, ab, rel, x.size, B)
        assert size == stream.size
        dec = szp.szp_decompress(1, 0, stream, size, x.size, B)
    finally:
        szp.set_threads(0)
    return stream, dec


def test_golden_fixtures_legacy_api(szp, golden):
    cases, arrays, _ = golden
    for c in cases:
        k = c["id"]
        x = arrays["in_%d" % k].view(np.float32)
        say("golden case", k, c["kind"], "n=%d B=%d T=%d mode=%d" % (c["n"], c["B"], c["T"], c["mode"]))
        stream, dec = _legacy_roundtrip(szp, x, c["mode"], c["abs_eb"], c["rel"], c["B"], c["T"])
        ref = arrays["stream_%d" % k]
        assert stream.size == ref.size, (k, stream.size, ref.size)
        bad = np.nonzero(stream != ref)[0]
        assert bad.size == 0, (k, bad[:8])
        assert np.array_equal(dec.view(np.uint32), arrays["out_%d" % k]), k


CASES = [
    # kind, n, B, T, mode, abs_eb, rel
    ("walk", 1, 8, 1, 0, 0.01, 0.0),
    ("walk", 2, 8, 1, 0, 0.01, 0.0),
    ("walk", 3, 8, 8, 0, 0.01, 0.0),
    ("walk", 7, 32, 8, 0, 0.01, 0.0),
    ("walk", 16, 8, 16, 0, 0.01, 0.0),
    ("walk", 40, 8, 16, 0, 0.01, 0.0),
    ("walk", 100, 1, 3, 0, 0.01, 0.0),
    ("walk", 1000, 2, 4, 0, 0.01, 0.0),
    ("walk", 4099, 5, 7, 0, 0.01, 0.0),
    ("walk", 10000, 9, 2, 1, 0.0, 1e-3),
    ("walk", 65537, 127, 5, 1, 0.0, 1e-4),
    ("walk", 65536, 256, 8, 1, 0.0, 1e-4),
    # B = 2048: one item per round (general tiles only).  An earlier 32-bit
    # index of the element before an octet (x[Et + (uint32_t)(0 - 1)]) read
    # 16 GiB past the tile here; indices are 64-bit and tests/cpu/geom_check.cc
    # checks every load of these shapes on the CPU.
    ("walk", 70000, 2048, 3, 1, 0.0, 1e-4),
    ("waves", 1 << 20, 128, 16, 1, 0.0, 1e-4),
    ("waves", 1 << 20, 32, 16, 1, 0.0, 1e-2),
    ("waves", (1 << 20) + 13, 64, 12, 1, 0.0, 1e-3),
    ("uniform", 300000, 128, 8, 0, 0.002, 0.0),
    ("zeroblocks", 500000, 128, 8, 1, 0.0, 1e-3),
    ("zeroblocks", 500003, 32, 3, 1, 0.0, 1e-3),
    ("constant", 100000, 32, 4, 1, 0.0, 1e-3),
    ("constant", 100000, 128, 4, 0, 0.5, 0.0),
    ("specials", 200000, 32, 4, 0, 0.5, 0.0),
    ("nearint", 200000, 128, 4, 0, 0.001, 0.0),  # products within ulps of integers
    ("nearint", 65536, 32, 2, 0, 0.001, 0.0),
    ("sparsewide", 200000, 64, 2, 0, 1.0, 0.0),
    ("sparsewide", 200000, 16, 5, 0, 0.25, 0.0),
    ("tiny", 50000, 32, 2, 0, 5e-39, 0.0),
    ("signedzero", 50000, 32, 4, 1, 0.0, 1e-3),
    ("nanfirst", 20000, 32, 2, 1, 0.0, 1e-3),
    ("specials", 300001, 32, 4, 1, 0.0, 1e-3),   # REL over +-Inf: eb = Inf
    ("posinf", 10000, 32, 2, 1, 0.0, 1e-3),      # REL, all +Inf: eb = Inf - Inf = NaN
    ("signedzero", (1 << 22) + 5, 128, 16, 1, 0.0, 1e-3),  # many range partials of +-0
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: "%s-n%d-B%d-T%d" % (c[0], c[1], c[2], c[3]))
def test_matches_oracle(szp, case):
    kind, n, B, T, mode, ab, rel = case
    x = make_input(kind, n, 7 + n)
    say("oracle case", case)
    want = oracle_lib.compress(x, mode, ab, rel, B, T)
    stream, dec = _legacy_roundtrip(szp, x, mode, ab, rel, B, T)
    assert stream.size == want.size
    bad = np.nonzero(stream != want)[0]
    assert bad.size == 0, bad[:8]
    ref_dec = oracle_lib.decompress(want, n, B, T)
    assert np.array_equal(dec.view(np.uint32), ref_dec.view(np.uint32))


def test_device_api_and_hash_fixtures(szp, torch_mod, golden):
    torch = torch_mod
    _, _, hashes = golden
    for h in hashes:
        x = make_input(h["kind"], h["n"], h["seed"])
        assert hashlib.sha256(x.tobytes()).hexdigest() == h["input_sha256"]
        say("hash case", h["kind"], h["n"], h["B"], h["T"])
        codec = szp.DeviceCodec(block_size=h["B"], threads=h["T"])
        xd = torch.from_numpy(x).cuda()
        out, size_t = codec.compress(xd, h["mode"], h["abs_eb"], h["rel"])
        codec.check()
        size = int(size_t.item())
        assert size == h["stream_bytes"]
        stream = out[:size].cpu().numpy()
        assert hashlib.sha256(stream.tobytes()).hexdigest() == h["stream_sha256"]
        dec = codec.decompress(out, size, h["n"])
        codec.check()
        assert hashlib.sha256(dec.cpu().numpy().tobytes()).hexdigest() == h["out_sha256"]


def test_decompress_accepts_upper_bound_length(szp, torch_mod):
    """The device decoder takes any upper bound of the stream length."""
    torch = torch_mod
    x = make_input("waves", 300001, 5)
    B, T = 128, 6
    want = oracle_lib.compress(x, 1, 0.0, 1e-3, B, T)
    codec = szp.DeviceCodec(block_size=B, threads=T)
    buf = torch.zeros(want.size + 5000, dtype=torch.uint8)
    buf[:want.size] = torch.from_numpy(want)
    buf[want.size:] = torch.randint(0, 256, (5000,), dtype=torch.uint8)
    d = buf.cuda()
    dec = codec.decompress(d, d.numel(), x.size)
    codec.check()
    ref = oracle_lib.decompress(want, x.size, B, T)
    assert np.array_equal(dec.cpu().numpy().view(np.uint32), ref.view(np.uint32))


INDEXED_CASES = [
    ("waves", (1 << 20) + 13, 128, 16, 1, 0.0, 1e-4),
    ("waves", 1 << 20, 32, 8, 1, 0.0, 1e-2),
    ("zeroblocks", 500003, 64, 3, 1, 0.0, 1e-3),
    ("walk", 70000, 9, 5, 0, 0.01, 0.0),
    ("walk", 40, 8, 16, 0, 0.01, 0.0),
    ("specials", 200000, 32, 4, 0, 0.5, 0.0),
    ("uniform", 300000, 2048, 2, 0, 0.002, 0.0),
    # Pipelined decode: tiles over 16 KiB of stream (decoded after the
    # pipeline), B = 8 / 16 / 256, and more general tiles than the list holds
    # (T = 80: the look-back-free tile kernel instead).
    ("uniform", 1 << 20, 64, 8, 0, 1e-5, 0.0),
    ("walk", 500000, 8, 4, 0, 0.01, 0.0),
    ("waves", 3000017, 16, 31, 1, 0.0, 1e-3),
    ("waves", 1 << 20, 256, 8, 1, 0.0, 1e-4),
    ("waves", 1 << 20, 32, 80, 1, 0.0, 1e-3),
]


@pytest.mark.parametrize("case", INDEXED_CASES, ids=lambda c: "%s-n%d-B%d-T%d" % (c[0], c[1], c[2], c[3]))
def test_side_index_roundtrip(szp, torch_mod, case):
    """compress(index=True) leaves the stream byte-identical; decompress with
    the side index is bit-identical to the oracle's decompression."""
    torch = torch_mod
    kind, n, B, T, mode, ab, rel = case
    x = make_input(kind, n, 11 + n)
    say("indexed case", case)
    want = oracle_lib.compress(x, mode, ab, rel, B, T)
    codec = szp.DeviceCodec(block_size=B, threads=T)
    xd = torch.from_numpy(x).cuda()
    out, size_t, idx = codec.compress(xd, mode, ab, rel, index=True)
    codec.check()
    size = int(size_t.item())
    assert size == want.size
    assert np.array_equal(out[:size].cpu().numpy(), want)
    dec = codec.decompress(out, size, n, index=idx)
    codec.check()
    ref = oracle_lib.decompress(want, n, B, T)
    assert np.array_equal(dec.cpu().numpy().view(np.uint32), ref.view(np.uint32))


PIPE_CASES = [
    ("waves", (1 << 20) + 13, 128, 16, 1, 0.0, 1e-4),
    ("uniform", 1 << 20, 64, 8, 0, 1e-5, 0.0),
    ("walk", 500000, 8, 4, 0, 0.01, 0.0),
    ("waves", 1 << 20, 256, 8, 1, 0.0, 1e-4),
]


@pytest.mark.parametrize("case", PIPE_CASES, ids=lambda c: "%s-n%d-B%d-T%d" % (c[0], c[1], c[2], c[3]))
def test_pipelined_decode_many_steps(szp, torch_mod, case, monkeypatch):
    """The pipelined decode with its grid capped (SZP_AMD_PIPE_GRID), so that
    each workgroup runs many pipeline steps and defers several tiles: the
    result is bit-identical to the oracle's decompression."""
    torch = torch_mod
    kind, n, B, T, mode, ab, rel = case
    x = make_input(kind, n, 23 + n)
    say("pipelined case", case)
    want = oracle_lib.compress(x, mode, ab, rel, B, T)
    codec = szp.DeviceCodec(block_size=B, threads=T)
    xd = torch.from_numpy(x).cuda()
    out, size_t, idx = codec.compress(xd, mode, ab, rel, index=True)
    codec.check()
    size = int(size_t.item())
    assert np.array_equal(out[:size].cpu().numpy(), want)
    ref = oracle_lib.decompress(want, n, B, T)
    for cap in ("3", "13"):
        monkeypatch.setenv("SZP_AMD_PIPE_GRID", cap)
        dec = codec.decompress(out, size, n, index=idx)
        codec.check()
        assert np.array_equal(dec.cpu().numpy().view(np.uint32), ref.view(np.uint32)), cap
