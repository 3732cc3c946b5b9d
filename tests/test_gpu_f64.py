"""GPU parity of the float64 path (szp_compress / szp_decompress with
SZ_DOUBLE, SZp/src/szp_double.cc:148-318, SZp/src/szpd_double.cc:23-106)
through the C ABI of libszp_amd.so, against the reference's fixtures and the
CPU oracle.  Bar: streams byte-identical, decompressed values bit-identical.
"""
import ctypes
import json
import os
import sys

import numpy as np
import pytest

import oracle_lib
from golden_inputs import make_input_f64

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def szp():
    import torch
    assert torch.cuda.is_available(), "no HIP device visible"
    import szp_amd
    return szp_amd


def _roundtrip(szp, x, mode, ab, rel, B, T):
    szp.set_threads(T)
    try:
        stream, size = szp.szp_compress(1, szp.SZ_DOUBLE, x, mode, ab, rel, x.size, B)
        assert size == stream.size
        dec = szp.szp_decompress(1, szp.SZ_DOUBLE, stream, size, x.size, B)
    finally:
        szp.set_threads(0)
    return stream, dec


def test_f64_golden_fixtures(szp):
    d = os.path.join(HERE, "golden")
    cases = json.load(open(os.path.join(d, "cases_f64.json")))
    arrays = np.load(os.path.join(d, "cases_f64.npz"))
    for c in cases:
        k = c["id"]
        x = arrays["in_%d" % k].view(np.float64)
        say("f64 golden case", k, c["kind"], "n=%d B=%d T=%d" % (c["n"], c["B"], c["T"]))
        stream, dec = _roundtrip(szp, x, c["mode"], c["abs_eb"], c["rel"], c["B"], c["T"])
        ref = arrays["stream_%d" % k]
        assert stream.size == ref.size, (k, stream.size, ref.size)
        bad = np.nonzero(stream != ref)[0]
        assert bad.size == 0, (k, bad[:8])
        assert np.array_equal(dec.view(np.uint64), arrays["out_%d" % k]), k


F64_CASES = [
    # kind, n, B, T, mode, abs_eb, rel
    ("walk", 1, 8, 1, 0, 0.01, 0.0),
    ("walk", 3, 8, 8, 0, 0.01, 0.0),
    ("walk", 40, 8, 16, 0, 0.01, 0.0),
    ("walk", 4099, 5, 7, 0, 0.01, 0.0),
    ("walk", 70000, 9, 5, 1, 0.0, 1e-6),
    ("waves", (1 << 20) + 13, 128, 16, 1, 0.0, 1e-4),
    ("waves", 1 << 20, 32, 8, 1, 0.0, 1e-8),
    ("uniform", 300000, 64, 8, 0, 1e-7, 0.0),
    ("zeroblocks", 500003, 32, 3, 1, 0.0, 1e-3),
    ("specials", 200000, 32, 4, 0, 0.5, 0.0),
    ("sparsewide", 200000, 64, 2, 0, 1.0, 0.0),
    ("signedzero", 50000, 32, 4, 1, 0.0, 1e-3),
    ("nanfirst", 20000, 32, 2, 1, 0.0, 1e-3),
    ("constant", 100000, 128, 4, 1, 0.0, 1e-3),
]


@pytest.mark.parametrize("case", F64_CASES, ids=lambda c: "%s-n%d-B%d-T%d" % (c[0], c[1], c[2], c[3]))
def test_f64_matches_oracle(szp, case):
    """The drop-in API (parse on decompress) and the device API with the side
    index, both against the oracle."""
    import torch
    kind, n, B, T, mode, ab, rel = case
    x = make_input_f64(kind, n, 13 + n)
    say("f64 oracle case", case)
    want = oracle_lib.compress(x, mode, ab, rel, B, T, dtype=np.float64)
    ref_dec = oracle_lib.decompress(want, n, B, T, dtype=np.float64)
    stream, dec = _roundtrip(szp, x, mode, ab, rel, B, T)
    assert stream.size == want.size
    bad = np.nonzero(stream != want)[0]
    assert bad.size == 0, bad[:8]
    assert np.array_equal(dec.view(np.uint64), ref_dec.view(np.uint64))

    codec = szp.DeviceCodec(block_size=B, threads=T)
    xd = torch.from_numpy(x).cuda()
    out, size_t, idx = codec.compress(xd, mode, ab, rel, index=True)
    codec.check()
    size = int(size_t.item())
    assert size == want.size and np.array_equal(out[:size].cpu().numpy(), want)
    d2 = codec.decompress(out, size, n, index=idx, dtype=torch.float64)
    codec.check()
    assert np.array_equal(d2.cpu().numpy().view(np.uint64), ref_dec.view(np.uint64))


def test_f64_lower_level_entry_points(szp):
    """szp_double_openmp_threadblock(_arg) and
    szp_double_decompress_openmp_threadblock(_arg) (SZp/include/szp_double.h:30-34,
    SZp/include/szpd_double.h:21-23)."""
    x = make_input_f64("waves", 100001, 4)
    B, T, eb = 64, 8, 1e-5
    want = oracle_lib.compress(x, 0, eb, 0.0, B, T, dtype=np.float64)
    L = szp.lib
    szp.set_threads(T)
    try:
        size = ctypes.c_size_t()
        p = L.szp_double_openmp_threadblock(x.ctypes.data, ctypes.byref(size), eb, x.size, B)
        assert p
        got = np.ctypeslib.as_array((ctypes.c_uint8 * size.value).from_address(p)).copy()
        szp._libc.free(p)
        assert got.size == want.size and np.array_equal(got, want)

        cap = szp.compress_bound(x.size, B, T, szp.SZ_DOUBLE)
        buf = np.zeros(cap, np.uint8)
        size2 = ctypes.c_size_t()
        L.szp_double_openmp_threadblock_arg(buf.ctypes.data, x.ctypes.data, ctypes.byref(size2),
                                            eb, x.size, B)
        assert size2.value == want.size and np.array_equal(buf[:size2.value], want)

        ref = oracle_lib.decompress(want, x.size, B, T, dtype=np.float64)
        body = np.ascontiguousarray(want[8:])
        q = L.szp_double_decompress_openmp_threadblock(x.size, eb, B, body.ctypes.data)
        assert q
        d = np.ctypeslib.as_array((ctypes.c_double * x.size).from_address(q)).copy()
        szp._libc.free(q)
        assert np.array_equal(d.view(np.uint64), ref.view(np.uint64))
        d2 = np.zeros(x.size, np.float64)
        L.szp_double_decompress_openmp_threadblock_arg(d2.ctypes.data, x.size, eb, B,
                                                       body.ctypes.data)
        assert np.array_equal(d2.view(np.uint64), ref.view(np.uint64))
    finally:
        szp.set_threads(0)
