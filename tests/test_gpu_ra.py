"""GPU parity of fastMode 2 ("random access", SZp/src/szp_float.cc:752-912,
SZp/src/szpd_float.cc:425-530, double: SZp/src/szp_double.cc:759-920,
SZp/src/szpd_double.cc:431-534) through the C ABI of libszp_amd.so, against
the reference's fixtures and the CPU oracle.  Bar: streams byte-identical,
decompressed values bit-identical.
"""
import ctypes
import json
import os
import sys

import numpy as np
import pytest

import oracle_lib
from golden_inputs import make_input, make_input_f64

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
    dt = szp.SZ_DOUBLE if x.dtype == np.float64 else szp.SZ_FLOAT
    szp.set_threads(T)
    try:
        stream, size = szp.szp_compress(2, dt, x, mode, ab, rel, x.size, B)
        assert size == stream.size
        dec = szp.szp_decompress(2, dt, stream, size, x.size, B)
    finally:
        szp.set_threads(0)
    return stream, dec


def test_ra_golden_fixtures(szp):
    d = os.path.join(HERE, "golden")
    cases = json.load(open(os.path.join(d, "cases_ra.json")))
    arrays = np.load(os.path.join(d, "cases_ra.npz"))
    for c in cases:
        k = c["id"]
        vt, ut = (np.float64, np.uint64) if c["dtype"] else (np.float32, np.uint32)
        x = arrays["in_%d" % k].view(vt)
        say("fastMode 2 golden case", k, c["dtype"], c["kind"],
            "n=%d B=%d T=%d" % (c["n"], c["B"], c["T"]))
        stream, dec = _roundtrip(szp, x, c["mode"], c["abs_eb"], c["rel"], c["B"], c["T"])
        ref = arrays["stream_%d" % k]
        assert stream.size == ref.size, (k, stream.size, ref.size)
        bad = np.nonzero(stream != ref)[0]
        assert bad.size == 0, (k, bad[:8])
        assert np.array_equal(dec.view(ut), arrays["out_%d" % k]), k


RA_CASES = [
    # dtype, kind, n, B, T, mode, abs_eb, rel
    (0, "walk", 1, 8, 1, 0, 0.01, 0.0),
    (0, "walk", 2, 8, 1, 0, 0.01, 0.0),
    (0, "walk", 17, 8, 2, 0, 0.01, 0.0),         # units of one value at chunk ends
    (0, "walk", 40, 8, 16, 0, 0.01, 0.0),
    (0, "walk", 3000, 1, 3, 0, 0.01, 0.0),       # every unit is one value
    (0, "walk", 4099, 5, 7, 0, 0.01, 0.0),
    (0, "walk", 70001, 9, 5, 1, 0.0, 1e-4),
    (0, "waves", (1 << 20) + 13, 128, 16, 1, 0.0, 1e-4),
    (0, "waves", 1 << 20, 32, 8, 1, 0.0, 1e-2),
    (0, "uniform", 300000, 64, 8, 0, 0.002, 0.0),
    (0, "zeroblocks", 500003, 32, 3, 1, 0.0, 1e-3),
    (0, "constant", 100000, 128, 4, 1, 0.0, 1e-3),
    (0, "specials", 200000, 32, 4, 0, 0.5, 0.0),
    (0, "sparsewide", 200000, 64, 2, 0, 1.0, 0.0),
    (0, "walk", 70000, 1000, 3, 0, 0.01, 0.0),
    (1, "waves", (1 << 20) + 1, 128, 16, 1, 0.0, 1e-6),
    (1, "walk", 70001, 9, 5, 0, 1e-9, 0.0),
    (1, "specials", 200000, 32, 4, 0, 0.5, 0.0),
]


@pytest.mark.parametrize("case", RA_CASES,
                         ids=lambda c: "%s-%s-n%d-B%d-T%d" % ("f64" if c[0] else "f32", c[1], c[2],
                                                              c[3], c[4]))
def test_ra_matches_oracle(szp, case):
    """The drop-in API (parse on decompress) and the device API with the side
    index, both against the oracle."""
    import torch
    dt, kind, n, B, T, mode, ab, rel = case
    x = (make_input_f64 if dt else make_input)(kind, n, 17 + n)
    say("fastMode 2 oracle case", case)
    want = oracle_lib.compress_mode(2, x, mode, ab, rel, B, T)
    vt = np.float64 if dt else np.float32
    ut = np.uint64 if dt else np.uint32
    ref_dec = oracle_lib.decompress_mode(2, want, n, B, T, dtype=vt)
    stream, dec = _roundtrip(szp, x, mode, ab, rel, B, T)
    assert stream.size == want.size
    bad = np.nonzero(stream != want)[0]
    assert bad.size == 0, bad[:8]
    assert np.array_equal(dec.view(ut), ref_dec.view(ut))

    codec = szp.DeviceCodec(block_size=B, threads=T, fast_mode=2)
    xd = torch.from_numpy(x).cuda()
    out, size_t, idx = codec.compress(xd, mode, ab, rel, index=True)
    codec.check()
    size = int(size_t.item())
    assert size == want.size and np.array_equal(out[:size].cpu().numpy(), want)
    d2 = codec.decompress(out, size, n, index=idx, dtype=xd.dtype)
    codec.check()
    assert np.array_equal(d2.cpu().numpy().view(ut), ref_dec.view(ut))


def test_ra_lower_level_entry_points(szp):
    """szp_float_openmp_threadblock_randomaccess(_arg) and
    szp_float_decompress_openmp_threadblock_randomaccess(_arg)
    (SZp/include/szp_float.h:42-48, SZp/include/szpd_float.h:29-31), and the
    double forms."""
    L = szp.lib
    B, T = 64, 8
    for dt, x, eb in ((0, make_input("waves", 100001, 4), 1e-3),
                      (1, make_input_f64("waves", 100001, 4), 1e-6)):
        want = oracle_lib.compress_mode(2, x, 0, eb, 0.0, B, T)
        vt = np.float64 if dt else np.float32
        ut = np.uint64 if dt else np.uint32
        ct = ctypes.c_double if dt else ctypes.c_float
        es = 8 if dt else 4
        pre = "szp_double" if dt else "szp_float"
        szp.set_threads(T)
        try:
            size = ctypes.c_size_t()
            p = getattr(L, pre + "_openmp_threadblock_randomaccess")(x.ctypes.data,
                                                                      ctypes.byref(size), eb,
                                                                      x.size, B)
            assert p
            got = np.ctypeslib.as_array((ctypes.c_uint8 * size.value).from_address(p)).copy()
            szp._libc.free(p)
            assert got.size == want.size and np.array_equal(got, want)
            cap = szp.compress_bound(x.size, B, T, dt, 2)
            buf = np.zeros(cap, np.uint8)
            size2 = ctypes.c_size_t()
            getattr(L, pre + "_openmp_threadblock_randomaccess_arg")(buf.ctypes.data, x.ctypes.data,
                                                                     ctypes.byref(size2), eb,
                                                                     x.size, B)
            assert size2.value == want.size and np.array_equal(buf[:size2.value], want)
            ref = oracle_lib.decompress_mode(2, want, x.size, B, T, dtype=vt)
            ebv = np.frombuffer(want[:es].tobytes()[::-1], vt)[0]
            body = np.ascontiguousarray(want[es:])
            q = getattr(L, pre + "_decompress_openmp_threadblock_randomaccess")(x.size, float(ebv),
                                                                                B, body.ctypes.data)
            assert q
            d = np.ctypeslib.as_array((ct * x.size).from_address(q)).copy()
            szp._libc.free(q)
            assert np.array_equal(d.view(ut), ref.view(ut))
            d2 = np.zeros(x.size, vt)
            getattr(L, pre + "_decompress_openmp_threadblock_randomaccess_arg")(
                d2.ctypes.data, x.size, float(ebv), B, body.ctypes.data)
            assert np.array_equal(d2.view(ut), ref.view(ut))
        finally:
            szp.set_threads(0)


def test_decompress_treats_other_modes_as_random_access(szp):
    """szp_decompress decodes every fastMode other than 1 as random access
    (SZp/src/szp.cc:135-151)."""
    x = make_input("walk", 5000, 3)
    want = oracle_lib.compress_mode(2, x, 0, 0.01, 0.0, 32, 4)
    ref = oracle_lib.decompress_mode(2, want, x.size, 32, 4)
    szp.set_threads(4)
    try:
        d = szp.szp_decompress(0, szp.SZ_FLOAT, want, want.size, x.size, 32)
    finally:
        szp.set_threads(0)
    assert np.array_equal(d.view(np.uint32), ref.view(np.uint32))
