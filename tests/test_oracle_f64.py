"""CPU checks that pin the oracle's float64 path (szp_compress with
SZ_DOUBLE, SZp/src/szp_double.cc:148-318, SZp/src/szpd_double.cc:23-106):

* reference fixtures: streams and decompressions written by the reference
  compiled from its own sources (tests/golden/make_golden.py, cases_f64.*);
* closed forms: the 8-byte big-endian eb header, the FP64 quantizer's x86
  truncation rule at the int32 edges, and the error bound (the double path
  honours |x - x_hat| <= eb up to the rounding of q * eb);
* the double and float paths share their blocks: a float64 field holding
  float32 values, with an eb whose FP32 and FP64 reciprocals give the same
  quantized values, yields the float stream with a 4-byte longer header;
* live comparison with oracle/_ref when it is built (skipped otherwise).
"""
import json
import math
import os
import struct

import numpy as np
import pytest

import oracle_lib
from golden_inputs import make_input, make_input_f64

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def golden_f64():
    d = os.path.join(HERE, "golden")
    with open(os.path.join(d, "cases_f64.json")) as f:
        cases = json.load(f)
    return cases, np.load(os.path.join(d, "cases_f64.npz"))


def test_f64_fixtures(golden_f64):
    cases, arrays = golden_f64
    assert len(cases) >= 14
    for c in cases:
        k = c["id"]
        x = arrays["in_%d" % k].view(np.float64)
        assert np.array_equal(x.view(np.uint64),
                              make_input_f64(c["kind"], c["n"], c["seed"]).view(np.uint64))
        s = oracle_lib.compress(x, c["mode"], c["abs_eb"], c["rel"], c["B"], c["T"],
                                dtype=np.float64)
        ref = arrays["stream_%d" % k]
        assert s.size == ref.size and np.array_equal(s, ref), (k, c["kind"])
        d = oracle_lib.decompress(ref, c["n"], c["B"], c["T"], dtype=np.float64)
        assert np.array_equal(d.view(np.uint64), arrays["out_%d" % k]), (k, c["kind"])


def test_f64_header_is_big_endian_double():
    x = make_input_f64("walk", 1000, 3)
    s = oracle_lib.compress(x, 0, 0.0123, 0.0, 32, 4, dtype=np.float64)
    assert struct.unpack(">d", bytes(s[:8]))[0] == 0.0123
    offs = np.frombuffer(s[8:8 + 32].tobytes(), np.uint64)
    assert offs[0] == 0 and np.all(np.diff(offs.astype(np.int64)) > 0)


def test_quantize_f64_edges():
    O = oracle_lib.oracle()
    q = O.szpo_quantize_f64
    imin = -(1 << 31)
    assert q(float("nan"), 1.0) == imin
    assert q(float("inf"), 1.0) == imin
    assert q(2147483647.0, 1.0) == 2147483647
    assert q(2147483647.99, 1.0) == 2147483647
    assert q(2147483648.0, 1.0) == imin
    assert q(-2147483648.0, 1.0) == imin
    assert q(-2147483648.9, 1.0) == imin        # truncates to INT_MIN
    assert q(-2147483649.0, 1.0) == imin        # out of range
    assert q(-2.75, 1.0) == -2 and q(2.75, 1.0) == 2   # toward zero
    # The product is rounded to double before truncation: 0.1 * 30 rounds up
    # to 3.0000000000000004 in double -> 3, and 0.7 * 10 rounds to exactly 7.
    assert q(0.1, 30.0) == 3 and q(0.7, 10.0) == 7
    assert q(0.29, 100.0) == 28                 # 0.29 * 100 = 28.999999999999996


def test_f64_error_bound():
    x = make_input_f64("waves", 200000, 9)
    for rel in (1e-2, 1e-6):
        s = oracle_lib.compress(x, 1, 0.0, rel, 128, 8, dtype=np.float64)
        eb = struct.unpack(">d", bytes(s[:8]))[0]
        d = oracle_lib.decompress(s, x.size, 128, 8, dtype=np.float64)
        err = np.abs(x - d)
        assert np.all(err <= eb * (1 + 2.0 ** -40) + 2.0 ** -50 * np.abs(d))


def test_f64_shares_the_float_blocks():
    """Blocks depend only on q: a double field of float values with eb = 2^-10
    (exact reciprocal in both types) gives the float stream's chunks."""
    x32 = make_input("walk", 30001, 5)
    x64 = x32.astype(np.float64)
    eb = 2.0 ** -10
    s32 = oracle_lib.compress(x32, 0, eb, 0.0, 64, 5)
    s64 = oracle_lib.compress(x64, 0, eb, 0.0, 64, 5, dtype=np.float64)
    assert s64.size == s32.size + 4
    assert np.array_equal(s64[8:], s32[4:])
    d32 = oracle_lib.decompress(s32, x32.size, 64, 5)
    d64 = oracle_lib.decompress(s64, x32.size, 64, 5, dtype=np.float64)
    assert np.array_equal(d64.astype(np.float32), d32)


def test_rel_bound_f64_semantics():
    O = oracle_lib.oracle()
    x = np.array([np.nan, 1.0, 2.0], np.float64)
    assert math.isnan(O.szpo_rel_bound_f64(x.ctypes.data, x.size, 1e-3))
    x = np.array([1.0, np.nan, 3.0, -1.0], np.float64)
    assert O.szpo_rel_bound_f64(x.ctypes.data, x.size, 0.1) == 4.0 * 0.1


@pytest.mark.skipif(oracle_lib.reference() is None, reason="oracle/_ref not built")
def test_oracle_f64_matches_reference_live():
    rng = np.random.default_rng(5)
    n_checked = 0
    for T in (1, 3, 8):
        for B in (1, 7, 32, 128, 300):
            for n in (100, 4097, 30000):
                x = np.cumsum(rng.standard_normal(n)) * 0.05
                for mode, ab, rel in ((0, 0.01, 0.0), (1, 0.0, 1e-5)):
                    s = oracle_lib.compress(x, mode, ab, rel, B, T, dtype=np.float64)
                    if not oracle_lib.ref_fits(s, n, T, es=8):
                        continue
                    r = oracle_lib.ref_compress(x, mode, ab, rel, B, T, dtype=np.float64)
                    assert r.size == s.size and np.array_equal(r, s), (T, B, n, mode)
                    d1 = oracle_lib.ref_decompress(r, n, B, T, dtype=np.float64)
                    d2 = oracle_lib.decompress(s, n, B, T, dtype=np.float64)
                    assert np.array_equal(d1.view(np.uint64), d2.view(np.uint64))
                    n_checked += 1
    assert n_checked > 50
