"""CPU checks that pin the oracle's fastMode 2 ("random access",
SZp/src/szp_float.cc:752-912, SZp/src/szpd_float.cc:425-530, and the double
forms SZp/src/szp_double.cc:759-920, SZp/src/szpd_double.cc:431-534):

* reference fixtures (tests/golden/make_golden.py, cases_ra.*), float32 and
  float64, including chunks whose last unit holds one value and units of one
  value (B = 1);
* an independent pure-Python decoder written from the format: each chunk is
  a run of units [q(first) i32][width byte][signs][planes][residuals];
* closed forms: the unit count and byte length of a field whose deltas are
  all zero (every unit is 4 + 1 bytes), and that the units restart the
  Lorenzo chain (a unit decodes alone, from its own bytes);
* live comparison with oracle/_ref when it is built (skipped otherwise).
"""
import json
import os
import struct

import numpy as np
import pytest

import oracle_lib
from golden_inputs import make_input, make_input_f64

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def golden_ra():
    d = os.path.join(HERE, "golden")
    with open(os.path.join(d, "cases_ra.json")) as f:
        cases = json.load(f)
    return cases, np.load(os.path.join(d, "cases_ra.npz"))


def _input(c):
    return (make_input_f64 if c["dtype"] else make_input)(c["kind"], c["n"], c["seed"])


def test_ra_fixtures(golden_ra):
    cases, arrays = golden_ra
    assert len(cases) >= 20
    for c in cases:
        k = c["id"]
        ut = np.uint64 if c["dtype"] else np.uint32
        vt = np.float64 if c["dtype"] else np.float32
        x = arrays["in_%d" % k].view(vt)
        assert np.array_equal(x.view(ut), _input(c).view(ut))
        s = oracle_lib.compress_mode(2, x, c["mode"], c["abs_eb"], c["rel"], c["B"], c["T"])
        ref = arrays["stream_%d" % k]
        assert s.size == ref.size and np.array_equal(s, ref), (k, c["kind"])
        d = oracle_lib.decompress_mode(2, ref, c["n"], c["B"], c["T"], dtype=vt)
        assert np.array_equal(d.view(ut), arrays["out_%d" % k]), (k, c["kind"])


def _i32(v):
    v &= 0xFFFFFFFF
    return v - (1 << 32) if v >> 31 else v


def py_decompress_ra(stream, n, B, T, es=4):
    """Independent decoder of a fastMode-2 stream, written from the format."""
    stream = bytes(stream)
    eb = struct.unpack(">f" if es == 4 else ">d", bytes(stream[:es]))[0]
    offs = struct.unpack("<%dQ" % T, bytes(stream[es:es + 8 * T]))
    body = es + 8 * T
    tbs = n // T
    q = np.zeros(n, np.int64)
    for t in range(T):
        lo, hi = t * tbs, (n if t == T - 1 else (t + 1) * tbs)
        p = body + offs[t]
        for i in range(lo, hi, B):
            ln = min(B, hi - i)
            prior = _i32(int.from_bytes(stream[p:p + 4], "little"))
            p += 4
            q[i] = prior
            if ln == 1:
                # The reference's decompressor never reads a one-value unit's
                # width byte (SZp/src/szpd_float.cc:483), so with B = 1 it
                # reads every unit at a 4-byte stride although the units are
                # 5 bytes long; decoded exactly as it does.
                continue
            L = ln - 1
            b = stream[p]
            p += 1
            if b == 0:
                q[i + 1:i + ln] = prior
                continue
            signs = stream[p:p + (L + 7) // 8]
            p += (L + 7) // 8
            nb, r = b // 8, b % 8
            planes = stream[p:p + nb * L]
            p += nb * L
            res = stream[p:p + (r * L + 7) // 8]
            p += (r * L + 7) // 8
            bit = 0
            for j in range(L):
                v = int.from_bytes(planes[j * nb:(j + 1) * nb], "little") if nb else 0
                low = 0
                for _ in range(r):
                    low = (low << 1) | ((res[bit >> 3] >> (7 - (bit & 7))) & 1)
                    bit += 1
                m = (v << r) | low
                neg = (signs[j >> 3] >> (7 - (j & 7))) & 1
                prior = _i32(prior + (-m if neg else m))
                q[i + 1 + j] = prior
    vt = np.float32 if es == 4 else np.float64
    return q.astype(vt) * vt(eb)


def test_independent_decoder(golden_ra):
    cases, arrays = golden_ra
    for c in cases:
        if c["n"] > 20000:
            continue
        k = c["id"]
        es = 8 if c["dtype"] else 4
        ut = np.uint64 if c["dtype"] else np.uint32
        ref = arrays["stream_%d" % k]
        d = py_decompress_ra(ref, c["n"], c["B"], c["T"], es)
        assert np.array_equal(d.view(ut), arrays["out_%d" % k]), (k, c["kind"])


def test_constant_field_closed_form():
    """All deltas zero: every unit is its 4-byte head and a zero width byte."""
    n, B, T = 10007, 64, 3
    x = np.full(n, 0.25, np.float32)
    s = oracle_lib.compress_mode(2, x, 0, 1e-3, 0.0, B, T)
    tbs = n // T
    units = 2 * ((tbs + B - 1) // B) + (n - 2 * tbs + B - 1) // B
    assert s.size == 4 + 8 * T + 5 * units
    q = int(np.float64(np.float32(0.25)) * np.float64(np.float32(1) / np.float32(1e-3)))
    assert int.from_bytes(bytes(s[4 + 8 * T:4 + 8 * T + 4]), "little") == q


def test_units_restart_the_chain():
    """A unit's bytes decode alone: cutting a field at a unit boundary and
    compressing the halves gives the same unit bytes (T = 1)."""
    x = make_input("walk", 4096, 8)
    B = 128
    full = oracle_lib.compress_mode(2, x, 0, 0.01, 0.0, B, 1)
    a = oracle_lib.compress_mode(2, x[:2048], 0, 0.01, 0.0, B, 1)
    b = oracle_lib.compress_mode(2, x[2048:], 0, 0.01, 0.0, B, 1)
    assert np.array_equal(full[12:], np.concatenate([a[12:], b[12:]]))


@pytest.mark.skipif(oracle_lib.reference() is None, reason="oracle/_ref not built")
def test_oracle_ra_matches_reference_live():
    rng = np.random.default_rng(9)
    n_checked = 0
    for dt in (np.float32, np.float64):
        es = 8 if dt == np.float64 else 4
        for T in (1, 3, 8):
            for B in (2, 7, 32, 128, 300):
                for n in (100, 4097, 30001):
                    x = (np.cumsum(rng.standard_normal(n)) * 0.05).astype(dt)
                    for mode, ab, rel in ((0, 0.01, 0.0), (1, 0.0, 1e-4)):
                        s = oracle_lib.compress_mode(2, x, mode, ab, rel, B, T)
                        if not oracle_lib.ref_fits(s, n, T, es=es):
                            continue
                        r = oracle_lib.ref_compress(x, mode, ab, rel, B, T, dtype=dt, fast=2)
                        assert r.size == s.size and np.array_equal(r, s), (dt, T, B, n, mode)
                        d1 = oracle_lib.ref_decompress(r, n, B, T, dtype=dt, fast=2)
                        d2 = oracle_lib.decompress_mode(2, s, n, B, T, dtype=dt)
                        assert d1.tobytes() == d2.tobytes()
                        n_checked += 1
    assert n_checked > 100
