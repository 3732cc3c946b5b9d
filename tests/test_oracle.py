"""CPU checks that pin the oracle (oracle/szp_oracle.c) before it is trusted
as the GPU path's checker.

* reference fixtures: streams and decompressions written by the reference
  compiled from its own sources (tests/golden/make_golden.py);
* an independent pure-Python decoder written from the stream format;
* a hand-derived known-answer stream;
* closed forms: the quantizer's x86 truncation rule, fl32(1/eb), the glibc
  log2f bit width over all 2^32 inputs, the chunk partition;
* live comparison with oracle/_ref when it is built (skipped otherwise).
"""
import ctypes
import hashlib
import json
import math
import os
import struct

import numpy as np
import pytest

import oracle_lib
from golden_inputs import HASH_CASES, make_input

HERE = os.path.dirname(os.path.abspath(__file__))


# ---------------------------------------------------------------- reference fixtures

def test_small_fixtures(golden):
    cases, arrays, _ = golden
    assert len(cases) >= 25
    for c in cases:
        k = c["id"]
        x = arrays["in_%d" % k].view(np.float32)
        s = oracle_lib.compress(x, c["mode"], c["abs_eb"], c["rel"], c["B"], c["T"])
        ref = arrays["stream_%d" % k]
        assert s.size == ref.size and np.array_equal(s, ref), (k, c["kind"])
        d = oracle_lib.decompress(ref, c["n"], c["B"], c["T"])
        assert np.array_equal(d.view(np.uint32), arrays["out_%d" % k]), (k, c["kind"])


def test_hash_fixtures(golden):
    _, _, hashes = golden
    assert len(hashes) == len(HASH_CASES)
    for h in hashes:
        x = make_input(h["kind"], h["n"], h["seed"])
        assert hashlib.sha256(x.tobytes()).hexdigest() == h["input_sha256"]
        s = oracle_lib.compress(x, h["mode"], h["abs_eb"], h["rel"], h["B"], h["T"])
        assert s.size == h["stream_bytes"]
        assert hashlib.sha256(s.tobytes()).hexdigest() == h["stream_sha256"]
        d = oracle_lib.decompress(s, h["n"], h["B"], h["T"])
        assert hashlib.sha256(d.tobytes()).hexdigest() == h["out_sha256"]


# ---------------------------------------------------------------- independent decoder

def _i32(v):
    v &= 0xFFFFFFFF
    return v - (1 << 32) if v >> 31 else v


def py_decode_block(buf, L):
    """One block from the format description: width byte, MSB-first signs,
    element-major little-endian bytes of mag >> r, MSB-first r-bit residuals."""
    b = int(buf[0])
    if b == 0:
        return [0] * L, 1
    assert b <= 32
    S = (L + 7) // 8
    signs = [(int(buf[1 + i // 8]) >> (7 - i % 8)) & 1 for i in range(L)]
    nb, r = b // 8, b % 8
    p = 1 + S
    hi = [int.from_bytes(bytes(buf[p + i * nb:p + (i + 1) * nb]), "little") for i in range(L)]
    p += nb * L
    nres = (r * L + 7) // 8
    bits = "".join(format(int(v), "08b") for v in buf[p:p + nres])
    lo = [int(bits[i * r:(i + 1) * r], 2) if r else 0 for i in range(L)]
    out = []
    for i in range(L):
        mag = (hi[i] << r) | lo[i]
        out.append(_i32(-mag if signs[i] else mag))
    return out, p + nres


def py_decompress(stream, n, B, T):
    """Whole-stream decoder: BE eb, T u64 offsets, per chunk a raw q then blocks;
    values fl32((float)q * eb)."""
    eb = struct.unpack(">f", bytes(stream[:4]))[0]
    eb32 = np.float32(eb)
    body = 4 + 8 * T
    offs = struct.unpack("<%dQ" % T, bytes(stream[4:body]))
    tbs = n // T
    out = np.zeros(n, np.float32)
    for t in range(T):
        lo = t * tbs
        hi = n if t == T - 1 else (t + 1) * tbs
        if lo >= hi:
            continue
        p = body + offs[t]
        q = struct.unpack("<i", bytes(stream[p:p + 4]))[0]
        p += 4
        out[lo] = np.float32(q) * eb32
        i = lo + 1
        while i < hi:
            L = min(B, hi - i)
            d, used = py_decode_block(stream[p:], L)
            p += used
            for k in range(L):
                q = _i32(q + d[k])
                out[i + k] = np.float32(q) * eb32
            i += L
    return out


def test_independent_decoder_on_fixtures(golden):
    cases, arrays, _ = golden
    for c in cases:
        if c["n"] > 6000:
            continue
        k = c["id"]
        got = py_decompress(arrays["stream_%d" % k], c["n"], c["B"], c["T"])
        assert np.array_equal(got.view(np.uint32), arrays["out_%d" % k]), (k, c["kind"])


def test_block_codec_every_width():
    """Oracle block encoder vs the independent decoder, every width 0..32 and
    block lengths 1..17, 32, 128 (short last blocks included)."""
    O = oracle_lib.oracle()
    rng = np.random.default_rng(5)
    for L in list(range(1, 18)) + [32, 128]:
        for b in range(0, 33):
            if b == 0:
                mags = np.zeros(L, np.uint64)
            else:
                mags = rng.integers(0, 1 << (b - 1), L, dtype=np.uint64)
                mags[rng.integers(0, L)] = (1 << (b - 1)) + rng.integers(0, 1 << (b - 1))
                mags = np.minimum(mags, 1 << 31)
            signs = rng.integers(0, 2, L)
            d = np.array([_i32(-int(m) if s and m else int(m)) for m, s in zip(mags, signs)],
                         dtype=np.int32)
            buf = np.zeros(1 + (L + 7) // 8 + 4 * L + 8, np.uint8)
            used = O.szpo_encode_block(d.ctypes.data, L, buf.ctypes.data)
            mx = int(np.max(np.abs(d.astype(np.int64)))) if L else 0
            width = O.szpo_bit_width(min(mx, 0xFFFFFFFF))
            assert buf[0] == width
            assert used == O.szpo_block_bytes(width, L)
            back, n2 = py_decode_block(buf, L)
            assert n2 == used
            assert back == [int(v) for v in d], (L, b)


def test_hand_derived_stream():
    """x = [5,6,4,4,4,4,4,4,4,4,1000,-3], ABS eb = 1, B = 4, T = 1, derived by hand:
    eb 1.0 -> 3f800000 (big-endian); one chunk offset 0; q0 = 5;
    block [6,4,4,4]: deltas 1,-2,0,0 -> width 2, signs 0100_0000, residuals 01 10 00 00;
    block [4,4,4,4]: all-zero deltas -> width byte 0 only;
    block [4,1000,-3]: deltas 0,996,-1003 -> width 10, signs 0010_0000,
    bytes of mag>>2 = 00 f9 fa, 2-bit residuals 00 00 11 -> 0c."""
    x = np.array([5, 6, 4, 4, 4, 4, 4, 4, 4, 4, 1000, -3], np.float32)
    want = bytes.fromhex("3f800000" "0000000000000000" "05000000" "024060" "00" "0a20" "00f9fa" "0c")
    s = oracle_lib.compress(x, 0, 1.0, 0.0, 4, 1)
    assert bytes(s) == want
    assert np.array_equal(oracle_lib.decompress(s, 12, 4, 1), x)


# ---------------------------------------------------------------- closed forms

def _x86_trunc(p):
    if math.isnan(p) or not (-2147483649.0 < p < 2147483648.0):
        return -(1 << 31)
    return int(p)  # Python int() truncates toward zero


def test_quantizer_x86_rule():
    O = oracle_lib.oracle()
    inv = np.float32(1.0) / np.float32(3e-3)
    specials = [0.0, -0.0, 1e-3, -1e-3, 0.4999, 1.5, -1.5, 6.4e6, -6.4e6, 6.5e6, -6.5e6,
                float("nan"), float("inf"), float("-inf"), 3.4e38, -3.4e38, 1e-45]
    rng = np.random.default_rng(1)
    vals = [np.float32(v) for v in specials] + list(rng.standard_normal(2000).astype(np.float32) * 1e7)
    for v in vals:
        p = float(np.float32(v)) * float(inv)  # exact in double
        assert O.szpo_quantize(float(v), float(inv)) == _x86_trunc(p), v
    # boundaries of the int32 range with inv = 1
    for v in (2147483520.0, 2147483648.0, -2147483648.0, -2147483904.0):
        assert O.szpo_quantize(v, 1.0) == _x86_trunc(float(np.float32(v)))


def test_inverse_bound_is_fp32_division():
    O = oracle_lib.oracle()
    for eb in (1e-4, 2e-4, 3e-3, 0.1, 1.0, 5e-39, 0.0, 7.0):
        want = np.float32(1.0) / np.float32(eb) if eb else np.float32(np.inf)
        got = np.float32(O.szpo_inverse_bound(float(np.float32(eb))))
        assert got.view(np.uint32) == np.float32(want).view(np.uint32)


def test_bit_width_table_exhaustive():
    """clz + threshold table == (int)log2f((float)m) + 1 for every u32."""
    O = oracle_lib.oracle()
    first = ctypes.c_uint32()
    assert O.szpo_check_bit_width_range(0, 0xFFFFFFFF, ctypes.byref(first)) == 0


def test_bit_width_fixture():
    with open(os.path.join(HERE, "golden", "bitwidth.json")) as f:
        w = json.load(f)["widths"]
    O = oracle_lib.oracle()
    for m, b in w.items():
        assert O.szpo_bit_width(int(m)) == b
        assert O.szpo_bit_width_table(int(m)) == b
        # the width is never below the significant bits, never more than one above
        assert int(m).bit_length() <= b <= int(m).bit_length() + 1


def test_chunk_partition():
    O = oracle_lib.oracle()
    O.szpo_chunk.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_int,
                             ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64)]
    lo, hi = ctypes.c_uint64(), ctypes.c_uint64()
    for n, T in [(0, 8), (5, 8), (8, 8), (17, 8), (1000, 7), (1 << 32, 1), (1 << 32, 3), (1 << 32, 16)]:
        covered = 0
        for t in range(T):
            O.szpo_chunk(n, T, t, ctypes.byref(lo), ctypes.byref(hi))
            assert lo.value == t * (n // T)
            assert hi.value == (n if t == T - 1 else (t + 1) * (n // T))
            covered += hi.value - lo.value
        assert covered == n


def test_rel_bound_serial_semantics():
    O = oracle_lib.oracle()
    x = np.array([0.0, -0.0, 0.0], np.float32)
    eb = O.szpo_rel_bound(x.ctypes.data, x.size, 1e-3)
    assert eb == 0.0 and math.copysign(1, eb) == 1  # first zero kept: 0 - 0
    x = np.array([-0.0, 0.0, -0.0], np.float32)
    eb = O.szpo_rel_bound(x.ctypes.data, x.size, 1e-3)
    assert eb == 0.0  # range -0 - -0 = +0
    x = np.array([np.nan, 1.0, 2.0], np.float32)
    assert math.isnan(O.szpo_rel_bound(x.ctypes.data, x.size, 1e-3))
    x = np.array([1.0, np.nan, 3.0, -1.0], np.float32)
    eb = O.szpo_rel_bound(x.ctypes.data, x.size, 0.5)
    assert eb == np.float32(np.float32(4.0) * np.float32(0.5))


def test_error_bound_closed_form():
    """|x - x_hat| <= eb + 2^-22 (eb + |x_hat|) on a smooth field (normal range)."""
    x = make_input("waves", 200000, 9)
    for rel in (1e-2, 1e-4):
        s = oracle_lib.compress(x, 1, 0.0, rel, 128, 8)
        eb = struct.unpack(">f", bytes(s[:4]))[0]
        d = oracle_lib.decompress(s, x.size, 128, 8)
        err = np.abs(x.astype(np.float64) - d.astype(np.float64))
        assert np.all(err <= eb + 2.0 ** -22 * (eb + np.abs(d.astype(np.float64))))


# ---------------------------------------------------------------- live reference

@pytest.mark.skipif(oracle_lib.reference() is None, reason="oracle/_ref not built")
def test_oracle_matches_reference_live():
    rng = np.random.default_rng(3)
    n_checked = 0
    for T in (1, 3, 8):
        for B in (1, 7, 32, 128, 300):
            for n in (100, 4097, 30000):
                x = (np.cumsum(rng.standard_normal(n)) * 0.05).astype(np.float32)
                for mode, ab, rel in ((0, 0.01, 0.0), (1, 0.0, 1e-3)):
                    s = oracle_lib.compress(x, mode, ab, rel, B, T)
                    if not oracle_lib.ref_fits(s, n, T):
                        continue
                    r = oracle_lib.ref_compress(x, mode, ab, rel, B, T)
                    assert r.size == s.size and np.array_equal(r, s), (T, B, n, mode)
                    d1 = oracle_lib.ref_decompress(r, n, B, T)
                    d2 = oracle_lib.decompress(s, n, B, T)
                    assert np.array_equal(d1.view(np.uint32), d2.view(np.uint32))
                    n_checked += 1
    assert n_checked > 50


def test_record_count_matches_reference_live():
    """The oracle's tally of the reference's instrumented single-thread forms
    (SZp/src/szp_float.cc:607-750, SZp/src/szpd_float.cc:216-317) equals the
    count the reference itself returns, and the reference's single-thread
    stream is the oracle's T = 1 stream."""
    if oracle_lib.reference() is None:
        pytest.skip("oracle/_ref not built (needs the reference sources)")
    rng = np.random.default_rng(5)
    checked = 0
    for B in (1, 8, 9, 32, 128):
        for n in (0, 1, 2, 9, 129, 5000):
            x = (np.cumsum(rng.standard_normal(n)) * 0.05).astype(np.float32)
            x[n // 3: n // 2] = 0.0  # runs of zero-width blocks
            stream, c1, dec, c2 = oracle_lib.ref_single_thread_record(x, np.float32(0.01), B)
            mine = oracle_lib.compress(x, 0, 0.01, 0.0, B, 1)
            assert np.array_equal(stream, mine), (B, n)
            assert oracle_lib.record_count(stream, n, B, False) == c1, (B, n)
            assert oracle_lib.record_count(stream, n, B, True) == c2, (B, n)
            ref = oracle_lib.decompress(mine, n, B, 1)
            assert np.array_equal(dec.view(np.uint32), ref.view(np.uint32))
            checked += 1
    assert checked == 30
