"""Host emulation of the stream parser (tests/cpu/parse_check.cc).

The kernel that recovers block boundaries from a stream without a side index
(k_parse) runs speculative chains per sub-tile, a candidate check, and a
serial true-chain resolution; the per-thread walks live in
szp_amd/csrc/szp_geom.h and are shared with the emulation.  Here the
emulation runs under AddressSanitizer on oracle streams (zero runs, constant
fields, every block width, tiny and huge blocks) and on damaged streams
(garbage tails, truncation, flipped width bytes, bad chunk offsets): the
recovered positions must equal a serial walk, and a damaged stream must be
reported exactly when the serial walk rejects it.
"""
import os
import subprocess

import numpy as np
import pytest

import oracle_lib
from golden_inputs import make_input

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

VALID = [
    # kind, n, B, T, mode, abs_eb, rel
    ("waves", 1 << 20, 128, 4, 1, 0.0, 1e-4),
    ("waves", 1 << 20, 32, 3, 1, 0.0, 1e-2),
    ("waves", 1 << 20, 8, 2, 1, 0.0, 1e-5),
    ("waves", (1 << 19) + 7, 1, 2, 1, 0.0, 1e-3),
    ("walk", 300000, 2048, 2, 1, 0.0, 1e-4),
    ("walk", 200000, 9, 5, 0, 0.01, 0.0),
    ("uniform", 400000, 128, 3, 0, 0.002, 0.0),
    ("uniform", 100000, 2048, 1, 0, 1e-6, 0.0),
    ("zeroblocks", 800003, 128, 3, 1, 0.0, 1e-3),
    ("zeroblocks", 500000, 16, 2, 1, 0.0, 1e-3),
    ("constant", 2000000, 8, 2, 1, 0.0, 1e-3),
    ("constant", 300000, 1, 3, 0, 0.5, 0.0),
    ("specials", 300000, 32, 4, 0, 0.5, 0.0),
    ("sparsewide", 400000, 64, 2, 0, 1.0, 0.0),
    ("signedzero", 200000, 32, 4, 1, 0.0, 1e-3),
    ("walk", 40, 8, 16, 0, 0.01, 0.0),
    ("walk", 3, 8, 8, 0, 0.01, 0.0),
]


def _stream(case, seed):
    kind, n, B, T, mode, ab, rel = case
    x = make_input(kind, n, seed)
    return oracle_lib.compress(x, mode, ab, rel, B, T)


def _width_positions(st, n, B, T):
    """Positions of every block's width byte (serial walk of a valid stream)."""
    body = 4 + 8 * T
    offs = [int.from_bytes(st[4 + 8 * t:12 + 8 * t].tobytes(), "little") for t in range(T)]
    O = oracle_lib.oracle()
    per = n // T
    out = []
    for t in range(T):
        nt = per if t < T - 1 else n - per * (T - 1)
        if nt == 0:
            continue
        p = body + offs[t] + 4
        nb = (nt - 1 + B - 1) // B
        for j in range(nb):
            L = B if j + 1 < nb else (nt - 1) - (nb - 1) * B
            out.append(p)
            p += O.szpo_block_bytes(int(st[p]), L)
    return out


@pytest.fixture(scope="module")
def parse_check(tmp_path_factory):
    d = tmp_path_factory.mktemp("parse")
    exe = str(d / "parse_check")
    subprocess.run(["g++", "-O1", "-g", "-std=c++17", "-Wall", "-fsanitize=address,undefined",
                    "-fno-sanitize-recover=undefined", "-I", os.path.join(ROOT, "szp_amd", "csrc"),
                    os.path.join(ROOT, "tests", "cpu", "parse_check.cc"), "-o", exe], check=True)
    return exe, d


def _run(parse_check, entries):
    exe, d = parse_check
    lst = d / ("cases_%d.txt" % len(os.listdir(d)))
    lines = []
    for k, e in enumerate(entries):
        st, n, B, T = e[:4]
        mode = e[4] if len(e) > 4 else 1
        p = d / ("s_%s_%d.bin" % (lst.stem, k))
        np.ascontiguousarray(st, dtype=np.uint8).tofile(p)
        lines.append("%s %d %d %d %d" % (p, n, B, T, mode))
    lst.write_text("\n".join(lines) + "\n")
    r = subprocess.run([exe, str(lst)], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr[-4000:]
    assert "0 failures" in r.stdout
    print(r.stdout)
    return r.stdout


def test_parse_valid_streams(parse_check):
    entries = []
    for k, c in enumerate(VALID):
        st = _stream(c, 50 + k)
        entries.append((st, c[1], c[2], c[3]))
    out = _run(parse_check, entries)
    assert "(0 invalid streams)" in out, out


def test_parse_upper_bound_lengths(parse_check):
    """Garbage after the last chunk (the caller passed a buffer size)."""
    rng = np.random.default_rng(3)
    entries = []
    for k, c in enumerate(VALID[:8] + VALID[8:12]):
        st = _stream(c, 80 + k)
        for tail, kind in ((7, "rand"), (70000, "rand"), (5000, "zero"), (3000, "small")):
            if kind == "rand":
                t = rng.integers(0, 256, tail, dtype=np.uint8)
            elif kind == "zero":
                t = np.zeros(tail, np.uint8)
            else:
                t = rng.integers(0, 33, tail, dtype=np.uint8)
            entries.append((np.concatenate([st, t]), c[1], c[2], c[3]))
    out = _run(parse_check, entries)
    assert "(0 invalid streams)" in out, out


def test_parse_damaged_streams(parse_check):
    """Truncation, flipped width bytes and bad chunk offsets: the parser's
    verdict must equal the serial walk's, and its positions too whenever the
    serial walk still accepts the stream."""
    rng = np.random.default_rng(4)
    entries = []
    for k, c in enumerate(VALID[:14]):
        kind, n, B, T = c[0], c[1], c[2], c[3]
        st = _stream(c, 120 + k)
        widths = _width_positions(st, n, B, T)
        # truncated by one byte and by a large amount
        entries.append((st[:-1].copy(), n, B, T))
        entries.append((st[:max(4 + 8 * T, st.size * 2 // 3)].copy(), n, B, T))
        for trial in range(6):
            s2 = st.copy()
            if widths:
                p = widths[int(rng.integers(0, len(widths)))]
                choice = trial % 3
                if choice == 0:
                    s2[p] = 33 + int(rng.integers(0, 223))  # invalid width
                elif choice == 1:
                    s2[p] = int(rng.integers(0, 33))        # another valid width
                else:
                    s2[p] = 0                               # a zero block
            entries.append((s2, n, B, T))
        if T > 1:
            s3 = st.copy()
            s3[4 + 8:4 + 16] = np.frombuffer(np.uint64(1 << 40).tobytes(), np.uint8)
            entries.append((s3, n, B, T))
            s4 = st.copy()
            off = int.from_bytes(st[12:20].tobytes(), "little")
            s4[12:20] = np.frombuffer(np.uint64(off + 1).tobytes(), np.uint8)
            entries.append((s4, n, B, T))
        s5 = st.copy()
        s5[4 + 8 * T:] = rng.integers(0, 256, s5.size - 4 - 8 * T, dtype=np.uint8)
        entries.append((s5, n, B, T))
    out = _run(parse_check, entries)
    ninv = int(out.split("(")[1].split()[0])
    assert ninv > 40, out


# fastMode 2 (units): B >= 2 (B = 1 streams decode without a parse).
VALID_RA = [
    ("waves", 1 << 20, 128, 4, 1, 0.0, 1e-4),
    ("waves", 1 << 20, 32, 3, 1, 0.0, 1e-2),
    ("waves", (1 << 19) + 7, 2, 2, 1, 0.0, 1e-3),
    ("walk", 300001, 2048, 2, 1, 0.0, 1e-4),
    ("walk", 200001, 9, 5, 0, 0.01, 0.0),
    ("uniform", 400000, 128, 3, 0, 0.002, 0.0),
    ("zeroblocks", 800003, 64, 3, 1, 0.0, 1e-3),
    ("constant", 2000000, 8, 2, 1, 0.0, 1e-3),
    ("specials", 300000, 32, 4, 0, 0.5, 0.0),
    ("sparsewide", 400000, 64, 2, 0, 1.0, 0.0),
    ("walk", 41, 8, 16, 0, 0.01, 0.0),
    ("walk", 17, 8, 2, 0, 0.01, 0.0),
]


def _ra_stream(case, seed):
    kind, n, B, T, mode, ab, rel = case
    return oracle_lib.compress_mode(2, make_input(kind, n, seed), mode, ab, rel, B, T)


def _unit_width_positions(st, n, B, T):
    """Positions of the width byte of every unit with more than one value."""
    offs = [int.from_bytes(st[4 + 8 * t:12 + 8 * t].tobytes(), "little") for t in range(T)]
    body = 4 + 8 * T
    per = n // T
    O = oracle_lib.oracle()
    out = []
    for t in range(T):
        nt = per if t < T - 1 else n - per * (T - 1)
        p = body + offs[t]
        for i in range(0, nt, B):
            L = min(B, nt - i) - 1
            if L:
                out.append(p + 4)
                p += 4 + O.szpo_block_bytes(int(st[p + 4]), L)
            else:
                p += 5
    return out


def test_parse_random_access_streams(parse_check):
    """fastMode 2: valid streams, garbage tails, truncation, flipped width
    bytes and bad chunk offsets, against the serial walk of the units."""
    rng = np.random.default_rng(5)
    entries = []
    for k, c in enumerate(VALID_RA):
        kind, n, B, T = c[0], c[1], c[2], c[3]
        st = _ra_stream(c, 200 + k)
        entries.append((st, n, B, T, 2))
        entries.append((np.concatenate([st, rng.integers(0, 256, 5000, dtype=np.uint8)]), n, B, T, 2))
        entries.append((st[:-1].copy(), n, B, T, 2))
        widths = _unit_width_positions(st, n, B, T)
        for trial in range(4):
            s2 = st.copy()
            if widths:
                p = widths[int(rng.integers(0, len(widths)))]
                s2[p] = (33 + int(rng.integers(0, 223))) if trial % 2 == 0 else int(rng.integers(0, 33))
            entries.append((s2, n, B, T, 2))
        if T > 1:
            s3 = st.copy()
            off = int.from_bytes(st[12:20].tobytes(), "little")
            s3[12:20] = np.frombuffer(np.uint64(off + 3).tobytes(), np.uint8)
            entries.append((s3, n, B, T, 2))
    out = _run(parse_check, entries)
    ninv = int(out.split("(")[1].split()[0])
    assert ninv > 10, out
