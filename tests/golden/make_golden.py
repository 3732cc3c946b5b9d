"""Writes the golden fixtures for SZp fastmode1 float32 from the reference itself.

Run in the development container, where the reference sources exist:

    make -C oracle ref            # oracle/_ref/libszp_ref.so from the reference's sources
    python tests/golden/make_golden.py

The reference library is driven through its own entry points
(`szp_compress` / `szp_decompress`, SZp/include/szp.h:33-36) with the OpenMP
team size pinned by oracle/ref_shim.c.  Outputs:

  cases.npz / cases.json   small inputs, reference streams and reference
                           decompressions (bit patterns), one entry per case
  hashes.json              SHA-256 of reference streams and decompressions
                           for larger inputs built from integer PRNG draws and
                           IEEE basic operations only (bit-reproducible on any
                           machine, so the inputs are regenerated, not stored)
  bitwidth.json            the reference's bit width at every threshold of the
                           glibc log2f table (szp_float.cc:261)
  cases_f64.npz / .json    the same for the float64 path (szp_compress with
                           SZ_DOUBLE, SZp/src/szp_double.cc:148-318)
  cases_ra.npz / .json     fastMode 2 (SZP_RANDOMACCESS) for float32 and
                           float64 (SZp/src/szp_float.cc:752-912,
                           SZp/src/szp_double.cc:759-920)

The reference allocates 4N+4 output bytes and ceil(4N/T) per chunk
(szp_float.cc:156,186); cases whose stream would not fit are generated with
the oracle's size check first and skipped, never passed to the reference.
"""
import ctypes
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "tests"))

from golden_inputs import make_input, make_input_f64, HASH_CASES  # noqa: E402

REF = os.path.join(ROOT, "oracle", "_ref", "libszp_ref.so")
ORACLE = os.path.join(ROOT, "oracle", "libszp_oracle.so")


def load():
    R = ctypes.CDLL(REF)
    R.szp_compress.restype = ctypes.c_void_p
    R.szp_compress.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                               ctypes.POINTER(ctypes.c_size_t), ctypes.c_int, ctypes.c_double,
                               ctypes.c_double, ctypes.c_size_t, ctypes.c_int]
    R.szp_decompress.restype = ctypes.c_void_p
    R.szp_decompress.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t,
                                 ctypes.c_size_t, ctypes.c_int]
    R.szp_ref_set_threads.argtypes = [ctypes.c_int]
    O = ctypes.CDLL(ORACLE)
    O.szpo_compress_bound.restype = ctypes.c_uint64
    O.szpo_compress_bound.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_int]
    O.szpo_compress_api.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int, ctypes.c_double,
                                    ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_void_p, ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64)]
    O.szpo_compress_api_f64.argtypes = O.szpo_compress_api.argtypes
    O.szpo_compress_mode.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_uint64,
                                     ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                     ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_uint64,
                                     ctypes.POINTER(ctypes.c_uint64)]
    O.szpo_compress_bound_mode.restype = ctypes.c_uint64
    O.szpo_compress_bound_mode.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_uint64,
                                           ctypes.c_int, ctypes.c_int]
    O.szpo_bit_width.argtypes = [ctypes.c_uint32]
    libc = ctypes.CDLL("libc.so.6")
    libc.free.argtypes = [ctypes.c_void_p]
    return R, O, libc


def fits_reference(O, x, mode, ab, rel, B, T):
    """True when the reference's fixed-size buffers hold this stream."""
    n = x.size
    cap = O.szpo_compress_bound(n, B, T)
    out = np.zeros(cap, np.uint8)
    sz = ctypes.c_uint64()
    if O.szpo_compress_api(x.ctypes.data, n, mode, ab, rel, B, T, 0, out.ctypes.data, cap,
                           ctypes.byref(sz)):
        return False
    hdr = 4 + 8 * T
    offs = np.frombuffer(out[4:hdr].tobytes(), np.uint64).astype(np.int64)
    sizes = np.diff(np.append(offs, sz.value - hdr))
    return sz.value <= 4 * n + 4 and sizes.max() <= (4 * n + T - 1) // T


def fits_reference_f64(O, x, mode, ab, rel, B, T):
    """The float64 form: 8N+8 bytes in all, ceil(8N/T) per chunk
    (SZp/src/szp_double.cc:157,186)."""
    n = x.size
    cap = O.szpo_compress_bound(n, B, T) + 4
    out = np.zeros(cap, np.uint8)
    sz = ctypes.c_uint64()
    if O.szpo_compress_api_f64(x.ctypes.data, n, mode, ab, rel, B, T, 0, out.ctypes.data, cap,
                               ctypes.byref(sz)):
        return False
    hdr = 8 + 8 * T
    offs = np.frombuffer(out[8:hdr].tobytes(), np.uint64).astype(np.int64)
    sizes = np.diff(np.append(offs, sz.value - hdr))
    return sz.value <= 8 * n + 8 and sizes.max() <= (8 * n + T - 1) // T


def fits_reference_mode(O, x, fast, dtype, mode, ab, rel, B, T):
    """Any fastMode / dataType: es*N+es bytes in all and ceil(es*N/T) per
    chunk (fastMode 2: SZp/src/szp_float.cc:757-790, SZp/src/szp_double.cc:764-797)."""
    n = x.size
    es = 8 if dtype else 4
    cap = O.szpo_compress_bound_mode(fast, dtype, n, B, T)
    out = np.zeros(cap, np.uint8)
    sz = ctypes.c_uint64()
    if O.szpo_compress_mode(fast, dtype, x.ctypes.data, n, mode, ab, rel, B, T, 0,
                            out.ctypes.data, cap, ctypes.byref(sz)):
        return False
    hdr = es + 8 * T
    offs = np.frombuffer(out[es:hdr].tobytes(), np.uint64).astype(np.int64)
    sizes = np.diff(np.append(offs, sz.value - hdr))
    return sz.value <= es * n + es and sizes.max() <= (es * n + T - 1) // T


def ref_roundtrip(R, libc, x, mode, ab, rel, B, T, dtype=0, fast=1):
    R.szp_ref_set_threads(T)
    sz = ctypes.c_size_t()
    p = R.szp_compress(fast, dtype, x.ctypes.data, ctypes.byref(sz), mode, ab, rel, x.size, B)
    stream = np.ctypeslib.as_array((ctypes.c_uint8 * sz.value).from_address(p)).copy()
    libc.free(p)
    q = R.szp_decompress(fast, dtype, stream.ctypes.data, stream.size, x.size, B)
    ct = ctypes.c_uint64 if dtype == 1 else ctypes.c_uint32
    dec = np.ctypeslib.as_array((ct * x.size).from_address(q)).copy()
    libc.free(q)
    return stream, dec


# (kind, n, seed, mode, abs_eb, rel, B, T)
SMALL_CASES = [
    ("walk", 12, 1, 0, 1.0, 0.0, 4, 1),
    ("walk", 1000, 2, 0, 0.01, 0.0, 8, 1),
    ("walk", 1000, 3, 1, 0.0, 1e-3, 32, 2),
    ("walk", 4097, 4, 1, 0.0, 1e-4, 128, 8),
    ("walk", 20000, 5, 1, 0.0, 1e-2, 128, 8),
    ("walk", 20000, 6, 1, 0.0, 1e-4, 32, 16),
    ("walk", 20001, 7, 1, 0.0, 1e-3, 64, 3),
    ("waves", 16384, 8, 1, 0.0, 1e-3, 128, 8),
    ("waves", 16384, 9, 1, 0.0, 1e-4, 32, 4),
    ("uniform", 8192, 10, 0, 0.002, 0.0, 128, 8),
    ("uniform", 8193, 11, 1, 0.0, 1e-5, 64, 7),
    ("zeroblocks", 16384, 12, 1, 0.0, 1e-3, 128, 8),
    ("zeroblocks", 16385, 13, 1, 0.0, 1e-3, 32, 5),
    ("constant", 3000, 14, 0, 0.1, 0.0, 32, 4),
    ("constant", 3000, 15, 1, 0.0, 1e-3, 32, 4),      # range 0 -> eb 0 -> q = INT_MIN
    ("walk", 5000, 16, 0, 0.01, 0.0, 1, 4),
    ("walk", 5000, 17, 0, 0.01, 0.0, 3, 2),
    ("walk", 5000, 18, 0, 0.01, 0.0, 100, 6),
    ("walk", 5000, 19, 0, 0.01, 0.0, 255, 2),
    ("walk", 5000, 20, 0, 0.01, 0.0, 1024, 1),
    ("walk", 200, 21, 0, 0.05, 0.0, 8, 16),           # chunks of 12 elements
    ("specials", 6000, 23, 0, 0.5, 0.0, 32, 4),       # NaN, +-Inf, int32 overflow
    ("fewspecials", 6000, 24, 0, 1e-3, 0.0, 128, 3),
    ("sparsewide", 6000, 25, 0, 1.0, 0.0, 64, 2),     # widths up to 32
    ("sparsewide", 6000, 26, 0, 0.25, 0.0, 16, 5),
    ("tiny", 4000, 27, 0, 5e-39, 0.0, 32, 2),         # subnormal eb, finite 1/eb
    ("signedzero", 4000, 28, 1, 0.0, 1e-3, 32, 4),
    ("walk", 9000, 29, 1, 0.0, 0.5, 128, 8),
    ("nanfirst", 2000, 30, 1, 0.0, 1e-3, 32, 2),      # x[0] = NaN -> eb = NaN
]


# float64 cases (same tuple layout): SZ_DOUBLE through the reference.
SMALL_CASES_F64 = [
    ("walk", 12, 1, 0, 1.0, 0.0, 4, 1),
    ("walk", 1000, 2, 0, 0.01, 0.0, 8, 1),
    ("walk", 4097, 4, 1, 0.0, 1e-4, 128, 8),
    ("walk", 20001, 7, 1, 0.0, 1e-3, 64, 3),
    ("waves", 16384, 8, 1, 0.0, 1e-3, 128, 8),
    ("waves", 16384, 9, 1, 0.0, 1e-6, 32, 4),
    ("uniform", 8193, 11, 1, 0.0, 1e-5, 64, 7),
    ("zeroblocks", 16385, 13, 1, 0.0, 1e-3, 32, 5),
    ("constant", 3000, 15, 1, 0.0, 1e-3, 32, 4),
    ("walk", 5000, 17, 0, 0.01, 0.0, 3, 2),
    ("walk", 5000, 19, 0, 1e-9, 0.0, 255, 2),
    ("walk", 200, 21, 0, 0.05, 0.0, 8, 16),
    ("specials", 6000, 23, 0, 0.5, 0.0, 32, 4),
    ("sparsewide", 6000, 25, 0, 1.0, 0.0, 64, 2),
    ("signedzero", 4000, 28, 1, 0.0, 1e-3, 32, 4),
    ("nanfirst", 2000, 30, 1, 0.0, 1e-3, 32, 2),
]


def make_f64():
    R, O, libc = load()
    arrays = {}
    index = []
    for k, (kind, n, seed, mode, ab, rel, B, T) in enumerate(SMALL_CASES_F64):
        x = make_input_f64(kind, n, seed)
        if not fits_reference_f64(O, x, mode, ab, rel, B, T):
            print("skip f64 (reference buffer too small):", kind, n, B, T)
            continue
        stream, dec = ref_roundtrip(R, libc, x, mode, ab, rel, B, T, dtype=1)
        arrays["in_%d" % k] = x.view(np.uint64)
        arrays["stream_%d" % k] = stream
        arrays["out_%d" % k] = dec
        index.append(dict(id=k, kind=kind, n=n, seed=seed, mode=mode, abs_eb=ab, rel=rel, B=B, T=T,
                          stream_bytes=int(stream.size),
                          stream_sha256=hashlib.sha256(stream.tobytes()).hexdigest()))
    np.savez_compressed(os.path.join(HERE, "cases_f64.npz"), **arrays)
    with open(os.path.join(HERE, "cases_f64.json"), "w") as f:
        json.dump(index, f, indent=1)
    print("f64 cases", len(index))


# fastMode 2 cases: (dtype, kind, n, seed, mode, abs_eb, rel, B, T).  Chunk
# lengths that leave a unit of one value (n_t % B == 1) are included.
RA_CASES = [
    (0, "walk", 12, 1, 0, 1.0, 0.0, 4, 1),
    (0, "walk", 13, 2, 0, 1.0, 0.0, 4, 1),          # last unit of one value
    (0, "walk", 1000, 3, 0, 0.01, 0.0, 8, 1),
    (0, "walk", 4097, 4, 1, 0.0, 1e-4, 128, 8),
    (0, "walk", 20001, 7, 1, 0.0, 1e-3, 64, 3),
    (0, "waves", 16384, 8, 1, 0.0, 1e-3, 128, 8),
    (0, "waves", 16387, 9, 1, 0.0, 1e-4, 32, 4),
    (0, "uniform", 8193, 11, 1, 0.0, 1e-5, 64, 7),
    (0, "zeroblocks", 16385, 13, 1, 0.0, 1e-3, 32, 5),
    (0, "constant", 3000, 15, 1, 0.0, 1e-3, 32, 4),
    (0, "walk", 5000, 16, 0, 0.01, 0.0, 1, 4),       # units of one value
    (0, "walk", 5000, 17, 0, 0.01, 0.0, 2, 3),
    (0, "walk", 5000, 19, 0, 0.01, 0.0, 255, 2),
    (0, "walk", 200, 21, 0, 0.05, 0.0, 8, 16),
    (0, "specials", 6000, 23, 0, 0.5, 0.0, 32, 4),
    (0, "sparsewide", 6000, 25, 0, 1.0, 0.0, 64, 2),
    (0, "signedzero", 4000, 28, 1, 0.0, 1e-3, 32, 4),
    (0, "nanfirst", 2000, 30, 1, 0.0, 1e-3, 32, 2),
    (1, "walk", 4097, 4, 1, 0.0, 1e-6, 128, 8),
    (1, "waves", 16387, 9, 1, 0.0, 1e-4, 32, 4),
    (1, "walk", 5001, 17, 0, 1e-9, 0.0, 5, 3),
    (1, "specials", 6000, 23, 0, 0.5, 0.0, 32, 4),
    (1, "zeroblocks", 16385, 13, 1, 0.0, 1e-3, 32, 5),
    (1, "walk", 3000, 31, 0, 0.01, 0.0, 1, 4),        # units of one value (fit in 8N)
]


def make_ra():
    R, O, libc = load()
    arrays = {}
    index = []
    for k, (dt, kind, n, seed, mode, ab, rel, B, T) in enumerate(RA_CASES):
        x = make_input_f64(kind, n, seed) if dt else make_input(kind, n, seed)
        if not fits_reference_mode(O, x, 2, dt, mode, ab, rel, B, T):
            print("skip fastMode 2 (reference buffer too small):", dt, kind, n, B, T)
            continue
        stream, dec = ref_roundtrip(R, libc, x, mode, ab, rel, B, T, dtype=dt, fast=2)
        arrays["in_%d" % k] = x.view(np.uint64 if dt else np.uint32)
        arrays["stream_%d" % k] = stream
        arrays["out_%d" % k] = dec
        index.append(dict(id=k, dtype=dt, kind=kind, n=n, seed=seed, mode=mode, abs_eb=ab, rel=rel,
                          B=B, T=T, stream_bytes=int(stream.size),
                          stream_sha256=hashlib.sha256(stream.tobytes()).hexdigest()))
    np.savez_compressed(os.path.join(HERE, "cases_ra.npz"), **arrays)
    with open(os.path.join(HERE, "cases_ra.json"), "w") as f:
        json.dump(index, f, indent=1)
    print("fastMode 2 cases", len(index))


def main():
    R, O, libc = load()
    arrays = {}
    index = []
    for k, (kind, n, seed, mode, ab, rel, B, T) in enumerate(SMALL_CASES):
        x = make_input(kind, n, seed)
        if not fits_reference(O, x, mode, ab, rel, B, T):
            print("skip (reference buffer too small):", kind, n, B, T)
            continue
        stream, dec = ref_roundtrip(R, libc, x, mode, ab, rel, B, T)
        arrays["in_%d" % k] = x.view(np.uint32)
        arrays["stream_%d" % k] = stream
        arrays["out_%d" % k] = dec
        index.append(dict(id=k, kind=kind, n=n, seed=seed, mode=mode, abs_eb=ab, rel=rel, B=B, T=T,
                          stream_bytes=int(stream.size),
                          stream_sha256=hashlib.sha256(stream.tobytes()).hexdigest()))
    np.savez_compressed(os.path.join(HERE, "cases.npz"), **arrays)
    with open(os.path.join(HERE, "cases.json"), "w") as f:
        json.dump(index, f, indent=1)

    hashes = []
    for (kind, n, seed, mode, ab, rel, B, T) in HASH_CASES:
        x = make_input(kind, n, seed)
        if not fits_reference(O, x, mode, ab, rel, B, T):
            print("skip hash case:", kind, n, B, T)
            continue
        stream, dec = ref_roundtrip(R, libc, x, mode, ab, rel, B, T)
        hashes.append(dict(kind=kind, n=n, seed=seed, mode=mode, abs_eb=ab, rel=rel, B=B, T=T,
                           input_sha256=hashlib.sha256(x.tobytes()).hexdigest(),
                           stream_bytes=int(stream.size),
                           stream_sha256=hashlib.sha256(stream.tobytes()).hexdigest(),
                           out_sha256=hashlib.sha256(dec.tobytes()).hexdigest()))
        print("hash case", kind, n, B, T, stream.size)
    with open(os.path.join(HERE, "hashes.json"), "w") as f:
        json.dump(hashes, f, indent=1)

    # Reference bit widths (through its own log2f call) at every table edge.
    probes = []
    for k in range(1, 33):
        for m in sorted({(1 << (k - 1)), (1 << k) - 1, max(1, (1 << k) - 2)}):
            probes.append(m)
    bump = [2097151, 4194302, 8388603, 16777205, 33554410, 67108819, 134217637, 268435273,
            536870545, 1073741089, 2147482177, 4294964353]
    for m in bump:
        probes += [m - 1, m]
    probes = sorted(set(p for p in probes if 0 < p < 2 ** 32))
    widths = {str(m): int(O.szpo_bit_width(m)) for m in probes}
    with open(os.path.join(HERE, "bitwidth.json"), "w") as f:
        json.dump(dict(source="glibc log2f via SZp/src/szp_float.cc:261", widths=widths), f,
                  indent=0)
    print("cases", len(index), "hash cases", len(hashes))


if __name__ == "__main__":
    if sys.argv[1:] == ["f64"]:
        make_f64()
    elif sys.argv[1:] == ["ra"]:
        make_ra()
    else:
        main()
        make_f64()
        make_ra()
