"""CPU checks of the product library's C ABI (no kernel launches): it loads,
exports every entry point include/*.h declares, and its host-side helpers
agree with the oracle and with the format."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import oracle_lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "szp_amd", "libszp_amd.so")


def declared_functions():
    names = set()
    for h in ("szp.h", "szp_amd.h", "szp_rw.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"([A-Za-z_][A-Za-z0-9_]*)\s*\(", text):
            name = m.group(1)
            if name.startswith("szp_") and not name.endswith("_H"):
                names.add(name)
    return names


@pytest.fixture(scope="module")
def lib():
    assert os.path.exists(LIB), "build first: python szp_amd/build.py"
    return ctypes.CDLL(LIB)


def test_exports_every_declared_symbol(lib):
    names = declared_functions()
    assert len(names) >= 20
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = sorted(names - exported)
    assert not missing, missing
    for data in ("szp_versionNumber", "szp_dataEndianType", "szp_sysEndianType"):
        assert data in exported
    for n in names:
        getattr(lib, n)


def test_package_imports():
    import szp_amd
    assert szp_amd.lib is not None
    assert szp_amd.SZP_NONRANDOMACCESS == 1 and szp_amd.SZ_FLOAT == 0
    assert szp_amd.ABS == 0 and szp_amd.REL == 1


def test_bit_width_matches_oracle():
    import szp_amd
    O = oracle_lib.oracle()
    rng = np.random.default_rng(0)
    probes = [0, 1, 2, 3, 255, 256, 2097150, 2097151, 4294964352, 4294964353, 0x80000000,
              0xFFFFFFFF] + list(rng.integers(0, 1 << 32, 20000, dtype=np.uint64))
    for m in probes:
        assert szp_amd.bit_width(int(m)) == O.szpo_bit_width(int(m)), m


def test_compress_bound_matches_oracle():
    import szp_amd
    O = oracle_lib.oracle()
    for n in (0, 1, 2, 7, 100, 4097, 1 << 20, (1 << 27) + 5):
        for B in (1, 8, 32, 128, 1000, 2048):
            for T in (1, 3, 16):
                assert szp_amd.compress_bound(n, B, T) == O.szpo_compress_bound(n, B, T)
    assert szp_amd.compress_bound(100, 2049, 1) == 0  # block sizes above 2048 are rejected


def test_stream_length_from_header(golden):
    import szp_amd
    cases, arrays, _ = golden
    for c in cases:
        s = arrays["stream_%d" % c["id"]]
        assert szp_amd.stream_length(s, c["n"], c["B"], c["T"]) == s.size


def test_threads_override():
    import szp_amd
    szp_amd.set_threads(5)
    assert szp_amd.default_threads() == 5
    szp_amd.set_threads(0)
    assert szp_amd.default_threads() >= 1


def test_unsupported_modes_return_null_without_a_device(lib):
    import szp_amd
    x = np.zeros(16, np.float32)
    with pytest.raises(szp_amd.SzpError):
        szp_amd.szp_compress(3, 0, x, 0, 0.1, 0.0, 16, 8)  # no such fastMode
    assert b"supported" in szp_amd.lib.szp_amd_last_error()
    size = ctypes.c_size_t()
    L = szp_amd.lib
    assert not L.szp_compress(1, 2, x.ctypes.data, ctypes.byref(size), 0, 0.1, 0.0, 16, 8)
    assert b"supported" in L.szp_amd_last_error()
    assert not L.szp_decompress(1, 7, x.ctypes.data, 64, 16, 8)
    with pytest.raises(szp_amd.SzpError):
        szp_amd.szp_compress(1, 2, x, 0, 0.1, 0.0, 16, 8)  # the mirror rejects it too


def test_ra_bound_and_stream_length():
    import json
    import szp_amd
    O = oracle_lib.oracle()
    for n in (0, 1, 100, 4097, 1 << 20):
        for B in (1, 2, 32, 128):
            for T in (1, 16):
                for dt in (0, 1):
                    assert (szp_amd.compress_bound(n, B, T, dt, 2)
                            == O.szpo_compress_bound_mode(2, dt, n, B, T))
    d = os.path.join(ROOT, "tests", "golden")
    cases = json.load(open(os.path.join(d, "cases_ra.json")))
    arrays = np.load(os.path.join(d, "cases_ra.npz"))
    for c in cases:
        s = arrays["stream_%d" % c["id"]]
        assert szp_amd.stream_length(s, c["n"], c["B"], c["T"], c["dtype"], 2) == s.size


def test_f64_bound_and_stream_length():
    import json
    import szp_amd
    O = oracle_lib.oracle()
    for n in (0, 1, 100, 4097, 1 << 20):
        for B in (1, 32, 128):
            for T in (1, 16):
                assert (szp_amd.compress_bound(n, B, T, szp_amd.SZ_DOUBLE)
                        == O.szpo_compress_bound(n, B, T) + 4)
    d = os.path.join(ROOT, "tests", "golden")
    cases = json.load(open(os.path.join(d, "cases_f64.json")))
    arrays = np.load(os.path.join(d, "cases_f64.npz"))
    for c in cases:
        s = arrays["stream_%d" % c["id"]]
        assert szp_amd.stream_length(s, c["n"], c["B"], c["T"], szp_amd.SZ_DOUBLE) == s.size


def test_sizes_are_consistent():
    import szp_amd
    L = szp_amd.lib
    for n, B, T in ((1000, 32, 2), (1 << 27, 128, 16), (5, 8, 16)):
        assert L.szp_amd_compress_workspace_size(n, B, T) > 0
        assert L.szp_amd_decompress_workspace_size(szp_amd.compress_bound(n, B, T), n, B, T) > 0
        assert L.szp_amd_index_size(n, B, T) >= 8 * ((n + B - 1) // B)


def test_raw_file_helpers(tmp_path):
    """szp_rw.h: byte files round-trip; float/double files are read in the
    byte order szp_dataEndianType names (swapped when it differs from the
    host's, SZp/src/szp_rw.cc:424-459)."""
    L = ctypes.CDLL(LIB)
    L.szp_readByteData.restype = ctypes.c_void_p
    L.szp_readFloatData.restype = ctypes.c_void_p
    L.szp_readDoubleData.restype = ctypes.c_void_p
    L.szp_checkFileSize.restype = ctypes.c_size_t
    libc = ctypes.CDLL("libc.so.6")
    libc.free.argtypes = [ctypes.c_void_p]
    st = ctypes.c_int(7)
    n = ctypes.c_size_t(0)
    data = np.arange(1000, dtype=np.uint8)
    f = str(tmp_path / "b.bin").encode()
    L.szp_writeByteData(data.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(data.size), f,
                        ctypes.byref(st))
    assert st.value == 0
    assert L.szp_checkFileSize(f, ctypes.byref(st)) == 1000 and st.value == 0
    p = L.szp_readByteData(f, ctypes.byref(n), ctypes.byref(st))
    assert st.value == 0 and n.value == 1000
    assert bytes((ctypes.c_uint8 * 1000).from_address(p)) == data.tobytes()
    libc.free(p)
    x = np.linspace(-1, 1, 257).astype(np.float32)
    g = str(tmp_path / "x.f32").encode()
    L.szp_writeFloatData_inBytes(x.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(x.size), g,
                                 ctypes.byref(st))
    assert st.value == 0 and open(g, "rb").read() == x.tobytes()
    p = L.szp_readFloatData(g, ctypes.byref(n), ctypes.byref(st))
    assert n.value == 257 and np.array_equal(np.ctypeslib.as_array((ctypes.c_float * 257).from_address(p)), x)
    libc.free(p)
    y = x.astype(np.float64)
    h = str(tmp_path / "y.f64").encode()
    L.szp_writeDoubleData_inBytes(y.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(y.size), h,
                                  ctypes.byref(st))
    endian = ctypes.c_int.in_dll(L, "szp_dataEndianType")
    endian.value = 1  # big-endian data on a little-endian host: values are swapped
    try:
        p = L.szp_readDoubleData(h, ctypes.byref(n), ctypes.byref(st))
        got = np.ctypeslib.as_array((ctypes.c_double * 257).from_address(p)).copy()
        libc.free(p)
    finally:
        endian.value = 0
    assert np.array_equal(got.view(np.uint64), y.byteswap().view(np.uint64))
    assert L.szp_readByteData(str(tmp_path / "missing").encode(), ctypes.byref(n),
                              ctypes.byref(st)) is None and st.value == -2


def test_damaged_header_is_refused_before_any_read(golden):
    """szp_amd_stream_length (and so szp_decompress) refuses a stream whose
    last chunk offset points past the longest body a valid stream can have,
    instead of walking the host buffer from there; a valid stream keeps its
    length.  The stream is embedded in a guard buffer whose tail is never
    reached by a refused walk."""
    import szp_amd
    cases, arrays, _ = golden
    L = szp_amd.lib
    checked = 0
    for c in cases:
        s = arrays["stream_%d" % c["id"]]
        n, B, T = c["n"], c["B"], c["T"]
        if n == 0:
            continue
        hdr_last = 4 + 8 * (T - 1)
        bound = szp_amd.compress_bound(n, B, T)
        for bad in (bound, bound - 4, 1 << 40, (1 << 64) - 1):
            d = s.copy()
            d[hdr_last:hdr_last + 8] = np.frombuffer(np.uint64(bad).tobytes(), np.uint8)
            assert L.szp_amd_stream_length(d.ctypes.data, n, B, T) == 0, (c["id"], bad)
        assert L.szp_amd_stream_length(s.ctypes.data, n, B, T) == s.size
        checked += 1
    assert checked >= 20
