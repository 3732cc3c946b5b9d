"""GPU parity of the float fastMode-1 lower-level entry points
(SZp/include/szp_float.h:30-40, SZp/include/szpd_float.h:21-27): the
OpenMP-team forms with and without a caller buffer, and the single-thread
forms with their memory-access tallies, all through the library's C ABI.

Bar: streams byte-identical to the oracle's (and, where the reference is
built, to the reference's own single-thread functions), values bit-identical,
record counts equal.
"""
import ctypes
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


def _bytes_at(szp, p, n):
    got = np.ctypeslib.as_array((ctypes.c_uint8 * max(n, 1)).from_address(p))[:n].copy()
    szp._libc.free(p)
    return got


@pytest.mark.parametrize("kind,n,B,T,eb", [
    ("waves", 100001, 128, 8, 1e-4),
    ("walk", 70000, 32, 3, 0.01),
    ("zeroblocks", 65537, 9, 16, 1e-3),
    ("walk", 5, 8, 16, 0.01),      # fewer elements than chunks
])
def test_openmp_threadblock_forms(szp, kind, n, B, T, eb):
    """szp_float_openmp_threadblock(_arg) (SZp/src/szp_float.cc:148-480) and
    szp_float_decompress_openmp_threadblock(_arg) (SZp/src/szpd_float.cc:23-128,
    320-423) with the team size T from szp_amd_set_threads."""
    x = make_input(kind, n, 31 + n)
    say("threadblock forms", kind, n, B, T, eb)
    eb32 = float(np.float32(eb))
    want = oracle_lib.compress(x, 0, eb32, 0.0, B, T)
    L = szp.lib
    szp.set_threads(T)
    try:
        size = ctypes.c_size_t()
        p = L.szp_float_openmp_threadblock(x.ctypes.data, ctypes.byref(size), eb32, n, B)
        assert p, szp._last_error()
        got = _bytes_at(szp, p, size.value)
        assert got.size == want.size and np.array_equal(got, want)

        buf = np.zeros(szp.compress_bound(n, B, T), np.uint8)
        size2 = ctypes.c_size_t()
        L.szp_float_openmp_threadblock_arg(buf.ctypes.data, x.ctypes.data, ctypes.byref(size2),
                                           eb32, n, B)
        assert size2.value == want.size and np.array_equal(buf[:size2.value], want)

        ref = oracle_lib.decompress(want, n, B, T)
        body = np.ascontiguousarray(want[4:])
        q = L.szp_float_decompress_openmp_threadblock(n, eb32, B, body.ctypes.data)
        assert q, szp._last_error()
        d = np.ctypeslib.as_array((ctypes.c_float * n).from_address(q)).copy()
        szp._libc.free(q)
        assert np.array_equal(d.view(np.uint32), ref.view(np.uint32))
        d2 = np.zeros(n, np.float32)
        L.szp_float_decompress_openmp_threadblock_arg(d2.ctypes.data, n, eb32, B,
                                                      body.ctypes.data)
        assert np.array_equal(d2.view(np.uint32), ref.view(np.uint32))
    finally:
        szp.set_threads(0)


@pytest.mark.parametrize("kind,n,B,eb", [
    ("waves", 200003, 128, 1e-4),
    ("walk", 30000, 32, 0.01),
    ("zeroblocks", 40000, 8, 1e-3),
    ("walk", 1, 16, 0.01),
    ("walk", 0, 16, 0.01),
])
def test_single_thread_forms(szp, kind, n, B, eb):
    """szp_float_single_thread_arg(_record) and
    szp_float_decompress_single_thread_arg(_record) (SZp/src/szp_float.cc:482-750,
    SZp/src/szpd_float.cc:130-317): the T = 1 stream whatever the default team
    size, and the reference's memory-access tallies."""
    x = make_input(kind, n, 41 + n) if n else np.zeros(0, np.float32)
    say("single-thread forms", kind, n, B, eb)
    eb32 = float(np.float32(eb))
    want = oracle_lib.compress(x, 0, eb32, 0.0, B, 1)
    L = szp.lib
    szp.set_threads(7)  # ignored by the single-thread forms
    try:
        buf = np.zeros(szp.compress_bound(n, B, 1), np.uint8)
        size = ctypes.c_size_t()
        L.szp_float_single_thread_arg(buf.ctypes.data, x.ctypes.data, ctypes.byref(size), eb32,
                                      n, B)
        assert size.value == want.size and np.array_equal(buf[:size.value], want)
        buf2 = np.zeros_like(buf)
        size2 = ctypes.c_size_t()
        c1 = L.szp_float_single_thread_arg_record(buf2.ctypes.data, x.ctypes.data,
                                                  ctypes.byref(size2), eb32, n, B)
        assert size2.value == want.size and np.array_equal(buf2[:size2.value], want)
        assert c1 == oracle_lib.record_count(want, n, B, False)

        ref = oracle_lib.decompress(want, n, B, 1)
        body = np.ascontiguousarray(want[4:])
        d = np.zeros(max(n, 1), np.float32)
        L.szp_float_decompress_single_thread_arg(d.ctypes.data, n, eb32, B, body.ctypes.data)
        assert np.array_equal(d[:n].view(np.uint32), ref.view(np.uint32))
        d2 = np.zeros(max(n, 1), np.float32)
        c2 = L.szp_float_decompress_single_thread_arg_record(d2.ctypes.data, n, eb32, B,
                                                             body.ctypes.data)
        assert np.array_equal(d2[:n].view(np.uint32), ref.view(np.uint32))
        assert c2 == oracle_lib.record_count(want, n, B, True)
    finally:
        szp.set_threads(0)
    if oracle_lib.reference() is not None and n:
        stream, r1, rdec, r2 = oracle_lib.ref_single_thread_record(x, np.float32(eb32), B)
        assert np.array_equal(stream, want)
        assert (r1, r2) == (c1, c2)
        assert np.array_equal(rdec.view(np.uint32), ref.view(np.uint32))
