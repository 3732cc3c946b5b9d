"""ctypes access to the CPU oracle (oracle/libszp_oracle.so) and, when built,
the reference compiled from its own sources (oracle/_ref/libszp_ref.so).

Test infrastructure only: the oracle is the checker, never the thing tested.
"""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "libszp_oracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libszp_ref.so")

_O = None
_R = None


def oracle():
    global _O
    if _O is None:
        if not os.path.exists(ORACLE_SO):
            subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)
        O = ctypes.CDLL(ORACLE_SO)
        u64, i32, vp, f32, dbl = (ctypes.c_uint64, ctypes.c_int, ctypes.c_void_p, ctypes.c_float,
                                  ctypes.c_double)
        O.szpo_bit_width.argtypes = [ctypes.c_uint32]
        O.szpo_bit_width_table.argtypes = [ctypes.c_uint32]
        O.szpo_check_bit_width_range.restype = u64
        O.szpo_check_bit_width_range.argtypes = [u64, u64, ctypes.POINTER(ctypes.c_uint32)]
        O.szpo_inverse_bound.restype = f32
        O.szpo_inverse_bound.argtypes = [f32]
        O.szpo_quantize.restype = ctypes.c_int32
        O.szpo_quantize.argtypes = [f32, f32]
        O.szpo_block_bytes.restype = u64
        O.szpo_block_bytes.argtypes = [i32, u64]
        O.szpo_encode_block.restype = u64
        O.szpo_encode_block.argtypes = [vp, u64, vp]
        O.szpo_decode_block.restype = u64
        O.szpo_decode_block.argtypes = [vp, u64, vp]
        O.szpo_rel_bound.restype = f32
        O.szpo_rel_bound.argtypes = [vp, u64, dbl]
        O.szpo_compress_bound.restype = u64
        O.szpo_compress_bound.argtypes = [u64, i32, i32]
        O.szpo_compress.argtypes = [vp, u64, f32, i32, i32, i32, vp, u64, ctypes.POINTER(u64)]
        O.szpo_decompress.argtypes = [vp, u64, u64, i32, i32, i32, vp]
        O.szpo_compress_api.argtypes = [vp, u64, i32, dbl, dbl, i32, i32, i32, vp, u64,
                                        ctypes.POINTER(u64)]
        O.szpo_quantize_f64.restype = ctypes.c_int32
        O.szpo_quantize_f64.argtypes = [dbl, dbl]
        O.szpo_rel_bound_f64.restype = dbl
        O.szpo_rel_bound_f64.argtypes = [vp, u64, dbl]
        O.szpo_decompress_f64.argtypes = [vp, u64, u64, i32, i32, i32, vp]
        O.szpo_compress_api_f64.argtypes = [vp, u64, i32, dbl, dbl, i32, i32, i32, vp, u64,
                                            ctypes.POINTER(u64)]
        O.szpo_compress_mode.argtypes = [i32, i32, vp, u64, i32, dbl, dbl, i32, i32, i32, vp, u64,
                                         ctypes.POINTER(u64)]
        O.szpo_decompress_mode.argtypes = [i32, i32, vp, u64, u64, i32, i32, i32, vp]
        O.szpo_compress_bound_mode.restype = u64
        O.szpo_compress_bound_mode.argtypes = [i32, i32, u64, i32, i32]
        O.szpo_record_count.restype = u64
        O.szpo_record_count.argtypes = [vp, u64, i32, i32]
        _O = O
    return _O


def compress(x, mode, abs_eb, rel, B, T, workers=0, dtype=np.float32):
    """Oracle stream (uint8 array) for szp_compress(1, SZ_FLOAT or SZ_DOUBLE,
    x, mode, ...) with team size T."""
    O = oracle()
    x = np.ascontiguousarray(x, dtype=dtype)
    f64 = x.dtype == np.float64
    cap = O.szpo_compress_bound(x.size, B, T) + (4 if f64 else 0)
    out = np.empty(cap, np.uint8)
    sz = ctypes.c_uint64()
    fn = O.szpo_compress_api_f64 if f64 else O.szpo_compress_api
    rc = fn(x.ctypes.data, x.size, mode, abs_eb, rel, B, T, workers, out.ctypes.data, cap,
            ctypes.byref(sz))
    if rc:
        raise ValueError("oracle compress rejected the arguments")
    return out[:sz.value].copy()


def decompress(stream, n, B, T, workers=0, dtype=np.float32):
    O = oracle()
    s = np.ascontiguousarray(stream, dtype=np.uint8)
    out = np.empty(n, dtype)
    fn = O.szpo_decompress_f64 if out.dtype == np.float64 else O.szpo_decompress
    rc = fn(s.ctypes.data, s.size, n, B, T, workers, out.ctypes.data)
    if rc:
        raise ValueError("oracle decompress found a malformed stream")
    return out


def compress_mode(fast, x, mode, abs_eb, rel, B, T, workers=0):
    """Oracle stream for szp_compress(fast, SZ_FLOAT or SZ_DOUBLE by x's
    dtype, x, mode, ...) with team size T; fast = 1 or 2."""
    O = oracle()
    x = np.ascontiguousarray(x)
    assert x.dtype in (np.float32, np.float64)
    dt = 1 if x.dtype == np.float64 else 0
    cap = O.szpo_compress_bound_mode(fast, dt, x.size, B, T)
    out = np.empty(cap, np.uint8)
    sz = ctypes.c_uint64()
    rc = O.szpo_compress_mode(fast, dt, x.ctypes.data, x.size, mode, abs_eb, rel, B, T, workers,
                              out.ctypes.data, cap, ctypes.byref(sz))
    if rc:
        raise ValueError("oracle compress rejected the arguments")
    return out[:sz.value].copy()


def decompress_mode(fast, stream, n, B, T, dtype=np.float32, workers=0):
    O = oracle()
    s = np.ascontiguousarray(stream, dtype=np.uint8)
    out = np.empty(n, dtype)
    dt = 1 if out.dtype == np.float64 else 0
    rc = O.szpo_decompress_mode(fast, dt, s.ctypes.data, s.size, n, B, T, workers,
                                out.ctypes.data)
    if rc:
        raise ValueError("oracle decompress found a malformed stream")
    return out


def reference():
    """The reference library compiled from its sources, or None when not built."""
    global _R
    if _R is None and os.path.exists(REF_SO):
        R = ctypes.CDLL(REF_SO)
        R.szp_compress.restype = ctypes.c_void_p
        R.szp_compress.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                   ctypes.POINTER(ctypes.c_size_t), ctypes.c_int, ctypes.c_double,
                                   ctypes.c_double, ctypes.c_size_t, ctypes.c_int]
        R.szp_decompress.restype = ctypes.c_void_p
        R.szp_decompress.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t,
                                     ctypes.c_size_t, ctypes.c_int]
        R.szp_ref_set_threads.argtypes = [ctypes.c_int]
        szt, vp, f32, i32 = ctypes.c_size_t, ctypes.c_void_p, ctypes.c_float, ctypes.c_int
        R.szp_float_single_thread_arg_record.restype = szt
        R.szp_float_single_thread_arg_record.argtypes = [vp, vp, ctypes.POINTER(szt), f32, szt,
                                                         i32]
        R.szp_float_decompress_single_thread_arg_record.restype = szt
        R.szp_float_decompress_single_thread_arg_record.argtypes = [vp, szt, f32, i32, vp]
        _R = R
    return _R


_libc = ctypes.CDLL("libc.so.6")
_libc.free.argtypes = [ctypes.c_void_p]


def ref_fits(stream, n, T, es=4):
    """True when the reference's fixed buffers (es*N+es total, ceil(es*N/T)
    per chunk, SZp/src/szp_float.cc:156,186, SZp/src/szp_double.cc:157,186)
    hold this stream; es = bytes per value."""
    hdr = es + 8 * T
    offs = np.frombuffer(stream[es:hdr].tobytes(), np.uint64).astype(np.int64)
    sizes = np.diff(np.append(offs, stream.size - hdr))
    return stream.size <= es * n + es and sizes.max() <= (es * n + T - 1) // T


def ref_compress(x, mode, abs_eb, rel, B, T, dtype=np.float32, fast=1):
    R = reference()
    R.szp_ref_set_threads(T)
    x = np.ascontiguousarray(x, dtype=dtype)
    sz = ctypes.c_size_t()
    p = R.szp_compress(fast, 1 if x.dtype == np.float64 else 0, x.ctypes.data, ctypes.byref(sz),
                       mode, abs_eb, rel, x.size, B)
    out = np.ctypeslib.as_array((ctypes.c_uint8 * sz.value).from_address(p)).copy()
    _libc.free(p)
    return out


def ref_decompress(stream, n, B, T, dtype=np.float32, fast=1):
    R = reference()
    R.szp_ref_set_threads(T)
    s = np.ascontiguousarray(stream, dtype=np.uint8)
    f64 = np.dtype(dtype) == np.float64
    p = R.szp_decompress(fast, 1 if f64 else 0, s.ctypes.data, s.size, n, B)
    ct = ctypes.c_double if f64 else ctypes.c_float
    out = np.ctypeslib.as_array((ct * n).from_address(p)).copy()
    _libc.free(p)
    return out


def record_count(stream, n, B, decompress):
    """The oracle's memory-access tally of the reference's instrumented
    single-thread forms for a T = 1 float stream (the whole stream, eb first)."""
    s = np.ascontiguousarray(stream[4:], dtype=np.uint8)
    return int(oracle().szpo_record_count(s.ctypes.data, n, B, 1 if decompress else 0))


def ref_single_thread_record(x, eb, B):
    """(stream, compress count, decompressed values, decompress count) from the
    reference's szp_float_single_thread_arg_record and
    szp_float_decompress_single_thread_arg_record (x must compress to at most
    4 * x.size bytes: the reference's per-thread buffer,
    SZp/src/szp_float.cc:632-633)."""
    R = reference()
    x = np.ascontiguousarray(x, dtype=np.float32)
    out = np.zeros(4 * x.size + 64, np.uint8)
    sz = ctypes.c_size_t()
    c1 = R.szp_float_single_thread_arg_record(out.ctypes.data, x.ctypes.data, ctypes.byref(sz),
                                              eb, x.size, B)
    stream = out[:sz.value].copy()
    dec = np.zeros(x.size, np.float32)
    body = np.ascontiguousarray(stream[4:])
    c2 = R.szp_float_decompress_single_thread_arg_record(dec.ctypes.data, x.size, eb, B,
                                                         body.ctypes.data)
    return stream, int(c1), dec, int(c2)
