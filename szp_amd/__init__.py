"""szp_amd -- SZp compress/decompress (fastMode 1 and 2, float32 and float64) on MI355X.

Python mirror of the reference's C entry points (SZp/include/szp.h:33-36),
driven through the C ABI of the in-tree library ``libszp_amd.so`` (HIP kernels
for gfx950), plus a device-resident form on torch tensors.

    stream, size = szp_compress(SZP_NONRANDOMACCESS, SZ_FLOAT, data, ABS, 1e-3, 0.0,
                                data.size, 128)
    values = szp_decompress(SZP_NONRANDOMACCESS, SZ_FLOAT, stream, size, data.size, 128)

There is no CPU fallback: importing this package fails when the library is
missing, and every call runs the HIP kernels.
"""
import ctypes
import os

import numpy as np

__all__ = [
    "ABS", "REL", "SZP_NONRANDOMACCESS", "SZP_RANDOMACCESS", "SZ_FLOAT", "SZ_DOUBLE",
    "SzpError", "lib", "library_path", "szp_compress", "szp_decompress",
    "compress_bound", "default_threads", "set_threads", "bit_width", "stream_length",
    "DeviceCodec", "profile_enable", "profile_read",
]

# SZp/include/szp_defines.h:19-20,33-37
ABS = 0
REL = 1
SZP_NONRANDOMACCESS = 1
SZP_RANDOMACCESS = 2
SZ_FLOAT = 0
SZ_DOUBLE = 1

STATUS = {
    0: "ok", 1: "invalid argument", 2: "capacity too small", 3: "HIP runtime error",
    4: "malformed compressed stream", 5: "device wait timed out", 6: "unsupported mode",
    7: "allocation failed",
}

# SZP_AMD_LIBRARY names another build of the same library (a side-by-side
# comparison of two builds in one process tree); by default the in-tree one.
library_path = os.environ.get("SZP_AMD_LIBRARY") or os.path.join(
    os.path.dirname(os.path.abspath(__file__)), "libszp_amd.so")


class SzpError(RuntimeError):
    def __init__(self, what, code=None):
        msg = what if code is None else "%s: %s (%d)" % (what, STATUS.get(code, "?"), code)
        super().__init__(msg)
        self.code = code


def _load():
    if not os.path.exists(library_path):
        raise ImportError(
            "libszp_amd.so is not built (%s); run `python szp_amd/build.py` or "
            "__graft_entry__.build()" % library_path)
    L = ctypes.CDLL(library_path)
    u64, i32, vp, dbl, f32 = (ctypes.c_uint64, ctypes.c_int, ctypes.c_void_p, ctypes.c_double,
                              ctypes.c_float)
    szt = ctypes.c_size_t
    sig = {
        "szp_compress": (vp, [i32, i32, vp, ctypes.POINTER(szt), i32, dbl, dbl, szt, i32]),
        "szp_decompress": (vp, [i32, i32, vp, szt, szt, i32]),
        "szp_float_openmp_threadblock": (vp, [vp, ctypes.POINTER(szt), f32, szt, i32]),
        "szp_float_openmp_threadblock_arg": (None, [vp, vp, ctypes.POINTER(szt), f32, szt, i32]),
        "szp_float_decompress_openmp_threadblock": (vp, [szt, f32, i32, vp]),
        "szp_float_decompress_openmp_threadblock_arg": (None, [vp, szt, f32, i32, vp]),
        "szp_float_single_thread_arg": (None, [vp, vp, ctypes.POINTER(szt), f32, szt, i32]),
        "szp_float_single_thread_arg_record": (szt, [vp, vp, ctypes.POINTER(szt), f32, szt, i32]),
        "szp_float_decompress_single_thread_arg": (None, [vp, szt, f32, i32, vp]),
        "szp_float_decompress_single_thread_arg_record": (szt, [vp, szt, f32, i32, vp]),
        "szp_amd_compress_bound": (u64, [u64, i32, i32]),
        "szp_amd_compress_workspace_size": (u64, [u64, i32, i32]),
        "szp_amd_decompress_workspace_size": (u64, [u64, u64, i32, i32]),
        "szp_amd_compress_dev": (i32, [vp, u64, i32, dbl, dbl, i32, i32, vp, u64, vp, vp, u64, vp]),
        "szp_amd_decompress_dev": (i32, [vp, u64, u64, i32, i32, vp, vp, u64, vp]),
        "szp_amd_check_status": (i32, [vp, vp]),
        "szp_amd_stream_length": (u64, [vp, u64, i32, i32]),
        "szp_amd_default_threads": (i32, []),
        "szp_amd_set_threads": (None, [i32]),
        "szp_amd_bit_width": (i32, [ctypes.c_uint32]),
        "szp_amd_status_string": (ctypes.c_char_p, [i32]),
        "szp_amd_last_error": (ctypes.c_char_p, []),
        "szp_amd_index_size": (u64, [u64, i32, i32]),
        "szp_amd_compress_dev_indexed": (i32, [vp, u64, i32, dbl, dbl, i32, i32, vp, u64, vp, vp,
                                               u64, vp, u64, vp]),
        "szp_amd_decompress_dev_indexed": (i32, [vp, u64, vp, u64, u64, i32, i32, vp, vp, u64, vp]),
        "szp_double_openmp_threadblock": (vp, [vp, ctypes.POINTER(szt), dbl, szt, i32]),
        "szp_double_openmp_threadblock_arg": (None, [vp, vp, ctypes.POINTER(szt), dbl, szt, i32]),
        "szp_double_decompress_openmp_threadblock": (vp, [szt, dbl, i32, vp]),
        "szp_double_decompress_openmp_threadblock_arg": (None, [vp, szt, dbl, i32, vp]),
        "szp_float_openmp_threadblock_randomaccess": (vp, [vp, ctypes.POINTER(szt), f32, szt, i32]),
        "szp_float_openmp_threadblock_randomaccess_arg": (None, [vp, vp, ctypes.POINTER(szt), f32,
                                                                 szt, i32]),
        "szp_float_decompress_openmp_threadblock_randomaccess": (vp, [szt, f32, i32, vp]),
        "szp_float_decompress_openmp_threadblock_randomaccess_arg": (None, [vp, szt, f32, i32, vp]),
        "szp_double_openmp_threadblock_randomaccess": (vp, [vp, ctypes.POINTER(szt), dbl, szt,
                                                            i32]),
        "szp_double_openmp_threadblock_randomaccess_arg": (None, [vp, vp, ctypes.POINTER(szt),
                                                                  dbl, szt, i32]),
        "szp_double_decompress_openmp_threadblock_randomaccess": (vp, [szt, dbl, i32, vp]),
        "szp_double_decompress_openmp_threadblock_randomaccess_arg": (None, [vp, szt, dbl, i32,
                                                                             vp]),
        "szp_amd_compress_bound_typed": (u64, [i32, i32, u64, i32, i32]),
        "szp_amd_compress_workspace_size_typed": (u64, [i32, u64, i32, i32]),
        "szp_amd_decompress_workspace_size_typed": (u64, [i32, u64, u64, i32, i32]),
        "szp_amd_index_size_typed": (u64, [i32, u64, i32, i32]),
        "szp_amd_compress_dev_typed": (i32, [i32, i32, vp, u64, i32, dbl, dbl, i32, i32, vp, u64,
                                             vp, vp, u64, vp, u64, vp]),
        "szp_amd_decompress_dev_typed": (i32, [i32, i32, vp, u64, vp, u64, u64, i32, i32, vp, vp,
                                               u64, vp]),
        "szp_amd_stream_length_typed": (u64, [i32, i32, vp, u64, i32, i32]),
        "szp_amd_profile_enable": (None, [i32]),
        "szp_amd_profile_enable_only": (None, [ctypes.c_char_p]),
        "szp_amd_profile_count": (i32, []),
        "szp_amd_profile_entry": (i32, [i32, ctypes.c_char_p, i32, ctypes.POINTER(u64),
                                        ctypes.POINTER(dbl)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    return L


lib = _load()
_libc = ctypes.CDLL("libc.so.6")
_libc.free.argtypes = [ctypes.c_void_p]
_libc.free.restype = None


def _last_error():
    return (lib.szp_amd_last_error() or b"").decode()


def profile_enable(on=True, only=None):
    """Starts (or stops) per-kernel event timing; both clear the totals.
    only: time just the launches of this kernel name."""
    if on and only:
        lib.szp_amd_profile_enable_only(only.encode())
    else:
        lib.szp_amd_profile_enable(1 if on else 0)


def profile_read():
    """{kernel name: (launches, total ms)} since profile_enable(True)."""
    out = {}
    for i in range(lib.szp_amd_profile_count()):
        name = ctypes.create_string_buffer(64)
        n = ctypes.c_uint64()
        ms = ctypes.c_double()
        if lib.szp_amd_profile_entry(i, name, 64, ctypes.byref(n), ctypes.byref(ms)) == 0:
            out[name.value.decode()] = (n.value, ms.value)
    return out


def default_threads():
    """T used by szp_compress/szp_decompress (the reference's OpenMP team size)."""
    return lib.szp_amd_default_threads()


def set_threads(n):
    """Pins T for the drop-in entry points; 0 restores the automatic choice."""
    lib.szp_amd_set_threads(int(n))


def compress_bound(n, block_size, threads, dataType=SZ_FLOAT, fastMode=SZP_NONRANDOMACCESS):
    return int(lib.szp_amd_compress_bound_typed(int(fastMode), int(dataType), int(n),
                                                int(block_size), int(threads)))


def bit_width(m):
    return int(lib.szp_amd_bit_width(int(m)))


def stream_length(stream, n, block_size, threads, dataType=SZ_FLOAT,
                  fastMode=SZP_NONRANDOMACCESS):
    s = np.ascontiguousarray(stream, dtype=np.uint8)
    return int(lib.szp_amd_stream_length_typed(int(fastMode), int(dataType), s.ctypes.data,
                                               int(n), int(block_size), int(threads)))


def _np_type(dataType):
    if dataType == SZ_FLOAT:
        return np.float32
    if dataType == SZ_DOUBLE:
        return np.float64
    raise SzpError("unknown dataType %r" % (dataType,))


def szp_compress(fastMode, dataType, data, errBoundMode, absErrBound, relBoundRatio, nbEle,
                 blockSize):
    """Mirror of szp_compress (SZp/include/szp.h:33-34): returns (stream, outSize).

    The stream is a numpy uint8 array copied out of the library's malloc'd
    buffer (which is then freed, as the reference's callers do)."""
    x = np.ascontiguousarray(data, dtype=_np_type(dataType))
    if x.size < nbEle:
        raise ValueError("data holds %d elements, nbEle=%d" % (x.size, nbEle))
    size = ctypes.c_size_t(0)
    p = lib.szp_compress(int(fastMode), int(dataType), x.ctypes.data, ctypes.byref(size),
                         int(errBoundMode), float(absErrBound), float(relBoundRatio), int(nbEle),
                         int(blockSize))
    if not p:
        raise SzpError("szp_compress failed: " + _last_error())
    try:
        out = np.ctypeslib.as_array((ctypes.c_uint8 * size.value).from_address(p)).copy()
    finally:
        _libc.free(p)
    return out, size.value


def szp_decompress(fastMode, dataType, bytes_, byteLength, nbEle, blockSize):
    """Mirror of szp_decompress (SZp/include/szp.h:36): returns float32[nbEle]
    (SZ_FLOAT) or float64[nbEle] (SZ_DOUBLE)."""
    s = np.ascontiguousarray(np.frombuffer(bytes_, dtype=np.uint8) if isinstance(bytes_, (bytes, bytearray)) else bytes_,
                             dtype=np.uint8)
    p = lib.szp_decompress(int(fastMode), int(dataType), s.ctypes.data, int(byteLength),
                           int(nbEle), int(blockSize))
    if not p:
        raise SzpError("szp_decompress failed: " + _last_error())
    try:
        n = int(nbEle)
        vt = _np_type(dataType)
        ct = ctypes.c_double if vt == np.float64 else ctypes.c_float
        out = (np.ctypeslib.as_array((ct * n).from_address(p)).copy() if n
               else np.zeros(0, vt))
    finally:
        _libc.free(p)
    return out


class DeviceCodec:
    """Device-resident compress/decompress on torch CUDA(HIP) tensors.

    fast_mode is the reference's fastMode: SZP_NONRANDOMACCESS (1, default)
    or SZP_RANDOMACCESS (2); float32 and float64 tensors are both served.
    Holds the output and workspace tensors of the last shape it served, so
    repeated calls on one field allocate nothing.  Calls are asynchronous on
    the given torch stream (default: the current stream); `check()` waits and
    raises on a device-side error.
    """

    def __init__(self, block_size=128, threads=None, device=None, fast_mode=SZP_NONRANDOMACCESS):
        import torch
        self.torch = torch
        self.B = int(block_size)
        self.T = int(threads) if threads else default_threads()
        self.mode = int(fast_mode)
        if self.mode not in (SZP_NONRANDOMACCESS, SZP_RANDOMACCESS):
            raise SzpError("unsupported fastMode %r" % (fast_mode,))
        self.device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        self._bufs = {}
        self.last_work = None

    def _buf(self, key, nbytes, dtype=None):
        t = self._bufs.get(key)
        torch = self.torch
        if t is None or t.numel() < nbytes:
            t = torch.empty(max(int(nbytes), 256), dtype=torch.uint8, device=self.device)
            self._bufs[key] = t
        return t

    def _stream(self, stream):
        torch = self.torch
        s = stream if stream is not None else torch.cuda.current_stream(self.device)
        return ctypes.c_void_p(s.cuda_stream)

    def compress(self, x, err_mode=REL, abs_eb=0.0, rel=1e-4, out=None, stream=None, index=None):
        """Returns (stream_tensor uint8[capacity], size_tensor int64[1]); the
        valid stream is stream_tensor[:size].  With index=True (or a uint8
        tensor of index_size bytes) it also writes the side index and returns
        (stream_tensor, size_tensor, index_tensor)."""
        torch = self.torch
        assert x.is_cuda and x.dtype in (torch.float32, torch.float64) and x.is_contiguous()
        dt = SZ_DOUBLE if x.dtype == torch.float64 else SZ_FLOAT
        n = x.numel()
        cap = compress_bound(n, self.B, self.T, dt, self.mode)
        wsb = lib.szp_amd_compress_workspace_size_typed(self.mode, n, self.B, self.T)
        if out is None:
            out = self._buf("cout", cap)
        size = self._buf("csize", 256).view(torch.int64)
        work = self._buf("cwork", wsb)
        self.last_work = work
        idx = None
        if index is not None and index is not False:
            idx = self._buf("cidx", self.index_size(n)) if index is True else index
        if dt == SZ_FLOAT and self.mode == SZP_NONRANDOMACCESS:
            # The float fastMode-1 entry points of the device API.
            if idx is None:
                rc = lib.szp_amd_compress_dev(x.data_ptr(), n, int(err_mode), float(abs_eb),
                                              float(rel), self.B, self.T, out.data_ptr(),
                                              out.numel(), size.data_ptr(), work.data_ptr(),
                                              work.numel(), self._stream(stream))
            else:
                rc = lib.szp_amd_compress_dev_indexed(x.data_ptr(), n, int(err_mode),
                                                      float(abs_eb), float(rel), self.B, self.T,
                                                      out.data_ptr(), out.numel(),
                                                      size.data_ptr(), idx.data_ptr(),
                                                      idx.numel(), work.data_ptr(), work.numel(),
                                                      self._stream(stream))
        else:
            rc = lib.szp_amd_compress_dev_typed(self.mode, dt, x.data_ptr(), n, int(err_mode),
                                                float(abs_eb), float(rel), self.B, self.T,
                                                out.data_ptr(), out.numel(), size.data_ptr(),
                                                idx.data_ptr() if idx is not None else None,
                                                idx.numel() if idx is not None else 0,
                                                work.data_ptr(), work.numel(),
                                                self._stream(stream))
        if rc:
            raise SzpError("compress", rc)
        return (out, size[:1]) if idx is None else (out, size[:1], idx)

    def index_size(self, n):
        return int(lib.szp_amd_index_size_typed(self.mode, int(n), self.B, self.T))

    def decompress(self, stream_t, stream_bytes, n, out=None, stream=None, index=None,
                   dtype=None):
        """Decodes stream_t[:stream_bytes] (stream_bytes may be any upper bound
        inside the tensor).  With `index` (from compress(index=...)) the block
        boundaries and carries come from the side index instead of a parse.
        dtype (torch.float32 default, or torch.float64) is the field's type;
        `out`, when given, fixes it."""
        torch = self.torch
        assert stream_t.is_cuda and stream_t.dtype == torch.uint8
        wsb = lib.szp_amd_decompress_workspace_size_typed(self.mode, int(stream_bytes), int(n),
                                                          self.B, self.T)
        work = self._buf("dwork", wsb)
        self.last_work = work
        if out is None:
            out = torch.empty(int(n), dtype=dtype or torch.float32, device=self.device)
        dt = SZ_DOUBLE if out.dtype == torch.float64 else SZ_FLOAT
        if dt == SZ_FLOAT and self.mode == SZP_NONRANDOMACCESS:
            if index is None:
                rc = lib.szp_amd_decompress_dev(stream_t.data_ptr(), int(stream_bytes), int(n),
                                                self.B, self.T, out.data_ptr(), work.data_ptr(),
                                                work.numel(), self._stream(stream))
            else:
                rc = lib.szp_amd_decompress_dev_indexed(stream_t.data_ptr(), int(stream_bytes),
                                                        index.data_ptr(), index.numel(), int(n),
                                                        self.B, self.T, out.data_ptr(),
                                                        work.data_ptr(), work.numel(),
                                                        self._stream(stream))
        else:
            rc = lib.szp_amd_decompress_dev_typed(self.mode, dt, stream_t.data_ptr(),
                                                  int(stream_bytes),
                                                  index.data_ptr() if index is not None else None,
                                                  index.numel() if index is not None else 0,
                                                  int(n), self.B, self.T, out.data_ptr(),
                                                  work.data_ptr(), work.numel(),
                                                  self._stream(stream))
        if rc:
            raise SzpError("decompress", rc)
        return out

    def check(self, stream=None):
        """Waits for the stream and raises on an error recorded by the device."""
        for key in ("cwork", "dwork"):
            work = self._bufs.get(key)
            if work is None:
                continue
            rc = lib.szp_amd_check_status(work.data_ptr(), self._stream(stream))
            if rc:
                raise SzpError("device status (%s)" % ("compress" if key == "cwork" else "decompress"), rc)
