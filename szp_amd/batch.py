"""Batched compress/decompress: many fields of any sizes in one persistent
launch (szp_amd_batch_* in include/szp_amd.h), its capture in a HIP graph,
and an autotuner for the launch shape.

The reference compresses one field per call (SZp/src/szp.cc:23-115); a
batch produces, for every field, the same bytes as a single-field call with
the same block size B and team size T.
"""
import ctypes

from . import ABS, REL, SzpError, compress_bound, default_threads, lib

_u64, _i32, _vp, _dbl = ctypes.c_uint64, ctypes.c_int, ctypes.c_void_p, ctypes.c_double


class Field(ctypes.Structure):
    """struct szp_amd_field."""
    _fields_ = [("in_", _vp), ("out", _vp), ("n", _u64), ("in_bytes", _u64), ("out_cap", _u64),
                ("index", _vp), ("err_mode", _i32), ("abs_eb", _dbl), ("rel", _dbl)]


class _Plan(ctypes.Structure):
    """struct szp_amd_batch (opaque)."""
    _fields_ = [("opaque", _u64 * 16)]


lib.szp_amd_batch_workspace_size.restype = _u64
lib.szp_amd_batch_workspace_size.argtypes = [ctypes.POINTER(Field), _i32, _i32, _i32, _i32]
lib.szp_amd_batch_prepare.restype = _i32
lib.szp_amd_batch_prepare.argtypes = [ctypes.POINTER(Field), _i32, _i32, _i32, _i32, _vp, _vp,
                                      _u64, ctypes.POINTER(_Plan)]
lib.szp_amd_batch_launch.restype = _i32
lib.szp_amd_batch_launch.argtypes = [ctypes.POINTER(_Plan), ctypes.c_uint32, ctypes.c_uint32, _vp]
lib.szp_amd_batch_resident_grid.restype = ctypes.c_uint32
lib.szp_amd_batch_resident_grid.argtypes = [ctypes.POINTER(_Plan)]

# grid value of szp_amd_batch_launch for one compress workgroup per tile
# (SZP_AMD_BATCH_GRID_PER_TILE in include/szp_amd.h).
GRID_PER_TILE = 0xFFFFFFFF


def _align(v, a=256):
    return (v + a - 1) // a * a


def _stream_ptr(torch, stream, device):
    s = stream if stream is not None else torch.cuda.current_stream(device)
    return ctypes.c_void_p(s.cuda_stream)


class _Batch:
    def __init__(self, torch, fields, B, T, decode, device, sizes=None):
        self.torch = torch
        self.B, self.T, self.decode = B, T, decode
        self.device = device
        self.F = len(fields)
        self._fields = (Field * self.F)(*fields)
        wsb = lib.szp_amd_batch_workspace_size(self._fields, self.F, B, T, int(decode))
        if wsb == 0:
            raise SzpError("szp_amd_batch_workspace_size rejected the fields")
        self.work = torch.empty(int(wsb), dtype=torch.uint8, device=device)
        self.sizes = sizes
        self.plan = _Plan()
        rc = lib.szp_amd_batch_prepare(self._fields, self.F, B, T, int(decode),
                                       sizes.data_ptr() if sizes is not None else None,
                                       self.work.data_ptr(), self.work.numel(),
                                       ctypes.byref(self.plan))
        if rc:
            raise SzpError("szp_amd_batch_prepare", rc)
        # Defaults: decompress two tiles per ticket on the resident grid;
        # compress one workgroup per tile.
        self.ticket, self.grid = (2, 0) if decode else (1, GRID_PER_TILE)

    def launch(self, stream=None, ticket=None, grid=None):
        """Enqueues the batch (kernels only) on `stream`."""
        t = self.ticket if ticket is None else int(ticket)
        g = self.grid if grid is None else int(grid)
        rc = lib.szp_amd_batch_launch(ctypes.byref(self.plan), t, g,
                                      _stream_ptr(self.torch, stream, self.device))
        if rc:
            raise SzpError("szp_amd_batch_launch", rc)

    def resident_grid(self):
        """Workgroups of the persistent kernel that fit the device at once."""
        return int(lib.szp_amd_batch_resident_grid(ctypes.byref(self.plan)))

    def check(self, stream=None):
        rc = lib.szp_amd_check_status(self.work.data_ptr(),
                                      _stream_ptr(self.torch, stream, self.device))
        if rc:
            raise SzpError("batch device status", rc)

    def graph(self, stream=None):
        """Captures one launch of the batch in a HIP graph (torch.cuda.CUDAGraph);
        g.replay() reruns it."""
        torch = self.torch
        s = stream if stream is not None else torch.cuda.Stream(self.device)
        torch.cuda.synchronize(self.device)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=s):
            self.launch(stream=s)
        return g


class CompressBatch(_Batch):
    """Compress a list of float32 CUDA tensors in one launch.

    modes: per field (err_mode, abs_eb, rel) or one tuple for all.  Each
    field's stream is written to its own slice of one output tensor; with
    index=True each field also gets its side index (for DecompressBatch).
    """

    def __init__(self, xs, modes=(REL, 0.0, 1e-3), block_size=128, threads=None, index=True,
                 device=None):
        import torch
        device = torch.device(device) if device is not None else xs[0].device
        B = int(block_size)
        T = int(threads) if threads else default_threads()
        if isinstance(modes, tuple):
            modes = [modes] * len(xs)
        self.xs = xs
        self.offsets, self.caps = [], []
        off = 0
        for x in xs:
            assert x.is_cuda and x.dtype == torch.float32 and x.is_contiguous() and x.numel() > 0
            cap = compress_bound(x.numel(), B, T)
            self.offsets.append(off)
            self.caps.append(cap)
            off += _align(cap)
        self.out = torch.empty(max(off, 256), dtype=torch.uint8, device=device)
        self.index_offsets, self.index_sizes = [], []
        ioff = 0
        for x in xs:
            ib = int(lib.szp_amd_index_size(x.numel(), B, T))
            self.index_offsets.append(ioff)
            self.index_sizes.append(ib)
            ioff += _align(ib)
        self.index = (torch.empty(max(ioff, 256), dtype=torch.uint8, device=device)
                      if index else None)
        sizes = torch.zeros(len(xs), dtype=torch.int64, device=device)
        fields = []
        for k, x in enumerate(xs):
            em, ab, rel = modes[k]
            fields.append(Field(x.data_ptr(), self.out.data_ptr() + self.offsets[k], x.numel(), 0,
                                self.caps[k],
                                (self.index.data_ptr() + self.index_offsets[k]) if index else None,
                                int(em), float(ab), float(rel)))
        super().__init__(torch, fields, B, T, False, device, sizes)

    def stream(self, k, size=None):
        """Field k's stream (uint8 tensor view); size: from sizes() if None."""
        if size is None:
            size = int(self.sizes[k].item())
        return self.out[self.offsets[k]:self.offsets[k] + size]

    def index_of(self, k):
        return self.index[self.index_offsets[k]:self.index_offsets[k] + self.index_sizes[k]]

    def byte_sizes(self):
        return [int(v) for v in self.sizes.tolist()]


class DecompressBatch(_Batch):
    """Decompress fields compressed with a side index, in one launch.

    streams: uint8 CUDA tensors (16-byte aligned starts); sizes: stream
    lengths; ns: element counts; indexes: their side indexes."""

    def __init__(self, streams, sizes, ns, indexes, block_size=128, threads=None, device=None,
                 outs=None):
        import torch
        device = torch.device(device) if device is not None else streams[0].device
        B = int(block_size)
        T = int(threads) if threads else default_threads()
        self.outs = outs if outs is not None else [torch.empty(int(n), dtype=torch.float32,
                                                               device=device) for n in ns]
        fields = [Field(s.data_ptr(), o.data_ptr(), int(n), int(sz), 0, ix.data_ptr(), 0, 0.0, 0.0)
                  for s, sz, n, ix, o in zip(streams, sizes, ns, indexes, self.outs)]
        self._keep = (streams, indexes)
        super().__init__(torch, fields, B, T, True, device)


# ---- autotuner --------------------------------------------------------------

_TUNED = {}


def _time_launch(batch, ticket, grid, reps):
    torch = batch.torch
    s = torch.cuda.current_stream(batch.device)
    batch.launch(stream=s, ticket=ticket, grid=grid)  # warm-up
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(s)
    for _ in range(reps):
        batch.launch(stream=s, ticket=ticket, grid=grid)
    e1.record(s)
    e1.synchronize()
    batch.check(s)
    return e0.elapsed_time(e1) / reps


def autotune(batch, tickets=(1, 2, 4, 8, 16), grid_fractions=(1.0, 0.75, 0.5, 0.25), reps=3,
             key=None):
    """Times the batch under each launch shape and keeps the fastest on the
    batch (and in a cache under `key`).  The shapes searched:

    * form: for compress, the persistent kernel (workgroups loop over
      tickets) or one workgroup per tile (GRID_PER_TILE: no loop, so the
      kernel runs at the single-field kernel's occupancy);
    * grid: the persistent kernel's workgroups, as fractions of the resident
      grid (CUs x workgroups per CU, from the kernel's occupancy);
    * ticket: tiles a workgroup takes at a time (decompress; a compress tile
      waits for the tiles before it to start, so compress takes one).

    Workgroup size (256 threads, four wave64s) and values per thread (one
    octet of 8 per round) are fixed by the tile geometry, which maps a
    round's 2048 values onto the workgroup; the stream does not depend on
    any of the shapes.  Returns {"ticket", "grid", "form", "ms", "trials"}."""
    if key is not None and key in _TUNED:
        best = _TUNED[key]
        batch.ticket, batch.grid = best["ticket"], best["grid"]
        return best
    resident = batch.resident_grid() or 1024
    if not batch.decode:
        tickets = (1,)
    trials = []
    for frac in grid_fractions:
        grid = 0 if frac == 1.0 else max(1, int(resident * frac))
        for t in tickets:
            ms = _time_launch(batch, t, grid, reps)
            trials.append({"ticket": t, "grid": grid, "form": "persistent", "ms": ms})
    if not batch.decode:
        ms = _time_launch(batch, 1, GRID_PER_TILE, reps)
        trials.append({"ticket": 1, "grid": GRID_PER_TILE, "form": "per_tile", "ms": ms})
    best = min(trials, key=lambda r: r["ms"])
    batch.ticket, batch.grid = best["ticket"], best["grid"]
    res = dict(best, trials=trials, resident_grid=resident)
    if key is not None:
        _TUNED[key] = res
    return res


def plain_loop(xs, modes, block_size, threads, index=True, stream=None):
    """The per-field baseline: one single-field device call per field (what a
    caller without the batch API does).  Returns the per-field codecs."""
    from . import DeviceCodec
    codecs = []
    for k, x in enumerate(xs):
        em, ab, rel = modes[k] if isinstance(modes, list) else modes
        c = DeviceCodec(block_size=block_size, threads=threads, device=x.device)
        c.compress(x, em, ab, rel, index=True if index else None, stream=stream)
        codecs.append(c)
    return codecs
