"""Serving form: a request loop with continuous batching.

Requests (compress a resident float32 field, or decompress a resident stream
compressed earlier with its side index) arrive over time.  The loop admits
them against an HBM budget in arrival order, and whenever one of its streams
is free it launches everything admitted so far as one batch (one compress
batch and/or one decompress batch, szp_amd.batch).  Requests that arrive
while batches run are admitted as memory frees and go into the next batch,
so work from several batches overlaps on the device.  Each request's result
is the same bytes as a single-field call on the same input.

A request is charged, at admission, for everything the server allocates for
it: its output (stream buffer or values), its side index, and its share of a
batch workspace (the workspace of a batch of that request alone, which bounds
its share of any batch).  A request larger than the whole budget is refused.
A request can be cancelled until its batch is launched (Server.cancel or
Request.cancel_at); its budget is released at once.
"""
import collections
import time

from . import REL, compress_bound, lib
from .batch import GRID_PER_TILE, CompressBatch, DecompressBatch, Field


class Request:
    __slots__ = ("rid", "kind", "x", "mode", "stream", "nbytes", "n", "index", "arrival",
                 "admitted", "launched", "done", "result", "result_size", "result_index", "need",
                 "cancel_at", "state")

    def __init__(self, rid, kind, arrival, x=None, mode=(REL, 0.0, 1e-3), stream=None, nbytes=0,
                 n=0, index=None, cancel_at=None):
        self.rid, self.kind, self.arrival = rid, kind, arrival
        self.x, self.mode = x, mode
        self.stream, self.nbytes, self.n, self.index = stream, nbytes, n, index
        self.admitted = self.launched = self.done = None
        self.result = self.result_size = self.result_index = None
        self.need = 0
        # Seconds from the start at which the client cancels (None: never).
        self.cancel_at = cancel_at
        # "pending", "admitted", "launched", "done", "cancelled" or "refused".
        self.state = "pending"


class Server:
    """Continuous-batching request loop on one GPU.

    budget: bytes of HBM the in-flight requests may hold (outputs, side
    indexes, batch workspaces); inputs are resident and not counted.
    streams: batches in flight at once.  ticket, grid: the decompress
    batches' launch shape; compress_grid: the compress batches' (one
    workgroup per tile by default, the faster form on mixed sizes).
    """

    def __init__(self, block_size=128, threads=16, budget=64 << 30, streams=3, ticket=2, grid=0,
                 compress_grid=GRID_PER_TILE, device=None):
        import torch
        self.torch = torch
        self.B, self.T = int(block_size), int(threads)
        self.budget = int(budget)
        self.device = torch.device(device) if device is not None else torch.device("cuda")
        self.streams = [torch.cuda.Stream(self.device) for _ in range(streams)]
        self.ticket, self.grid, self.compress_grid = ticket, grid, compress_grid
        self.used = 0
        self.peak = 0
        self.batches = 0
        self.cancelled = 0
        self.refused = 0
        self._cancel = set()

    def cancel(self, rid):
        """Asks for request `rid` to be cancelled (effective if its batch has
        not been launched yet; may be called from another thread)."""
        self._cancel.add(rid)

    def _workspace(self, r):
        """Bytes of a batch workspace for request r alone."""
        if r.kind == "compress":
            em, ab, rel = r.mode
            f = Field(None, None, r.x.numel(), 0, 0, None, int(em), float(ab), float(rel))
        else:
            f = Field(None, None, int(r.n), int(r.nbytes), 0, None, 0, 0.0, 0.0)
        arr = (Field * 1)(f)
        return int(lib.szp_amd_batch_workspace_size(arr, 1, self.B, self.T,
                                                    0 if r.kind == "compress" else 1))

    def _need(self, r):
        """Bytes the server allocates for r: output, side index and its
        share of a batch workspace (each rounded to the 256-byte alignment
        the batch gives it)."""
        a = lambda v: (int(v) + 255) // 256 * 256  # noqa: E731
        if r.kind == "compress":
            n = r.x.numel()
            return (a(compress_bound(n, self.B, self.T)) +
                    a(lib.szp_amd_index_size(n, self.B, self.T)) + self._workspace(r))
        return a(4 * r.n) + self._workspace(r)

    def _cancelled(self, r, now):
        return r.rid in self._cancel or (r.cancel_at is not None and now >= r.cancel_at)

    def _launch(self, reqs, stream):
        """One batch for `reqs` on `stream`; returns the objects to keep alive."""
        torch = self.torch
        keep = []
        comp = [r for r in reqs if r.kind == "compress"]
        deco = [r for r in reqs if r.kind == "decompress"]
        with torch.cuda.stream(stream):
            if comp:
                cb = CompressBatch([r.x for r in comp], [r.mode for r in comp], block_size=self.B,
                                   threads=self.T, index=True, device=self.device)
                cb.launch(stream=stream, ticket=1, grid=self.compress_grid)
                keep.append(("c", cb, comp))
            if deco:
                db = DecompressBatch([r.stream for r in deco], [r.nbytes for r in deco],
                                     [r.n for r in deco], [r.index for r in deco],
                                     block_size=self.B, threads=self.T, device=self.device)
                db.launch(stream=stream, ticket=self.ticket, grid=self.grid)
                keep.append(("d", db, deco))
        ev = torch.cuda.Event()
        ev.record(stream)
        return ev, keep

    def _finish(self, keep, t):
        for tag, b, reqs in keep:
            b.check()
            if tag == "c":
                sizes = b.byte_sizes()
                for k, r in enumerate(reqs):
                    r.result_size = sizes[k]
                    r.result = b.stream(k, sizes[k])
                    r.result_index = b.index_of(k)
            else:
                for k, r in enumerate(reqs):
                    r.result = b.outs[k]
            for r in reqs:
                r.done = t
                r.state = "done"
                self.used -= r.need

    def run(self, requests, keep_results=True, poll_sleep=20e-6):
        """Serves `requests` (arrival times in seconds from the start, in
        order).  Returns the requests with admitted/launched/done times."""
        torch = self.torch
        pending = collections.deque(sorted(requests, key=lambda r: r.arrival))
        admitted = []
        inflight = {}  # stream index -> (event, keep)
        t0 = time.perf_counter()
        served = []
        while pending or admitted or inflight:
            now = time.perf_counter() - t0
            # Cancellations of requests not launched yet (their budget is
            # released at once).
            if self._cancel or any(r.cancel_at is not None for r in admitted):
                keep_adm = []
                for r in admitted:
                    if self._cancelled(r, now):
                        self.used -= r.need
                        r.state, r.done = "cancelled", now
                        self.cancelled += 1
                        served.append(r)
                    else:
                        keep_adm.append(r)
                admitted = keep_adm
            # Retire finished batches (frees their memory).
            for si in list(inflight):
                ev, keep = inflight[si]
                if ev.query():
                    self._finish(keep, time.perf_counter() - t0)
                    for _, _, reqs in keep:
                        served.extend(reqs)
                        if not keep_results:
                            for r in reqs:
                                r.result = r.result_index = None
                    del inflight[si]
            # Admit arrived requests in order while the budget allows.
            while pending and pending[0].arrival <= now:
                r = pending[0]
                if self._cancelled(r, now):
                    pending.popleft()
                    r.state, r.done = "cancelled", now
                    self.cancelled += 1
                    served.append(r)
                    continue
                r.need = self._need(r)
                if r.need > self.budget:
                    pending.popleft()
                    r.state, r.done = "refused", now
                    self.refused += 1
                    served.append(r)
                    continue
                if self.used + r.need > self.budget:
                    break
                pending.popleft()
                self.used += r.need
                self.peak = max(self.peak, self.used)
                r.admitted = now
                r.state = "admitted"
                admitted.append(r)
            # Launch everything admitted on a free stream.
            free = [i for i in range(len(self.streams)) if i not in inflight]
            if admitted and free:
                si = free[0]
                tl = time.perf_counter() - t0
                for r in admitted:
                    r.launched = tl
                    r.state = "launched"
                inflight[si] = self._launch(admitted, self.streams[si])
                self.batches += 1
                admitted = []
                continue
            if not inflight and not admitted and pending and pending[0].arrival > now:
                wait = pending[0].arrival - (time.perf_counter() - t0)
                if wait > 0:
                    time.sleep(min(wait, 1e-3))
            else:
                time.sleep(poll_sleep)
        torch.cuda.synchronize(self.device)
        return served


def percentile(v, q):
    import numpy as np
    return float(np.percentile(np.asarray(v, dtype=np.float64), q)) if len(v) else 0.0


def report(served, wall):
    """p50/p99 latency (ms) per kind and aggregate input GB/s, over the
    requests served (cancelled and refused ones are counted apart)."""
    out = {}
    done = [r for r in served if r.state == "done"]
    out["cancelled"] = sum(1 for r in served if r.state == "cancelled")
    out["refused"] = sum(1 for r in served if r.state == "refused")
    served = done
    for kind in ("compress", "decompress"):
        lat = [(r.done - r.arrival) * 1e3 for r in served if r.kind == kind]
        qd = [(r.launched - r.arrival) * 1e3 for r in served if r.kind == kind]
        out[kind] = {"requests": len(lat), "latency_ms_p50": round(percentile(lat, 50), 3),
                     "latency_ms_p99": round(percentile(lat, 99), 3),
                     "queue_ms_p50": round(percentile(qd, 50), 3)}
    nbytes = sum((r.x.numel() * 4 if r.kind == "compress" else r.n * 4) for r in served)
    out["bytes"] = nbytes
    out["aggregate_GBps"] = round(nbytes / wall / 1e9, 2) if wall > 0 else 0.0
    return out
