"""Serving form: a request loop with continuous batching.

Requests (compress a resident float32 field, or decompress a resident stream
compressed earlier with its side index) arrive over time.  The loop admits
them against an HBM budget in arrival order, and whenever one of its streams
is free it launches everything admitted so far as one batch (one compress
batch and/or one decompress batch, szp_amd.batch).  Requests that arrive
while batches run are admitted as memory frees and go into the next batch,
so work from several batches overlaps on the device.  Each request's result
is the same bytes as a single-field call on the same input.
"""
import collections
import time

from . import REL, compress_bound, lib
from .batch import CompressBatch, DecompressBatch


class Request:
    __slots__ = ("rid", "kind", "x", "mode", "stream", "nbytes", "n", "index", "arrival",
                 "admitted", "launched", "done", "result", "result_size", "result_index", "need")

    def __init__(self, rid, kind, arrival, x=None, mode=(REL, 0.0, 1e-3), stream=None, nbytes=0,
                 n=0, index=None):
        self.rid, self.kind, self.arrival = rid, kind, arrival
        self.x, self.mode = x, mode
        self.stream, self.nbytes, self.n, self.index = stream, nbytes, n, index
        self.admitted = self.launched = self.done = None
        self.result = self.result_size = self.result_index = None
        self.need = 0


class Server:
    """Continuous-batching request loop on one GPU.

    budget: bytes of HBM the in-flight requests may hold (outputs, side
    indexes, batch workspaces); inputs are resident and not counted.
    streams: batches in flight at once.
    """

    def __init__(self, block_size=128, threads=16, budget=64 << 30, streams=3, ticket=2, grid=0,
                 device=None):
        import torch
        self.torch = torch
        self.B, self.T = int(block_size), int(threads)
        self.budget = int(budget)
        self.device = torch.device(device) if device is not None else torch.device("cuda")
        self.streams = [torch.cuda.Stream(self.device) for _ in range(streams)]
        self.ticket, self.grid = ticket, grid
        self.used = 0
        self.peak = 0
        self.batches = 0

    def _need(self, r):
        if r.kind == "compress":
            n = r.x.numel()
            return compress_bound(n, self.B, self.T) + int(lib.szp_amd_index_size(n, self.B, self.T)) + 4096
        return 4 * r.n + 4096

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
                cb.launch(stream=stream, ticket=self.ticket, grid=self.grid)
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
                r.need = self._need(r)
                if self.used + r.need > self.budget and (admitted or inflight):
                    break
                pending.popleft()
                self.used += r.need
                self.peak = max(self.peak, self.used)
                r.admitted = now
                admitted.append(r)
            # Launch everything admitted on a free stream.
            free = [i for i in range(len(self.streams)) if i not in inflight]
            if admitted and free:
                si = free[0]
                tl = time.perf_counter() - t0
                for r in admitted:
                    r.launched = tl
                inflight[si] = self._launch(admitted, self.streams[si])
                self.batches += 1
                admitted = []
                continue
            if not inflight and not admitted and pending:
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
    """p50/p99 latency (ms) per kind and aggregate input GB/s."""
    out = {}
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
