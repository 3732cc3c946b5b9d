#!/usr/bin/env python3
"""SZp fastmode1 float32 compress + decompress on one MI355X.

Workload: a 512^3 float32 Gaussian random field with isotropic power spectrum
proportional to k^-3, normalised to [-1, 1] (synthetic, seeded), compressed
with a relative error bound of 1e-4 and block size 128 into T = 16 chunks.
One step = one compress of the field (stream plus the codec's side index of
block positions) plus one decompress of its stream using that index, with the
field already resident in HBM.  Decompressing the same bytes as a foreign
stream (no index: block boundaries recovered by the speculative parse) and
compressing without the index are timed separately and reported under
"paths".  Variants (REL 1e-2, B = 32) are measured
after the main line and reported beside it.

value = 2 * 4N / step time: uncompressed bytes through the codec per second
(compress consumes 4N bytes, decompress produces 4N bytes).

The stream is checked byte-for-byte, and the decompression bit-for-bit,
against the reference compiled from its own sources (oracle/_ref) run on the
same field with the same T; the reference's OpenMP path on the host cores is
also timed as the CPU baseline.  Prints one JSON line.
"""
import argparse
import ctypes
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
HBM_PEAK_GBPS = 8000.0
sys.path.insert(0, os.path.join(ROOT, "tests"))
import workloads  # noqa: E402  (the configurations' synthetic fields)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def load_reference(threads):
    """The reference library compiled from its own sources, else the oracle."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib
    if oracle_lib.reference() is not None:
        kind = "reference"

        def comp(x, rel, B):
            return oracle_lib.ref_compress(x, 1, 0.0, rel, B, threads)

        def decomp(s, n, B):
            return oracle_lib.ref_decompress(s, n, B, threads)
    else:
        kind = "port"

        def comp(x, rel, B):
            return oracle_lib.compress(x, 1, 0.0, rel, B, threads, threads)

        def decomp(s, n, B):
            return oracle_lib.decompress(s, n, B, threads, threads)
    return kind, comp, decomp


def cpu_threads():
    """OpenMP threads of the CPU baseline: OMP_NUM_THREADS when set (the
    host's CPU share for this GPU), else the CPUs this process may run on."""
    v = os.environ.get("OMP_NUM_THREADS", "")
    if v.isdigit() and int(v) > 0:
        return int(v), "OMP_NUM_THREADS"
    return len(os.sched_getaffinity(0)), "sched_getaffinity"


def cpu_timing(comp, decomp, xh, rel, B, warmup=2, reps=10):
    """2 warm-ups, then the median (and min) of `reps` API-level wall times of
    compress and of decompress, as the reference's drivers time them."""
    N = xh.size
    tc, td = [], []
    stream = dec = None
    for k in range(warmup + reps):
        t0 = time.perf_counter()
        stream = comp(xh, rel, B)
        t1 = time.perf_counter()
        dec = decomp(stream, N, B)
        t2 = time.perf_counter()
        if k >= warmup:
            tc.append(t1 - t0)
            td.append(t2 - t1)
    return dict(compress_s=statistics.median(tc), decompress_s=statistics.median(td),
                compress_min_s=min(tc), decompress_min_s=min(td), stream=stream, dec=dec)


def c1_bench(torch, szp, n=1 << 26, rel=1e-3, B=128, T=16, seed=12345, steps=10, cpu=True):
    """The CPU-path configuration: a 64 Mi-element sum of 8 sinusoids plus
    N(0, 1e-3) noise, REL 1e-3 (eb = 1e-3 x value range), B = 128.  The
    reference on the host cores (2 warm-ups, median of 10) and the GPU on the
    same field, with byte parity at the GPU's team size T."""
    import numpy as np
    x = workloads.sinusoids(torch, n, seed)
    codec = szp.DeviceCodec(block_size=B, threads=T)
    out, size_t, idx = codec.compress(x, szp.REL, 0.0, rel, index=True)
    codec.check()
    C = int(size_t.item())
    dec = torch.empty_like(x)
    codec.decompress(out, C, n, out=dec, index=idx)
    codec.check()
    comp_ms = _timed(torch, lambda: codec.compress(x, szp.REL, 0.0, rel, out=out, index=idx), steps)
    dec_ms = _timed(torch, lambda: codec.decompress(out, C, n, out=dec, index=idx), steps)
    dec_parse_ms = _timed(torch, lambda: codec.decompress(out, C, n, out=dec), steps)
    codec.check()
    stream_head = out[:8].cpu().numpy()
    res = {"workload": "1-D float32, %d elements: 8 sinusoids (f log-uniform 1e-6..1e-3 "
                       "cycles/element, a ~ U(0.1,1)) + N(0,1e-3), seed %d; REL %g, B=%d"
                       % (n, seed, rel, B),
           "n_elements": n, "compressed_bytes": C, "ratio": round(4 * n / C, 4),
           "gpu": {"T": T, "compress_GBps": round(4 * n / (comp_ms * 1e-3) / 1e9, 1),
                   "decompress_GBps": round(4 * n / (dec_ms * 1e-3) / 1e9, 1),
                   "decompress_foreign_stream_GBps": round(4 * n / (dec_parse_ms * 1e-3) / 1e9, 1),
                   "compress_hbm_frac": round((8 * n + C) / (comp_ms * 1e-3) / 8e12, 4),
                   "decompress_hbm_frac": round((C + 4 * n) / (dec_ms * 1e-3) / 8e12, 4)},
           "error_bound": eb_report(torch, x, dec, stream_head)}
    if cpu:
        threads, src = cpu_threads()
        kind, rcomp, rdecomp = load_reference(threads)
        xh = x.cpu().numpy()
        t = cpu_timing(rcomp, rdecomp, xh, rel, B)
        err = np.abs(xh.astype(np.float64) - t["dec"].astype(np.float64))
        res["cpu"] = {"kind": kind, "cores": threads, "cores_from": src,
                      "compress_GBps": round(4 * n / t["compress_s"] / 1e9, 3),
                      "decompress_GBps": round(4 * n / t["decompress_s"] / 1e9, 3),
                      "compress_s_median_of_10": round(t["compress_s"], 4),
                      "decompress_s_median_of_10": round(t["decompress_s"], 4),
                      "ratio": round(4 * n / t["stream"].size, 4),
                      "max_abs_err": float(err.max())}
        if threads != T:
            kind_t, rcomp_t, rdecomp_t = load_reference(T)
            ref = rcomp_t(xh, rel, B)
            ref_dec = rdecomp_t(ref, n, B)
        else:
            ref, ref_dec = t["stream"], t["dec"]
        got = out[:C].cpu().numpy()
        res["parity"] = {"against": kind, "T": T,
                         "stream_byte_identical": bool(ref.size == C and np.array_equal(ref, got)),
                         "decompress_bit_identical": bool(np.array_equal(
                             ref_dec.view(np.uint32), dec.cpu().numpy().view(np.uint32)))}
    del x, out, dec, idx, codec
    torch.cuda.empty_cache()
    return res


def _timed(torch, fn, steps):
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(steps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps


def measure(torch, szp, codec, x, rel, steps, warmup, profile):
    """Step = compress (stream + side index) then decompress with the index.
    Also times, separately: compress without the index, and decompress of the
    same bytes as a foreign stream (block boundaries recovered by parsing)."""
    n = x.numel()
    out, size_t, idx = codec.compress(x, szp.REL, 0.0, rel, index=True)
    codec.check()
    nbytes = int(size_t.item())
    dec = torch.empty(n, dtype=torch.float32, device=x.device)
    for _ in range(warmup):
        codec.compress(x, szp.REL, 0.0, rel, out=out, index=idx)
        codec.decompress(out, nbytes, n, out=dec, index=idx)
        codec.decompress(out, nbytes, n, out=dec)
    codec.check()
    if int(size_t.item()) != nbytes:
        raise RuntimeError("compressed size changed between runs")

    def step():
        codec.compress(x, szp.REL, 0.0, rel, out=out, index=idx)
        codec.decompress(out, nbytes, n, out=dec, index=idx)

    all_kernels = {}
    if profile:
        # One untimed pass with every launch bracketed by events finds the
        # dominant kernel; the timed steps then bracket only that kernel, so
        # that the events disturb the step as little as possible.
        szp.profile_enable(True)
        for _ in range(max(warmup, 1)):
            step()
        all_kernels = szp.profile_read()
        avg = {k: v[1] / v[0] for k, v in all_kernels.items() if v[0]}
        szp.profile_enable(True, only=max(avg, key=avg.get))
    step_ms = _timed(torch, step, steps)
    kernels = szp.profile_read() if profile else {}
    if profile:
        szp.profile_enable(False)
    comp_ms = _timed(torch, lambda: codec.compress(x, szp.REL, 0.0, rel, out=out, index=idx), steps)
    dec_ms = _timed(torch, lambda: codec.decompress(out, nbytes, n, out=dec, index=idx), steps)
    comp_plain_ms = _timed(torch, lambda: codec.compress(x, szp.REL, 0.0, rel, out=out), steps)
    if profile:
        szp.profile_enable(True)
    dec_parse_ms = _timed(torch, lambda: codec.decompress(out, nbytes, n, out=dec), steps)
    parse_kernels = szp.profile_read() if profile else {}
    if profile:
        szp.profile_enable(False)
    codec.check()
    return dict(out=out, nbytes=nbytes, dec=dec, step_ms=step_ms, comp_ms=comp_ms, dec_ms=dec_ms,
                comp_plain_ms=comp_plain_ms, dec_parse_ms=dec_parse_ms, kernels=kernels,
                all_kernels=all_kernels, parse_kernels=parse_kernels,
                index_bytes=codec.index_size(n))


def eb_report(torch, x, dec, stream_head):
    eb = float(torch.frombuffer(bytearray(stream_head[:4][::-1]), dtype=torch.float32)[0])
    err = (x.double() - dec.double()).abs()
    tol = eb + 2.0 ** -22 * (eb + dec.double().abs())
    return dict(eb=eb, max_err_over_eb=float(err.max()) / eb,
                strict_over_eb=int((err > eb).sum()), within_rounding_bound=bool((err <= tol).all()))


def _pct(v, q):
    import numpy as np
    return float(np.percentile(np.asarray(v), q))


def batched_bench(torch, szp, count=256, rel=1e-3, B=128, T=16, seed=7, reps=5):
    """Uneven-size mix: per-field plain calls against one batched launch
    (autotuned, replayed from a graph), compress and decompress."""
    from szp_amd import batch as sb
    xs, kinds = workloads.mix(torch, count, 64 << 10, 64 << 20, seed)
    N = sum(x.numel() for x in xs)
    log("batched leg: %d fields, %.2f GB generated" % (len(xs), 4 * N / 1e9))
    mode = (szp.REL, 0.0, rel)
    s = torch.cuda.current_stream()

    # Plain: one single-field call per field, buffers allocated beforehand.
    codecs = [szp.DeviceCodec(block_size=B, threads=T) for _ in xs]
    for c, x in zip(codecs, xs):
        c.compress(x, *mode, index=True)
    torch.cuda.synchronize()
    outs = [c._bufs["cout"] for c in codecs]
    idxs = [c._bufs["cidx"] for c in codecs]
    sizes = [int(c._bufs["csize"].view(torch.int64)[0].item()) for c in codecs]
    decs = [torch.empty_like(x) for x in xs]
    for c, o, sz, x, d, ix in zip(codecs, outs, sizes, xs, decs, idxs):
        c.decompress(o, sz, x.numel(), out=d, index=ix)
    torch.cuda.synchronize()

    def plain_compress():
        ev = []
        for c, x, o, ix in zip(codecs, xs, outs, idxs):
            c.compress(x, *mode, out=o, index=ix)
            e = torch.cuda.Event(enable_timing=True)
            e.record(s)
            ev.append(e)
        return ev

    def plain_decompress():
        ev = []
        for c, o, sz, x, d, ix in zip(codecs, outs, sizes, xs, decs, idxs):
            c.decompress(o, sz, x.numel(), out=d, index=ix)
            e = torch.cuda.Event(enable_timing=True)
            e.record(s)
            ev.append(e)
        return ev

    def timed_plain(fn):
        best = None
        for _ in range(reps):
            torch.cuda.synchronize()
            e0 = torch.cuda.Event(enable_timing=True)
            e0.record(s)
            ev = fn()
            ev[-1].synchronize()
            lat = [e0.elapsed_time(e) for e in ev]
            if best is None or lat[-1] < best[-1]:
                best = lat
        return best

    def timed_graph(g):
        ms = []
        for _ in range(reps):
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(s)
            g.replay()
            e1.record(s)
            e1.synchronize()
            ms.append(e0.elapsed_time(e1))
        return min(ms)

    pc = timed_plain(plain_compress)
    pd = timed_plain(plain_decompress)
    log("batched leg: plain calls timed")

    # Host-free plain baseline: the same per-field calls captured in one
    # graph, so that the comparison isolates the persistent kernel from the
    # Python loop's launch cost.  Replays are checked against the eager
    # streams and values.
    def capture(fn):
        g = torch.cuda.CUDAGraph()
        torch.cuda.synchronize()
        with torch.cuda.graph(g):
            fn()
        torch.cuda.synchronize()
        return g

    def plain_compress_calls():
        for c, x, o, ix in zip(codecs, xs, outs, idxs):
            c.compress(x, *mode, out=o, index=ix)

    def plain_decompress_calls():
        for c, o, sz, x, d, ix in zip(codecs, outs, sizes, xs, decs, idxs):
            c.decompress(o, sz, x.numel(), out=d, index=ix)

    ref_streams = [o[:sz].clone() for o, sz in zip(outs, sizes)]
    ref_decs = [d.clone() for d in decs]
    gpc = capture(plain_compress_calls)
    gpd = capture(plain_decompress_calls)
    pgc = timed_graph(gpc)
    pgd = timed_graph(gpd)
    torch.cuda.synchronize()
    plain_graph_same = (all(torch.equal(o[:sz], r) for o, sz, r in zip(outs, sizes, ref_streams))
                        and all(torch.equal(d.view(torch.int32), r.view(torch.int32))
                                for d, r in zip(decs, ref_decs)))
    del gpc, gpd, ref_streams, ref_decs
    log("batched leg: graphed plain calls timed")

    # Batched: one persistent launch for all fields, autotuned, in a graph.
    cb = sb.CompressBatch(xs, mode, block_size=B, threads=T, index=True)
    tune_c = sb.autotune(cb)
    cb.launch()
    cb.check()
    bsizes = cb.byte_sizes()
    same = bsizes == sizes and all(torch.equal(cb.stream(k, sizes[k]), outs[k][:sizes[k]])
                                   for k in range(len(xs)))
    gc = cb.graph()
    db = sb.DecompressBatch([cb.stream(k, bsizes[k]) for k in range(len(xs))], bsizes,
                            [x.numel() for x in xs], [cb.index_of(k) for k in range(len(xs))],
                            block_size=B, threads=T)
    tune_d = sb.autotune(db)
    db.launch()
    db.check()
    same_dec = all(torch.equal(db.outs[k].view(torch.int32), decs[k].view(torch.int32))
                   for k in range(len(xs)))
    gd = db.graph()

    bc = timed_graph(gc)
    cb.check()
    bd = timed_graph(gd)
    db.check()
    log("batched leg: batch timed")
    C = sum(sizes)
    res = {
        "workload": "%d float32 fields, sizes log-uniform 64 KB-64 MB (seed %d), kinds smooth k^-3 / "
                    "U(-1,1) / 20%% zero blocks, REL %g, B=%d, T=%d" % (count, seed, rel, B, T),
        "fields": count, "elements": N, "input_bytes": 4 * N, "compressed_bytes": C,
        "kinds": {k: kinds.count(k) for k in set(kinds)},
        "plain_compress_GBps": round(4 * N / (pc[-1] * 1e-3) / 1e9, 1),
        "batched_compress_GBps": round(4 * N / (bc * 1e-3) / 1e9, 1),
        "compress_speedup": round(pc[-1] / bc, 2),
        "plain_decompress_GBps": round(4 * N / (pd[-1] * 1e-3) / 1e9, 1),
        "batched_decompress_GBps": round(4 * N / (bd * 1e-3) / 1e9, 1),
        "decompress_speedup": round(pd[-1] / bd, 2),
        "plain_graphed_compress_GBps": round(4 * N / (pgc * 1e-3) / 1e9, 1),
        "plain_graphed_decompress_GBps": round(4 * N / (pgd * 1e-3) / 1e9, 1),
        "compress_speedup_vs_graphed_plain": round(pgc / bc, 2),
        "decompress_speedup_vs_graphed_plain": round(pgd / bd, 2),
        "graphed_plain_replays_identical": bool(plain_graph_same),
        "plain_compress_latency_ms": {"p50": round(_pct(pc, 50), 3), "p99": round(_pct(pc, 99), 3)},
        "batched_compress_latency_ms": {"p50": round(bc, 3), "p99": round(bc, 3)},
        "plain_decompress_latency_ms": {"p50": round(_pct(pd, 50), 3), "p99": round(_pct(pd, 99), 3)},
        "batched_decompress_latency_ms": {"p50": round(bd, 3), "p99": round(bd, 3)},
        "latency_note": "time from the start of the sequence to each field's completion; the batch "
                        "completes all fields together",
        "autotune": {k: {"form": t["form"], "ticket": t["ticket"],
                         "grid": t["grid"] if t["form"] == "persistent" else "one per tile",
                         "resident_grid": t["resident_grid"], "shapes_tried": len(t["trials"])}
                     for k, t in (("compress", tune_c), ("decompress", tune_d))},
        "streams_byte_identical_to_plain": bool(same),
        "decompress_bit_identical_to_plain": bool(same_dec),
        "graph": "one hipGraph replay per batch (kernels only)",
    }
    del cb, db, gc, gd, codecs, outs, xs
    torch.cuda.empty_cache()
    return res


def serving_bench(torch, szp, rate=2000.0, count=400, pool_size=48, B=128, T=16, seed=11,
                  budget=64 << 30):
    """Request stream: Poisson arrivals at `rate`/s, 70% compress of a
    resident field / 30% decompress of a field compressed earlier, payloads
    log-uniform 16 KB-256 MB from the smooth or noisy generators, REL from
    {1e-2, 1e-3, 1e-4}, admission bounded by `budget` bytes of HBM."""
    import numpy as np
    from szp_amd import serve
    rng = np.random.default_rng(seed)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed)
    nbytes = np.exp(rng.uniform(np.log(16 << 10), np.log(256 << 20), pool_size)).astype(np.int64)
    pool = [workloads.field_1d(torch, ["smooth", "noise"][rng.integers(0, 2)], int(b // 4), gen)
            for b in nbytes]
    ebs = [1e-2, 1e-3, 1e-4]
    pre = []
    for x in pool:
        c = szp.DeviceCodec(block_size=B, threads=T)
        out, size_t, idx = c.compress(x, szp.REL, 0.0, ebs[rng.integers(0, 3)], index=True)
        c.check()
        size = int(size_t.item())
        pre.append((out[:size].clone(), size, x.numel(), idx.clone()))
    torch.cuda.synchronize()
    log("serving leg: pool of %d fields, %.2f GB" % (len(pool), sum(x.numel() for x in pool) * 4 / 1e9))

    def make_requests(k, lam, seed2):
        r2 = np.random.default_rng(seed2)
        reqs, t = [], 0.0
        for rid in range(k):
            t += r2.exponential(1.0 / lam)
            i = int(r2.integers(0, len(pool)))
            if r2.random() < 0.7:
                q = serve.Request(rid, "compress", t, x=pool[i],
                                  mode=(szp.REL, 0.0, ebs[int(r2.integers(0, 3))]))
            else:
                st, size, n, idx = pre[i]
                q = serve.Request(rid, "decompress", t, stream=st, nbytes=size, n=n, index=idx)
            q.result_index = i
            reqs.append(q)
        return reqs

    srv = serve.Server(block_size=B, threads=T, budget=budget, streams=3)
    srv.run(make_requests(32, rate, seed + 1), keep_results=False)  # warm-up
    # Heavier offered loads (same request mix) show throughput under queueing.
    loads = []
    # The last load runs against a 2 GiB budget, so that admission binds
    # (requests wait for budget to free up; none needs more than 2 GiB).
    small = serve.Server(block_size=B, threads=T, budget=2 << 30, streams=3)
    for k, (lam, sv_) in enumerate(((rate * 10, srv), (rate * 50, srv), (rate * 50, small))):
        rq = make_requests(count, lam, seed + 10 + k)
        sv_.batches = 0
        sv_.peak = 0
        t0 = time.perf_counter()
        sv = sv_.run(rq, keep_results=False)
        w = time.perf_counter() - t0
        rp = serve.report(sv, w)
        loads.append({"rate_per_s": lam, "budget_GiB": sv_.budget / (1 << 30),
                      "wall_seconds": round(w, 4), "batches": sv_.batches,
                      "peak_admitted_bytes": int(sv_.peak), "refused": rp["refused"],
                      "compress_latency_ms": {"p50": rp["compress"]["latency_ms_p50"],
                                              "p99": rp["compress"]["latency_ms_p99"]},
                      "decompress_latency_ms": {"p50": rp["decompress"]["latency_ms_p50"],
                                                "p99": rp["decompress"]["latency_ms_p99"]},
                      "aggregate_GBps": rp["aggregate_GBps"]})
        del sv, rq
    del small
    reqs = make_requests(count, rate, seed + 2)
    src = {r.rid: r.result_index for r in reqs}
    srv.batches = 0
    srv.peak = 0
    t0 = time.perf_counter()
    served = srv.run(reqs, keep_results=True)
    wall = time.perf_counter() - t0
    rep = serve.report(served, wall)
    # Each result against the same request alone (a sample of 40 of each kind).
    same, checked = True, 0
    for r in [q for q in served if q.kind == "compress"][:40]:
        c = szp.DeviceCodec(block_size=B, threads=T)
        out, size_t = c.compress(r.x, *r.mode)
        c.check()
        n = int(size_t.item())
        same = same and n == r.result_size and torch.equal(out[:n], r.result)
        checked += 1
    for r in [q for q in served if q.kind == "decompress"][:40]:
        st, size, n, idx = pre[src[r.rid]]
        c = szp.DeviceCodec(block_size=B, threads=T)
        d = c.decompress(st, size, n, index=idx)
        c.check()
        same = same and torch.equal(d.view(torch.int32), r.result.view(torch.int32))
        checked += 1
    log("serving leg: %d requests in %d batches" % (len(served), srv.batches))
    res = {
        "workload": "%d requests, Poisson arrivals at %g/s (seed %d), 70%% compress / 30%% "
                    "decompress, payloads log-uniform 16 KB-256 MB (pool of %d resident fields, "
                    "smooth or noise), REL in {1e-2,1e-3,1e-4}, B=%d, T=%d, budget %d GiB, "
                    "3 streams" % (count, rate, seed, pool_size, B, T, budget >> 30),
        "offered_seconds": round(reqs[-1].arrival, 4),
        "wall_seconds": round(wall, 4),
        "batches": srv.batches,
        "peak_admitted_bytes": int(srv.peak),
        "compress": rep["compress"],
        "decompress": rep["decompress"],
        "aggregate_GBps": rep["aggregate_GBps"],
        "results_same_alone_and_in_mix": bool(same),
        "results_checked": checked,
        "heavier_loads": loads,
    }
    del srv, served, reqs, pool, pre
    torch.cuda.empty_cache()
    return res


def other_paths_bench(torch, szp, x, rel=1e-4, B=128, T=16, steps=10, cpu_check=True):
    """The 512^3 field through the float64 path (the field widened to double)
    and through fastMode 2 (random access, float): compress with the side
    index, decompress with it and as a foreign stream; stream bytes and
    decompressed values against the reference on the same field and T."""
    import numpy as np
    n = x.numel()
    out = {}
    cases = (("double", x.double().contiguous(), 1), ("random_access", x, 2))
    for name, xv, mode in cases:
        es = xv.element_size()
        codec = szp.DeviceCodec(block_size=B, threads=T, fast_mode=mode)
        st, size_t, idx = codec.compress(xv, szp.REL, 0.0, rel, index=True)
        codec.check()
        C = int(size_t.item())
        dec = torch.empty_like(xv)
        for _ in range(2):
            codec.compress(xv, szp.REL, 0.0, rel, out=st, index=idx)
            codec.decompress(st, C, n, out=dec, index=idx)
            codec.decompress(st, C, n, out=dec)
        codec.check()
        comp_ms = _timed(torch, lambda: codec.compress(xv, szp.REL, 0.0, rel, out=st, index=idx), steps)
        dec_ms = _timed(torch, lambda: codec.decompress(st, C, n, out=dec, index=idx), steps)
        parse_ms = _timed(torch, lambda: codec.decompress(st, C, n, out=dec), steps)
        codec.check()
        r = {"workload": "512^3 k^-3 field (%s), %s, REL %g, B=%d, T=%d" % (
                 "float64" if es == 8 else "float32", "fastMode %d" % mode, rel, B, T),
             "compressed_bytes": C, "ratio": round(es * n / C, 4),
             "compress_GBps": round(es * n / (comp_ms * 1e-3) / 1e9, 1),
             "compress_hbm_frac": round((2 * es * n + C) / (comp_ms * 1e-3) / 8e12, 4),
             "decompress_GBps": round(es * n / (dec_ms * 1e-3) / 1e9, 1),
             "decompress_hbm_frac": round((C + es * n) / (dec_ms * 1e-3) / 8e12, 4),
             "decompress_foreign_stream_GBps": round(es * n / (parse_ms * 1e-3) / 1e9, 1)}
        if cpu_check:
            kind, _, _ = load_reference(T)
            sys.path.insert(0, os.path.join(ROOT, "tests"))
            import oracle_lib
            xh = xv.cpu().numpy()
            t0 = time.perf_counter()
            if kind == "reference":
                ref = oracle_lib.ref_compress(xh, 1, 0.0, rel, B, T, dtype=xh.dtype, fast=mode)
                t1 = time.perf_counter()
                ref_dec = oracle_lib.ref_decompress(ref, n, B, T, dtype=xh.dtype, fast=mode)
            else:
                ref = oracle_lib.compress_mode(mode, xh, 1, 0.0, rel, B, T, workers=T)
                t1 = time.perf_counter()
                ref_dec = oracle_lib.decompress_mode(mode, ref, n, B, T, dtype=xh.dtype, workers=T)
            t2 = time.perf_counter()
            got = st[:C].cpu().numpy()
            dh = dec.cpu().numpy()
            ut = np.uint64 if es == 8 else np.uint32
            r["stream_byte_identical"] = bool(ref.size == C and np.array_equal(ref, got))
            r["decompress_bit_identical"] = bool(np.array_equal(ref_dec.view(ut), dh.view(ut)))
            r["against"] = kind
            r["cpu_compress_GBps_%d_threads" % T] = round(es * n / (t1 - t0) / 1e9, 3)
            r["cpu_decompress_GBps_%d_threads" % T] = round(es * n / (t2 - t1) / 1e9, 3)
            del xh, ref, ref_dec, got, dh
        out[name] = r
        del codec, st, idx, dec
        log("%s leg: compress %.1f GB/s, decompress %.1f GB/s" % (name, r["compress_GBps"],
                                                                  r["decompress_GBps"]))
    torch.cuda.empty_cache()
    return out


def _blocks(n, B, T):
    """SZp blocks of a field: each chunk of n_t elements has ceil((n_t - 1) / B)."""
    tbs = n // T
    last = n - (T - 1) * tbs
    nb = lambda m: (m - 1 + B - 1) // B if m else 0
    return (T - 1) * nb(tbs) + nb(last)


def large_field_bench(torch, szp, B=64, T=16, rel=1e-4, seed=5, steps=3, cpu_check=True):
    """2048 x 2048 x 1024 float32 (2^32 elements, 16 GiB): k^-3 field plus 1%
    Gaussian noise, REL 1e-4, B=64, compressed and decompressed in one pass
    each (the prefix stage over ~67 M blocks)."""
    import numpy as np
    x = workloads.large_field(torch, seed)
    N = x.numel()
    log("large field: %d elements generated" % N)
    codec = szp.DeviceCodec(block_size=B, threads=T)
    out, size_t, idx = codec.compress(x, szp.REL, 0.0, rel, index=True)
    codec.check()
    C = int(size_t.item())
    dec = torch.empty_like(x)
    codec.decompress(out, C, N, out=dec, index=idx)
    codec.check()
    comp_ms = _timed(torch, lambda: codec.compress(x, szp.REL, 0.0, rel, out=out, index=idx), steps)
    dec_ms = _timed(torch, lambda: codec.decompress(out, C, N, out=dec, index=idx), steps)
    dec_parse_ms = _timed(torch, lambda: codec.decompress(out, C, N, out=dec), 1)
    codec.check()
    log("large field: timed")
    # eb honoured (rounding-aware bound, as for the 512^3 field), on the GPU in slices.
    eb = float(np.frombuffer(out[:4].cpu().numpy()[::-1].tobytes(), dtype=np.float32)[0])
    worst, strict = 0.0, 0
    ok = True
    for a in range(0, N, 1 << 28):
        xa, da = x[a:a + (1 << 28)].double(), dec[a:a + (1 << 28)].double()
        err = (xa - da).abs()
        worst = max(worst, float(err.max()))
        strict += int((err > eb).sum())
        ok = ok and bool((err <= eb + 2.0 ** -22 * (eb + da.abs())).all())
    res = {
        "workload": "2048x2048x1024 float32 (2^32 elements, 16 GiB) k^-3 field + 1%% Gaussian noise "
                    "in [-1,1] (seed %d), REL %g, B=%d, T=%d" % (seed, rel, B, T),
        "n_elements": N, "blocks": _blocks(N, B, T), "compressed_bytes": C,
        "ratio": round(4 * N / C, 4),
        "compress_GBps": round(4 * N / (comp_ms * 1e-3) / 1e9, 1),
        "compress_hbm_frac": round((8 * N + C) / (comp_ms * 1e-3) / 8e12, 4),
        "decompress_GBps": round(4 * N / (dec_ms * 1e-3) / 1e9, 1),
        "decompress_hbm_frac": round((C + 4 * N) / (dec_ms * 1e-3) / 8e12, 4),
        "decompress_foreign_stream_GBps": round(4 * N / (dec_parse_ms * 1e-3) / 1e9, 1),
        "compress_ms": round(comp_ms, 3), "decompress_ms": round(dec_ms, 3),
        "error_bound": {"eb": eb, "max_err_over_eb": worst / eb, "strict_over_eb": strict,
                        "within_rounding_bound": ok},
    }
    # Batched: the same field as the one job of a persistent launch (graph
    # replay), against the plain call's stream and values.
    from szp_amd import batch as sb
    cb = sb.CompressBatch([x], (szp.REL, 0.0, rel), block_size=B, threads=T, index=True)
    tune_c = sb.autotune(cb)
    cb.launch()
    cb.check()
    bsz = cb.byte_sizes()[0]
    gc = cb.graph()
    bc_ms = _timed(torch, gc.replay, steps)
    cb.check()
    bstream = cb.stream(0, bsz)
    db = sb.DecompressBatch([bstream], [bsz], [N], [cb.index_of(0)], block_size=B, threads=T)
    tune_d = sb.autotune(db)
    db.launch()
    db.check()
    gd = db.graph()
    bd_ms = _timed(torch, gd.replay, steps)
    db.check()
    res["batched"] = {
        "compress_GBps": round(4 * N / (bc_ms * 1e-3) / 1e9, 1),
        "compress_hbm_frac": round((8 * N + C) / (bc_ms * 1e-3) / 8e12, 4),
        "decompress_GBps": round(4 * N / (bd_ms * 1e-3) / 1e9, 1),
        "decompress_hbm_frac": round((C + 4 * N) / (bd_ms * 1e-3) / 8e12, 4),
        "stream_byte_identical_to_plain": bool(bsz == C and torch.equal(bstream, out[:C])),
        "decompress_bit_identical_to_plain": bool(torch.equal(db.outs[0].view(torch.int32),
                                                              dec.view(torch.int32))),
        "graph": "one hipGraph replay per launch (kernels only)",
        "autotune": {k: {"form": t["form"], "ticket": t["ticket"],
                         "grid": t["grid"] if t["form"] == "persistent" else "one per tile"}
                     for k, t in (("compress", tune_c), ("decompress", tune_d))},
    }
    del cb, db, gc, gd, bstream
    torch.cuda.empty_cache()
    log("large field: batched timed")
    if cpu_check:
        # The reference compiled from its sources (else the oracle) on the
        # host, T threads, on the same field: byte and bit parity.
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib
        xh = x.cpu().numpy()
        got = out[:C].cpu().numpy()
        if oracle_lib.reference() is not None:
            kind = "reference"
            t0 = time.perf_counter()
            want = oracle_lib.ref_compress(xh, szp.REL, 0.0, rel, B, T)
            tc = time.perf_counter() - t0
            same = bool(want.size == C and np.array_equal(want, got))
            del xh
            t0 = time.perf_counter()
            wdec = oracle_lib.ref_decompress(want, N, B, T)
            td = time.perf_counter() - t0
            res["decompress_bit_identical_to_reference"] = bool(
                np.array_equal(wdec.view(np.uint32), dec.cpu().numpy().view(np.uint32)))
            res["reference_decompress_seconds"] = round(td, 2)
            del wdec
        else:
            kind = "port"
            t0 = time.perf_counter()
            want = oracle_lib.compress(xh, szp.REL, 0.0, rel, B, T, workers=T)
            tc = time.perf_counter() - t0
            same = bool(want.size == C and np.array_equal(want, got))
            del xh
        res["cpu_check"] = {"kind": kind, "threads": T, "compress_seconds": round(tc, 2),
                            "stream_byte_identical": same}
        del want, got
    del x, out, dec, idx, codec
    torch.cuda.empty_cache()
    return res


def copy_calibration(torch, steps):
    """A plain copy of 512 MiB, 1 read + 1 write per byte: its FETCH_SIZE and
    WRITE_SIZE readings per dispatch give the counters' scale."""
    nbytes = 512 << 20
    src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    ms = _timed(torch, lambda: dst.copy_(src), steps)
    print(json.dumps({"copy_bytes": nbytes, "ms": round(ms, 5),
                      "GBps_read_plus_write": round(2 * nbytes / (ms * 1e-3) / 1e9, 1)}))


def traffic_for(kernel, workload):
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    e = t.get(workload, {}).get(kernel)
    return e.get("hbm_bytes_per_launch") if e else None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=512, help="field edge length")
    ap.add_argument("--rel", type=float, default=1e-4)
    ap.add_argument("--block", type=int, default=128)
    ap.add_argument("--threads", type=int, default=16, help="T, the reference's OpenMP team size")
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline and parity leg")
    ap.add_argument("--no-variants", action="store_true")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched (uneven-size) leg")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving leg")
    ap.add_argument("--serve-only", action="store_true", help="only the serving leg (JSON line)")
    ap.add_argument("--rate", type=float, default=2000.0, help="serving arrivals per second")
    ap.add_argument("--no-large", action="store_true", help="skip the 2^32-element field leg")
    ap.add_argument("--no-c1", action="store_true",
                    help="skip the 64 Mi sinusoid field leg (reference CPU path + GPU)")
    ap.add_argument("--no-other", action="store_true",
                    help="skip the float64 and fastMode 2 (random access) legs")
    ap.add_argument("--large-only", action="store_true",
                    help="only the 2^32-element field (16 GiB) leg (JSON line)")
    ap.add_argument("--batch-only", action="store_true", help="only the batched leg (JSON line)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only time a plain 512 MiB device-to-device copy (the known-size kernel "
                         "that scales the FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the codec has no multi-GPU path (a field's "
              "chunks would be split across GPUs, exchanging only their byte sizes).",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    import szp_amd as szp

    if args.copy_calibration:
        copy_calibration(torch, args.steps)
        return
    if args.batch_only:
        log("batched leg")
        print(json.dumps(batched_bench(torch, szp, T=args.threads)), flush=True)
        return
    if args.large_only:
        log("large field leg")
        print(json.dumps(large_field_bench(torch, szp, T=args.threads)), flush=True)
        return
    if args.serve_only:
        log("serving leg")
        print(json.dumps(serving_bench(torch, szp, rate=args.rate, T=args.threads)), flush=True)
        return

    n3, B, T, rel = args.n, args.block, args.threads, args.rel
    log("generating %d^3 k^-3 field" % n3)
    x = workloads.grf(torch, n3, args.seed)
    N = x.numel()
    codec = szp.DeviceCodec(block_size=B, threads=T)
    log("main: REL %g B=%d T=%d" % (rel, B, T))
    r = measure(torch, szp, codec, x, rel, args.steps, args.warmup, profile=True)
    C = r["nbytes"]
    stream_host = r["out"][:C].cpu().numpy()
    rep = eb_report(torch, x, r["dec"], stream_host)

    # Dominant kernel and its roofline (HBM-bound integer work; no MFMA).
    alg = {"k_compress": 4 * N + C, "k_decode_indexed": C + 4 * N, "k_decode": C + 4 * N,
           "k_parse": C, "k_range_partial": 4 * N}
    kern = {k: v[1] / v[0] for k, v in r["kernels"].items() if v[0]}  # timed steps: dominant only
    allk = {k: v[1] / v[0] for k, v in r["all_kernels"].items() if v[0]}  # untimed pass
    pkern = {k: v[1] / v[0] for k, v in r["parse_kernels"].items() if v[0]}
    dom = max((k for k in kern if k in alg), key=lambda k: kern[k])
    achieved = alg[dom] / (kern[dom] * 1e-3) / 1e9
    workload = "grf512_k3_rel%g_B%d_T%d" % (rel, B, T)
    roofline = dict(kernel=dom, bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBPS,
                    unit="GB/s", frac=round(achieved / HBM_PEAK_GBPS, 4),
                    traffic=traffic_for(dom, workload), alg_bytes_per_launch=alg[dom],
                    avg_launch_ms=round(kern[dom], 5),
                    timing="avg_launch_ms: HIP events on the launch stream around the dominant "
                           "kernel only, inside the timed steps; kernels_avg_ms: an untimed pass "
                           "of the same step with every launch bracketed",
                    kernels_avg_ms={k: round(v, 5) for k, v in sorted(allk.items())},
                    foreign_decompress_kernels_avg_ms={k: round(v, 5) for k, v in sorted(pkern.items())})

    comp_gbps = 4 * N / (r["comp_ms"] * 1e-3) / 1e9
    dec_gbps = 4 * N / (r["dec_ms"] * 1e-3) / 1e9
    value = 8 * N / (r["step_ms"] * 1e-3) / 1e9

    cpu = None
    parity = None
    if not args.no_cpu:
        xh = x.cpu().numpy()
        threads, src = cpu_threads()
        kind, rcomp_h, rdecomp_h = load_reference(threads)
        log("cpu %s: compress + decompress on %d threads (2 warm-ups, median of 10)"
            % (kind, threads))
        t = cpu_timing(rcomp_h, rdecomp_h, xh, rel, B)
        tc, td = t["compress_s"], t["decompress_s"]
        cpu = dict(value=round(8 * N / (tc + td) / 1e9, 3), unit="GB/s", cores=threads, kind=kind,
                   cores_from=src,
                   sample="the full %d^3 field, REL %g, B=%d, T=%d: 2 warm-ups, then the median "
                          "of 10 API-level calls (compress %.3f s, decompress %.3f s)"
                          % (n3, rel, B, threads, tc, td),
                   compress_GBps=round(4 * N / tc / 1e9, 3),
                   decompress_GBps=round(4 * N / td / 1e9, 3))
        # Parity at the GPU stream's team size T.
        kind, rcomp, rdecomp = load_reference(T)
        if threads == T:
            ref_stream, ref_dec = t["stream"], t["dec"]
        else:
            ref_stream = rcomp(xh, rel, B)
            ref_dec = rdecomp(ref_stream, N, B)
        gpu_dec = r["dec"].cpu().numpy()
        parity = dict(stream_byte_identical=bool(ref_stream.size == C and
                                                 np.array_equal(ref_stream, stream_host)),
                      decompress_bit_identical=bool(np.array_equal(ref_dec.view(np.uint32),
                                                                   gpu_dec.view(np.uint32))),
                      against=kind, T=T)

    variants = []
    if not args.no_variants:
        for vrel, vB in ((1e-2, 128), (1e-4, 32), (1e-2, 32)):
            log("variant: REL %g B=%d" % (vrel, vB))
            vc = szp.DeviceCodec(block_size=vB, threads=T)
            vr = measure(torch, szp, vc, x, vrel, max(5, args.steps // 2), 2, profile=False)
            v = dict(rel=vrel, B=vB, ratio=round(4 * N / vr["nbytes"], 4),
                     compress_GBps=round(4 * N / (vr["comp_ms"] * 1e-3) / 1e9, 1),
                     decompress_GBps=round(4 * N / (vr["dec_ms"] * 1e-3) / 1e9, 1),
                     decompress_foreign_stream_GBps=round(4 * N / (vr["dec_parse_ms"] * 1e-3) / 1e9, 1),
                     compress_hbm_frac=round((8 * N + vr["nbytes"]) / (vr["comp_ms"] * 1e-3) / 8e12, 4),
                     decompress_hbm_frac=round((vr["nbytes"] + 4 * N) / (vr["dec_ms"] * 1e-3) / 8e12, 4))
            if not args.no_cpu:
                vs = vr["out"][:vr["nbytes"]].cpu().numpy()
                rs = rcomp(xh, vrel, vB)
                v["stream_byte_identical"] = bool(rs.size == vs.size and np.array_equal(rs, vs))
            variants.append(v)
            del vc, vr

    c1 = None
    if not args.no_c1:
        log("C1: 64 Mi sinusoid field (the CPU-path configuration)")
        c1 = c1_bench(torch, szp, T=T, cpu=not args.no_cpu)

    other = None
    if not args.no_other:
        log("float64 and random-access legs")
        other = other_paths_bench(torch, szp, x, rel=rel, B=B, T=T, cpu_check=not args.no_cpu)

    batched = None
    if not args.no_batch:
        log("batched leg: 256 fields")
        del x
        torch.cuda.empty_cache()
        batched = batched_bench(torch, szp, T=T)
    served = None
    if not args.no_serve:
        log("serving leg")
        served = serving_bench(torch, szp, rate=args.rate, T=T)
    large = None
    if not args.no_large:
        log("large field leg")
        large = large_field_bench(torch, szp, T=T, cpu_check=not args.no_cpu)

    line = {
        "metric": "compress/decompress GB/s and %HBM peak, 1 GPU",
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["step_ms"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "512^3 float32 k^-3 Gaussian random field in [-1,1], SZp fastmode1, "
                               "REL %g, block %d, T=%d" % (rel, B, T),
                   "n_elements": N, "field_bytes": 4 * N, "error_bound_mode": "REL",
                   "rel": rel, "block_size": B, "threads_T": T},
        "compress_GBps": round(comp_gbps, 1),
        "decompress_GBps": round(dec_gbps, 1),
        "compress_hbm_frac": round((8 * N + C) / (r["comp_ms"] * 1e-3) / 8e12, 4),
        "decompress_hbm_frac": round((C + 4 * N) / (r["dec_ms"] * 1e-3) / 8e12, 4),
        "paths": {
            "step": "compress writing stream + side index, then decompress using the index",
            "side_index_bytes": r["index_bytes"],
            "compress_without_index_GBps": round(4 * N / (r["comp_plain_ms"] * 1e-3) / 1e9, 1),
            "decompress_foreign_stream_GBps": round(4 * N / (r["dec_parse_ms"] * 1e-3) / 1e9, 1),
            "decompress_foreign_stream_hbm_frac": round((C + 4 * N) / (r["dec_parse_ms"] * 1e-3) / 8e12, 4),
        },
        "compressed_bytes": C,
        "ratio": round(4 * N / C, 4),
        "error_bound": rep,
        "parity": parity,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "variants": variants,
        "c1": c1,
        "other_paths": other,
        "batched": batched,
        "serving": served,
        "large_field": large,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
