"""Serving loop: requests that arrive while others run are admitted and
batched together, and every result equals the same request run alone
(compressed streams byte-identical, decompressed values bit-identical)."""
import sys

import numpy as np
import pytest

from golden_inputs import make_input

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_results_same_alone_and_in_the_mix():
    import torch
    import szp_amd
    from szp_amd import serve
    B, T = 128, 8
    rng = np.random.default_rng(5)
    kinds = ["waves", "walk", "zeroblocks", "uniform", "sparsewide"]
    sizes = [4096, 100003, 777777, 2 ** 21, 30001, 65536, 1, 5000]
    pool = [torch.from_numpy(make_input(kinds[i % len(kinds)], n, 40 + i)).cuda()
            for i, n in enumerate(sizes)]
    ebs = [1e-2, 1e-3, 1e-4]
    # Reference ("alone") results and the pre-compressed streams that
    # decompress requests refer to.
    alone = {}
    pre = []
    for i, x in enumerate(pool):
        c = szp_amd.DeviceCodec(block_size=B, threads=T)
        out, size_t, idx = c.compress(x, szp_amd.REL, 0.0, 1e-3, index=True)
        c.check()
        size = int(size_t.item())
        stream = out[:size].clone()
        index = idx.clone()
        dec = c.decompress(stream, size, x.numel(), index=index)
        c.check()
        pre.append((stream, size, x.numel(), index, dec.clone()))
    reqs = []
    t = 0.0
    for rid in range(40):
        t += rng.exponential(1.0 / 20000)  # bursty enough that batches mix
        if rng.random() < 0.7:
            i = int(rng.integers(0, len(pool)))
            eb = ebs[int(rng.integers(0, 3))]
            reqs.append(serve.Request(rid, "compress", t, x=pool[i], mode=(szp_amd.REL, 0.0, eb)))
            reqs[-1].n = i
            if (i, eb) not in alone:
                c = szp_amd.DeviceCodec(block_size=B, threads=T)
                out, size_t = c.compress(pool[i], szp_amd.REL, 0.0, eb)
                c.check()
                alone[(i, eb)] = out[:int(size_t.item())].cpu().numpy().copy()
        else:
            i = int(rng.integers(0, len(pool)))
            s, size, n, index, _ = pre[i]
            r = serve.Request(rid, "decompress", t, stream=s, nbytes=size, n=n, index=index)
            r.result_index = i
            reqs.append(r)
    src = {r.rid: (r.n if r.kind == "compress" else r.result_index) for r in reqs}
    srv = serve.Server(block_size=B, threads=T, budget=1 << 30, streams=3)
    served = srv.run(reqs)
    say("served", len(served), "batches", srv.batches)
    assert len(served) == len(reqs)
    assert srv.batches < len(reqs), "no request shared a batch"
    for r in served:
        i = src[r.rid]
        if r.kind == "compress":
            got = r.result.cpu().numpy()
            want = alone[(i, r.mode[2])]
            assert got.size == want.size and np.array_equal(got, want), (r.rid, i, r.mode)
        else:
            assert torch.equal(r.result.view(torch.int32), pre[i][4].view(torch.int32)), (r.rid, i)
    rep = serve.report(served, max(r.done for r in served))
    say(rep)
    assert rep["compress"]["requests"] + rep["decompress"]["requests"] == len(reqs)
