"""Serving loop: requests that arrive while others run are admitted and
batched together; every served result equals the CPU oracle's (compressed
streams byte-identical, decompressed values bit-identical) and the same
request run alone; the HBM budget bounds what is admitted at once; requests
can be cancelled before launch and are refused when larger than the budget."""
import sys

import numpy as np
import pytest

import oracle_lib
from golden_inputs import make_input

pytestmark = pytest.mark.gpu

B, T = 128, 8


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def setup():
    import torch
    import szp_amd
    kinds = ["waves", "walk", "zeroblocks", "uniform", "sparsewide"]
    sizes = [4096, 100003, 777777, 2 ** 21, 30001, 65536, 1, 5000]
    host = [make_input(kinds[i % len(kinds)], n, 40 + i) for i, n in enumerate(sizes)]
    pool = [torch.from_numpy(x).cuda() for x in host]
    # Pre-compressed streams (REL 1e-3) and side indexes that decompress
    # requests refer to, with the oracle's decompression of each.
    pre = []
    for x, xh in zip(pool, host):
        c = szp_amd.DeviceCodec(block_size=B, threads=T)
        out, size_t, idx = c.compress(x, szp_amd.REL, 0.0, 1e-3, index=True)
        c.check()
        size = int(size_t.item())
        stream = out[:size].clone()
        want = oracle_lib.compress(xh, 1, 0.0, 1e-3, B, T)
        assert np.array_equal(stream.cpu().numpy(), want)
        pre.append((stream, size, x.numel(), idx.clone(), oracle_lib.decompress(want, xh.size, B, T)))
    return torch, szp_amd, host, pool, pre


def _oracle_stream(cache, host, i, eb):
    if (i, eb) not in cache:
        cache[(i, eb)] = oracle_lib.compress(host[i], 1, 0.0, eb, B, T)
    return cache[(i, eb)]


def _requests(szp_amd, serve, pool, pre, count, rate, seed, cancel_every=0):
    rng = np.random.default_rng(seed)
    ebs = [1e-2, 1e-3, 1e-4]
    reqs, src = [], {}
    t = 0.0
    for rid in range(count):
        t += rng.exponential(1.0 / rate) if rate else 0.0
        i = int(rng.integers(0, len(pool)))
        if rng.random() < 0.7:
            eb = ebs[int(rng.integers(0, 3))]
            r = serve.Request(rid, "compress", t, x=pool[i], mode=(szp_amd.REL, 0.0, eb))
        else:
            s, size, n, index, _ = pre[i]
            r = serve.Request(rid, "decompress", t, stream=s, nbytes=size, n=n, index=index)
        if cancel_every and rid % cancel_every == 1:
            r.cancel_at = t  # the client gives up as soon as it has asked
        reqs.append(r)
        src[rid] = i
    return reqs, src


def _check_against_oracle(torch, served, src, host, pre, cache):
    checked = 0
    for r in served:
        if r.state != "done":
            assert r.result is None, (r.rid, r.state)
            continue
        i = src[r.rid]
        if r.kind == "compress":
            got = r.result.cpu().numpy()
            want = _oracle_stream(cache, host, i, r.mode[2])
            assert got.size == want.size and np.array_equal(got, want), (r.rid, i, r.mode)
        else:
            got = r.result.cpu().numpy()
            assert np.array_equal(got.view(np.uint32), pre[i][4].view(np.uint32)), (r.rid, i)
        checked += 1
    return checked


def test_results_same_alone_in_the_mix_and_as_the_oracle(setup):
    torch, szp_amd, host, pool, pre = setup
    from szp_amd import serve
    reqs, src = _requests(szp_amd, serve, pool, pre, 40, 20000, 5)
    srv = serve.Server(block_size=B, threads=T, budget=1 << 30, streams=3)
    served = srv.run(reqs)
    say("served", len(served), "batches", srv.batches, "peak", srv.peak)
    assert len(served) == len(reqs)
    assert srv.batches < len(reqs), "no request shared a batch"
    cache = {}
    assert _check_against_oracle(torch, served, src, host, pre, cache) == len(reqs)
    # The same requests run alone (a single-field call each) give the same bytes.
    for r in served:
        if r.kind != "compress":
            continue
        c = szp_amd.DeviceCodec(block_size=B, threads=T)
        out, size_t = c.compress(r.x, szp_amd.REL, 0.0, r.mode[2])
        c.check()
        assert torch.equal(out[:int(size_t.item())], r.result), r.rid
    rep = serve.report(served, max(r.done for r in served))
    say(rep)
    assert rep["compress"]["requests"] + rep["decompress"]["requests"] == len(reqs)
    assert srv.used == 0


def test_budget_binds(setup):
    """Sixty requests arrive at once against a budget that holds one and a
    half of the largest: admission queues them, the admitted bytes never exceed
    the budget, and every result is still the oracle's."""
    torch, szp_amd, host, pool, pre = setup
    from szp_amd import serve
    reqs, src = _requests(szp_amd, serve, pool, pre, 60, 0, 9)
    srv = serve.Server(block_size=B, threads=T, budget=1, streams=3)
    needs = [srv._need(r) for r in reqs]
    offered = sum(needs)
    srv.budget = int(1.5 * max(needs))
    say("offered", offered, "budget", srv.budget)
    assert offered > 3 * srv.budget
    served = srv.run(reqs)
    say("served", len(served), "batches", srv.batches, "peak", srv.peak)
    assert srv.peak <= srv.budget
    assert srv.refused == 0 and srv.cancelled == 0
    assert srv.batches >= 3
    waited = [r for r in served if r.admitted > r.arrival + 1e-4]
    assert waited, "admission never queued a request"
    assert _check_against_oracle(torch, served, src, host, pre, {}) == len(reqs)
    assert srv.used == 0


def test_cancellation_and_refusal(setup):
    """Cancelled requests (before arrival via Server.cancel, at arrival via
    cancel_at) release their budget and produce nothing; a request larger
    than the budget is refused; the others are served as the oracle's."""
    torch, szp_amd, host, pool, pre = setup
    from szp_amd import serve
    reqs, src = _requests(szp_amd, serve, pool, pre, 24, 50000, 13, cancel_every=4)
    srv = serve.Server(block_size=B, threads=T, budget=1 << 30, streams=2)
    srv.cancel(reqs[2].rid)
    # Larger than any pool field; the budget admits everything but it.
    big = serve.Request(1000, "compress", reqs[-1].arrival,
                        x=torch.zeros(1 << 22, dtype=torch.float32, device="cuda"),
                        mode=(szp_amd.REL, 0.0, 1e-3))
    srv.budget = srv._need(big) - 1
    assert max(srv._need(r) for r in reqs) <= srv.budget
    served = srv.run(reqs + [big])
    states = {r.rid: r.state for r in served}
    assert len(served) == len(reqs) + 1
    assert states[1000] == "refused" and srv.refused == 1
    want_cancel = {r.rid for r in reqs if r.cancel_at is not None} | {reqs[2].rid}
    assert {k for k, v in states.items() if v == "cancelled"} == want_cancel
    assert srv.cancelled == len(want_cancel)
    n_done = _check_against_oracle(torch, served, src, host, pre, {})
    assert n_done == len(reqs) - len(want_cancel)
    assert srv.used == 0
    rep = serve.report(served, max(r.done for r in served))
    assert rep["cancelled"] == len(want_cancel) and rep["refused"] == 1
