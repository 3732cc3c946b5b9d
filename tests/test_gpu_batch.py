"""GPU parity of the batched forms: a batch of fields of mixed sizes, kinds
and error modes compresses to the oracle's bytes for every field, decodes
bit-identically, and a captured graph replays to the same result."""
import sys

import numpy as np
import pytest

import oracle_lib
from golden_inputs import make_input

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


# kind, n, mode, abs_eb, rel
FIELDS = [
    ("waves", 300001, 1, 0.0, 1e-3),
    ("walk", 5000, 0, 0.01, 0.0),
    ("zeroblocks", 1 << 18, 1, 0.0, 1e-3),
    ("uniform", 70001, 0, 0.002, 0.0),
    ("constant", 4096, 1, 0.0, 1e-3),
    ("specials", 100000, 0, 0.5, 0.0),
    ("waves", 1, 0, 0.01, 0.0),
    ("walk", 17, 1, 0.0, 1e-2),
    ("sparsewide", 123457, 0, 1.0, 0.0),
    ("waves", (1 << 20) + 3, 1, 0.0, 1e-4),
    ("signedzero", 9000, 1, 0.0, 1e-3),
    ("nanfirst", 30000, 1, 0.0, 1e-3),
]


@pytest.fixture(scope="module")
def env():
    import torch
    assert torch.cuda.is_available()
    import szp_amd
    from szp_amd import batch
    return torch, szp_amd, batch


def _inputs(torch):
    xs_h = [make_input(k, n, 300 + i) for i, (k, n, *_r) in enumerate(FIELDS)]
    xs = [torch.from_numpy(x).cuda() for x in xs_h]
    modes = [(m, a, r) for (_k, _n, m, a, r) in FIELDS]
    return xs_h, xs, modes


@pytest.mark.parametrize("B,T", [(128, 16), (32, 3), (24, 5), (2048, 2), (9, 4)])
def test_batch_compress_matches_oracle(env, B, T):
    torch, szp, batch = env
    xs_h, xs, modes = _inputs(torch)
    say("batch compress B=%d T=%d fields=%d" % (B, T, len(xs)))
    cb = batch.CompressBatch(xs, modes, block_size=B, threads=T, index=True)
    cb.launch()
    cb.check()
    sizes = cb.byte_sizes()
    for k, x in enumerate(xs_h):
        m, a, r = modes[k]
        want = oracle_lib.compress(x, m, a, r, B, T)
        assert sizes[k] == want.size, (k, sizes[k], want.size)
        got = cb.stream(k, sizes[k]).cpu().numpy()
        bad = np.nonzero(got != want)[0]
        assert bad.size == 0, (k, bad[:8])
    # Decompress the whole batch through the side indexes.
    streams = [cb.stream(k, sizes[k]) for k in range(len(xs))]
    db = batch.DecompressBatch(streams, sizes, [x.size for x in xs_h],
                               [cb.index_of(k) for k in range(len(xs))], block_size=B, threads=T)
    db.launch()
    db.check()
    for k, x in enumerate(xs_h):
        want = oracle_lib.decompress(streams[k].cpu().numpy(), x.size, B, T)
        assert np.array_equal(db.outs[k].cpu().numpy().view(np.uint32), want.view(np.uint32)), k


def test_batch_launch_shapes_agree(env):
    """Every ticket size and grid gives the same bytes."""
    torch, szp, batch = env
    xs_h, xs, modes = _inputs(torch)
    cb = batch.CompressBatch(xs, modes, block_size=64, threads=8, index=False)
    cb.launch(ticket=1, grid=0)
    cb.check()
    ref_sizes = cb.byte_sizes()
    ref = [cb.stream(k, ref_sizes[k]).cpu().numpy().copy() for k in range(len(xs))]
    for ticket, grid in ((2, 7), (8, 64), (32, 0), (3, 1)):
        say("ticket", ticket, "grid", grid)
        cb.out.zero_()
        cb.launch(ticket=ticket, grid=grid)
        cb.check()
        assert cb.byte_sizes() == ref_sizes
        for k in range(len(xs)):
            assert np.array_equal(cb.stream(k, ref_sizes[k]).cpu().numpy(), ref[k]), (ticket, grid, k)


@pytest.mark.parametrize("B,T", [(128, 16), (32, 3), (2048, 2), (9, 4)])
def test_batch_compress_per_tile_form(env, B, T):
    """One workgroup per tile (GRID_PER_TILE) gives the oracle's bytes, with
    and without the side index, twice in a row."""
    torch, szp, batch = env
    xs_h, xs, modes = _inputs(torch)
    want = [oracle_lib.compress(x, *modes[k], B, T) for k, x in enumerate(xs_h)]
    for index in (True, False):
        say("per-tile form B=%d T=%d index=%s" % (B, T, index))
        cb = batch.CompressBatch(xs, modes, block_size=B, threads=T, index=index)
        for rep in range(2):
            cb.out.zero_()
            cb.launch(grid=batch.GRID_PER_TILE)
            cb.check()
            sizes = cb.byte_sizes()
            for k in range(len(xs)):
                assert sizes[k] == want[k].size, (index, rep, k)
                assert np.array_equal(cb.stream(k, sizes[k]).cpu().numpy(), want[k]), (index, rep, k)


def test_batch_autotune_shapes(env):
    """The autotuner tries both compress forms and several grids from the
    resident grid, and its choice gives the same bytes."""
    torch, szp, batch = env
    xs_h, xs, modes = _inputs(torch)
    cb = batch.CompressBatch(xs, modes, block_size=128, threads=16, index=True)
    cb.launch()
    cb.check()
    sizes = cb.byte_sizes()
    ref = [cb.stream(k, sizes[k]).cpu().numpy().copy() for k in range(len(xs))]
    tune = batch.autotune(cb, reps=1)
    say("autotune", {k: v for k, v in tune.items() if k != "trials"})
    forms = {t["form"] for t in tune["trials"]}
    assert forms == {"persistent", "per_tile"}
    assert tune["resident_grid"] >= 256
    assert len({t["grid"] for t in tune["trials"]}) >= 4
    cb.out.zero_()
    cb.launch()  # the chosen shape
    cb.check()
    assert cb.byte_sizes() == sizes
    for k in range(len(xs)):
        assert np.array_equal(cb.stream(k, sizes[k]).cpu().numpy(), ref[k]), k
    db = batch.DecompressBatch([cb.stream(k, sizes[k]) for k in range(len(xs))], sizes,
                               [x.size for x in xs_h], [cb.index_of(k) for k in range(len(xs))],
                               block_size=128, threads=16)
    tuned = batch.autotune(db, reps=1)
    assert {t["form"] for t in tuned["trials"]} == {"persistent"}
    assert len({t["ticket"] for t in tuned["trials"]}) >= 4
    db.launch()
    db.check()
    for k, x in enumerate(xs_h):
        want = oracle_lib.decompress(cb.stream(k, sizes[k]).cpu().numpy(), x.size, 128, 16)
        assert np.array_equal(db.outs[k].cpu().numpy().view(np.uint32), want.view(np.uint32)), k


def test_batch_graph_replays(env):
    """A captured batch replays to the eager result, with eager work between
    replays (the graph records kernels only)."""
    torch, szp, batch = env
    xs_h, xs, modes = _inputs(torch)
    cb = batch.CompressBatch(xs, modes, block_size=128, threads=16, index=True)
    cb.launch()
    cb.check()
    sizes = cb.byte_sizes()
    ref = [cb.stream(k, sizes[k]).cpu().numpy().copy() for k in range(len(xs))]
    g = cb.graph()
    for rep in range(3):
        say("replay", rep)
        cb.out.zero_()
        cb.sizes.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        cb.check()
        assert cb.byte_sizes() == sizes
        for k in range(len(xs)):
            assert np.array_equal(cb.stream(k, sizes[k]).cpu().numpy(), ref[k]), (rep, k)
        # eager work between replays
        y = torch.randn(1 << 20, device="cuda").sum()
        torch.cuda.synchronize()
        del y
