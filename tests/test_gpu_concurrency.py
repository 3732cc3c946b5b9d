"""Single-field calls issued on two streams at once.  The look-back kernels
(compress, and decompress of a stream without the side index) each need the
lowest unfinished tile of their own launch to be resident; the library orders
such launches across streams, so concurrent calls complete (no device-wait
timeout) and every result equals the oracle's."""
import sys

import numpy as np
import pytest

import oracle_lib
from golden_inputs import make_input

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_two_streams_compress_and_decompress_at_once():
    import torch
    import szp_amd
    B, T = 128, 16
    n = (1 << 24) + 77
    xs = [make_input("waves", n, 101), make_input("walk", n, 102)]
    want = [oracle_lib.compress(x, 1, 0.0, 1e-4, B, T, workers=T) for x in xs]
    want_dec = [oracle_lib.decompress(w, n, B, T, workers=T) for w in want]
    xd = [torch.from_numpy(x).cuda() for x in xs]
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    codecs = [szp_amd.DeviceCodec(block_size=B, threads=T) for _ in range(2)]
    decs = [torch.empty(n, dtype=torch.float32, device="cuda") for _ in range(2)]
    torch.cuda.synchronize()
    for rep in range(3):
        say("round", rep)
        outs = []
        for k in range(2):
            with torch.cuda.stream(streams[k]):
                outs.append(codecs[k].compress(xd[k], szp_amd.REL, 0.0, 1e-4, stream=streams[k]))
        for k in range(2):
            out, _ = outs[k]
            with torch.cuda.stream(streams[k]):
                codecs[k].decompress(out, want[k].size, n, out=decs[k], stream=streams[k])
        for k in range(2):
            codecs[k].check(stream=streams[k])
            out, size_t = outs[k]
            size = int(size_t.item())
            assert size == want[k].size
            assert np.array_equal(out[:size].cpu().numpy(), want[k]), (rep, k)
            assert np.array_equal(decs[k].cpu().numpy().view(np.uint32),
                                  want_dec[k].view(np.uint32)), (rep, k)
