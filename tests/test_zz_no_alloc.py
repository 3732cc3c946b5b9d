"""szp_amd_compress_dev_noalloc (its workspace is a static device array): the
same stream and size as szp_amd_compress_dev on two calls in a row, the second
with a smaller field (the array keeps the first call's contents), both equal
to the oracle's; a field whose workspace exceeds the array is refused before
any launch."""
import ctypes
import sys

import numpy as np
import pytest

import oracle_lib
from golden_inputs import make_input

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def env():
    import torch
    assert torch.cuda.is_available(), "no HIP device visible"
    import szp_amd
    L = szp_amd.lib
    u64, i32, vp, dbl = ctypes.c_uint64, ctypes.c_int, ctypes.c_void_p, ctypes.c_double
    L.szp_amd_compress_dev_noalloc.restype = i32
    L.szp_amd_compress_dev_noalloc.argtypes = [vp, u64, i32, dbl, dbl, i32, i32, vp, u64, vp, vp]
    return torch, szp_amd, L


def _noalloc(torch, szp, L, xd, mode, ab, rel, B, T):
    n = xd.numel()
    cap = szp.compress_bound(n, B, T)
    out = torch.zeros(cap + 64, dtype=torch.uint8, device="cuda")
    size = torch.zeros(1, dtype=torch.int64, device="cuda")
    s = torch.cuda.current_stream()
    rc = L.szp_amd_compress_dev_noalloc(xd.data_ptr(), n, mode, ab, rel, B, T, out.data_ptr(),
                                        cap, size.data_ptr(), ctypes.c_void_p(s.cuda_stream))
    assert rc == 0, (rc, szp._last_error())
    torch.cuda.synchronize()
    return out, int(size.item())


@pytest.mark.parametrize("kind,n1,n2,B,T,mode,ab,rel", [
    ("waves", 3000017, 70001, 128, 16, 1, 0.0, 1e-4),
    ("walk", 1 << 20, 5000, 32, 3, 0, 0.01, 0.0),
])
def test_noalloc_matches_compress_dev(env, kind, n1, n2, B, T, mode, ab, rel):
    torch, szp, L = env
    for n in (n1, n2):
        x = make_input(kind, n, 17 + n)
        say("noalloc case", kind, n, B, T)
        xd = torch.from_numpy(x).cuda()
        codec = szp.DeviceCodec(block_size=B, threads=T)
        ref, ref_size_t = codec.compress(xd, mode, ab, rel)
        codec.check()
        ref_size = int(ref_size_t.item())
        out, size = _noalloc(torch, szp, L, xd, mode, ab, rel, B, T)
        assert size == ref_size, (size, ref_size)
        assert torch.equal(out[:size], ref[:size])
        want = oracle_lib.compress(x, mode, ab, rel, B, T)
        assert size == want.size and np.array_equal(out[:size].cpu().numpy(), want)


def test_noalloc_refuses_fields_beyond_the_scratch(env):
    torch, szp, L = env
    xd = torch.zeros(16, dtype=torch.float32, device="cuda")
    n = 2_000_000_000  # B = 1: a tile per 1024 elements, 15.6 MB of status words
    assert L.szp_amd_compress_workspace_size(n, 1, 4) > (8 << 20)
    size = torch.zeros(1, dtype=torch.int64, device="cuda")
    rc = L.szp_amd_compress_dev_noalloc(xd.data_ptr(), n, 0, 0.1, 0.0, 1, 4, xd.data_ptr(), 64,
                                        size.data_ptr(), None)
    assert rc == 2  # SZP_AMD_ERR_CAPACITY
    assert b"scratch" in L.szp_amd_last_error()
    torch.cuda.synchronize()
    assert int(size.item()) == 0  # nothing was launched
