"""Synthetic fields of the benchmark configurations, built on the GPU with
torch (seeded).  Shared by bench.py and the GPU tests, so that the tests check
exactly the workloads the benchmark measures.

  sinusoids(n, seed)     1-D sum of 8 sinusoids + N(0, 1e-3) noise (64 Mi elements
                         in the CPU-path configuration)
  grf(n, seed)           n^3 Gaussian random field, power spectrum k^-3, in [-1, 1]
  large_field(seed)      2048 x 2048 x 1024 k^-3 field + 1% Gaussian noise, in [-1, 1]
  field_1d(kind, n, gen) one field of the uneven-size mix (smooth / noise / zeroblocks)
  mix(count, lo, hi, seed)  the uneven-size mix
"""
import numpy as np


def sinusoids(torch, n, seed):
    """x[i] = sum_s a_s sin(2 pi f_s i + phi_s) + N(0, 1e-3), f_s log-uniform in
    [1e-6, 1e-3] cycles per element, a_s ~ U(0.1, 1), phi_s ~ U(0, 2 pi); the
    sum is formed in float64 and rounded to float32."""
    rng = np.random.default_rng(seed)
    f = 10.0 ** rng.uniform(-6.0, -3.0, 8)
    a = rng.uniform(0.1, 1.0, 8)
    ph = rng.uniform(0.0, 2.0 * np.pi, 8)
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    i = torch.arange(n, dtype=torch.float64, device="cuda")
    x = torch.zeros(n, dtype=torch.float64, device="cuda")
    for s in range(8):
        x += float(a[s]) * torch.sin(i * (2.0 * np.pi * float(f[s])) + float(ph[s]))
    del i
    x += 1e-3 * torch.randn(n, generator=g, dtype=torch.float64, device="cuda")
    x = x.float().contiguous()
    torch.cuda.synchronize()
    return x


def grf(torch, n, seed):
    """n^3 Gaussian random field with power spectrum k^-3 (white noise filtered
    by |k|^-1.5, zero mean), scaled to [-1, 1], flattened in C order."""
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    w = torch.randn((n, n, n), generator=g, device="cuda", dtype=torch.float32)
    F = torch.fft.rfftn(w)
    del w
    kx = torch.fft.fftfreq(n, device="cuda").view(n, 1, 1)
    ky = torch.fft.fftfreq(n, device="cuda").view(1, n, 1)
    kz = torch.fft.rfftfreq(n, device="cuda").view(1, 1, n // 2 + 1)
    k = torch.sqrt(kx * kx + ky * ky + kz * kz)
    k[0, 0, 0] = 1.0
    F *= k.pow(-1.5)
    F[0, 0, 0] = 0
    del k
    x = torch.fft.irfftn(F, s=(n, n, n))
    del F
    mn, mx = x.min(), x.max()
    x = ((x - mn) / (mx - mn) * 2 - 1).contiguous().view(-1)
    torch.cuda.synchronize()
    return x


def large_field(torch, seed, shape=(2048, 2048, 1024)):
    """2048 x 2048 x 1024 float32 (2^32 elements, 16 GiB): k^-3 field plus 1%
    Gaussian noise (relative to the field's standard deviation), in [-1, 1]."""
    nx, ny, nz = shape
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    w = torch.randn((nx, ny, nz), generator=g, device="cuda", dtype=torch.float32)
    F = torch.fft.rfftn(w)
    del w
    kx = torch.fft.fftfreq(nx, device="cuda").view(nx, 1, 1)
    ky = torch.fft.fftfreq(ny, device="cuda").view(1, ny, 1)
    kz = torch.fft.rfftfreq(nz, device="cuda").view(1, 1, nz // 2 + 1)
    F *= (kx * kx + ky * ky + kz * kz).clamp_min(1e-12).pow(-0.75)
    F[0, 0, 0] = 0
    x = torch.fft.irfftn(F, s=(nx, ny, nz)).view(-1)
    del F
    torch.cuda.empty_cache()
    x = (x - x.mean()) / x.std()
    x += 0.01 * torch.randn(x.numel(), generator=g, device="cuda")
    mn, mx = x.min(), x.max()
    x = ((x - mn) / (mx - mn) * 2 - 1).contiguous()
    torch.cuda.synchronize()
    return x


def field_1d(torch, kind, n, gen):
    """One field of the uneven-size mix: 'smooth' (1-D k^-3 spectrum, [-1, 1]),
    'noise' (U(-1, 1)), or 'zeroblocks' (smooth with 20% of the 128-element
    index-aligned blocks set to 0)."""
    if kind == "noise":
        return torch.rand(n, generator=gen, device="cuda") * 2 - 1
    m = 1 << max(int(n - 1).bit_length(), 4)  # power-of-two FFT length (few distinct plans)
    w = torch.randn(m, generator=gen, device="cuda")
    F = torch.fft.rfft(w)
    k = torch.fft.rfftfreq(m, device="cuda")
    k[0] = 1.0
    F *= k.pow(-1.5)
    F[0] = 0
    x = torch.fft.irfft(F, n=m)[:n]
    mn, mx = x.min(), x.max()
    x = ((x - mn) / (mx - mn) * 2 - 1).contiguous()
    if kind == "zeroblocks":
        nblk = (n + 127) // 128
        z = torch.rand(nblk, generator=gen, device="cuda") < 0.2
        m = z.repeat_interleave(128)[:n]
        x = torch.where(m, torch.zeros_like(x), x)
    return x.contiguous()


def mix(torch, count, lo_bytes, hi_bytes, seed):
    """`count` fields with sizes log-uniform in [lo, hi] bytes, kinds drawn
    from smooth / noise / zeroblocks (seeded)."""
    rng = np.random.default_rng(seed)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed)
    sizes = np.exp(rng.uniform(np.log(lo_bytes), np.log(hi_bytes), count)).astype(np.int64) // 4
    kinds = ["smooth", "noise", "zeroblocks"]
    xs, ks = [], []
    for i, n in enumerate(sizes):
        kind = kinds[rng.integers(0, 3)]
        xs.append(field_1d(torch, kind, int(max(n, 1)), gen))
        ks.append(kind)
    torch.cuda.synchronize()
    return xs, ks
