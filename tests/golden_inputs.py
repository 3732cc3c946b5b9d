"""Bit-reproducible float32 test inputs.

Every generator uses numpy's PCG64 integer draws and IEEE basic operations only
(no transcendental functions), so the same (kind, n, seed) gives the same bits
on any machine.  The golden hash fixtures rely on that.
"""
import numpy as np


def _walk(rng, n, step=1000, scale=2.0 ** -12):
    steps = rng.integers(-step, step + 1, size=n, dtype=np.int64)
    return (np.cumsum(steps).astype(np.float64) * scale).astype(np.float32)


def _waves(rng, n):
    i = np.arange(n, dtype=np.int64)
    acc = np.zeros(n, dtype=np.float64)
    for _ in range(4):
        period = int(rng.integers(1000, 200000))
        phase = int(rng.integers(0, period))
        amp = float(rng.integers(1, 1 << 10)) * 2.0 ** -10
        tri = np.abs(((i + phase) % period) - period // 2).astype(np.float64)
        acc += amp * (tri * (4.0 / period) - 1.0)
    noise = rng.integers(-64, 65, size=n, dtype=np.int64).astype(np.float64) * 2.0 ** -16
    return (acc + noise).astype(np.float32)


def make_input(kind, n, seed):
    rng = np.random.default_rng(seed)
    if kind == "walk":
        return _walk(rng, n)
    if kind == "waves":
        return _waves(rng, n)
    if kind == "uniform":
        return (rng.integers(-(1 << 23), 1 << 23, size=n, dtype=np.int64).astype(np.float64)
                * 2.0 ** -23).astype(np.float32)
    if kind == "zeroblocks":
        x = _walk(rng, n)
        nblk = (n + 127) // 128
        zero = rng.random(nblk) < 0.2
        for b in np.nonzero(zero)[0]:
            x[b * 128:(b + 1) * 128] = 0.0
        return x
    if kind == "constant":
        return np.full(n, 0.375, dtype=np.float32)
    if kind == "specials":
        x = _walk(rng, n) * np.float32(1000.0)
        for val, cnt in ((np.nan, 40), (np.inf, 15), (-np.inf, 15), (3e9, 30), (-3e9, 30),
                         (2.1e9, 30), (-2.1e9, 30), (2147483520.0, 10), (-2147483648.0, 10)):
            x[rng.integers(0, n, cnt)] = np.float32(val)
        return x
    if kind == "fewspecials":
        x = _walk(rng, n)
        for val in (np.nan, np.inf, -np.inf, 3e9, -3e9, 2147483520.0, -2147483648.0):
            x[rng.integers(0, n, 3)] = np.float32(val)
        return x
    if kind == "sparsewide":
        x = _walk(rng, n)
        hit = rng.integers(0, n, max(1, n // 100))
        x[hit] = rng.integers(-(1 << 31), 1 << 31, size=hit.size, dtype=np.int64).astype(np.float32)
        return x
    if kind == "tiny":
        return (_walk(rng, n).astype(np.float64) * 1e-33).astype(np.float32)
    if kind == "signedzero":
        x = np.zeros(n, dtype=np.float32)
        x[rng.random(n) < 0.5] = np.float32(-0.0)
        return x
    if kind == "posinf":
        return np.full(n, np.inf, dtype=np.float32)
    if kind == "nearint":
        # Values whose products x * fl32(1 / 0.001) lie within a few ulps of
        # an integer k, |k| from 1 to 2^31, on a walk: the quantizer's
        # truncation of the exact product is tested where a float product
        # would round onto the integer.
        x = _walk(rng, n)
        inv = np.float32(1.0) / np.float32(0.001)
        m = n // 2
        kk = np.floor(2.0 ** (31.0 * rng.integers(0, 1 << 20, size=m) / (1 << 20)))
        kk *= np.where(rng.integers(0, 2, size=m) == 1, -1.0, 1.0)
        near = (kk / np.float64(inv)).astype(np.float32)
        step = rng.integers(-3, 4, size=m)
        for j in range(3):
            near = np.where(step > j, np.nextafter(near, np.float32(np.inf)), near)
            near = np.where(step < -j, np.nextafter(near, np.float32(-np.inf)), near)
        x[rng.permutation(n)[:m]] = near
        return x
    if kind == "nanfirst":
        x = _walk(rng, n)
        x[0] = np.float32(np.nan)
        return x
    raise ValueError(kind)


def make_input_f64(kind, n, seed):
    """float64 inputs: the float32 generators widened, plus a fine integer
    perturbation (multiples of 2^-40) so that values use the full double
    mantissa and the FP64 product x * (1 / eb) rounds; "specials" adds
    doubles beyond the int32 range of q and NaN / Inf."""
    base = make_input(kind, n, seed).astype(np.float64)
    rng = np.random.default_rng(seed + 7919)
    if kind in ("signedzero", "constant", "nanfirst", "posinf"):
        return base
    fine = rng.integers(-(1 << 20), (1 << 20) + 1, size=n, dtype=np.int64).astype(np.float64)
    x = base + fine * 2.0 ** -40
    if kind == "specials":
        for val, cnt in ((1e300, 10), (-1e300, 10), (4294967296.5, 10), (2147483647.75, 10)):
            x[rng.integers(0, n, cnt)] = val
    return x


# (kind, n, seed, mode, abs_eb, rel, B, T): larger inputs whose reference
# streams are pinned by SHA-256 only.
HASH_CASES = [
    ("walk", 1 << 20, 101, 0, 0.05, 0.0, 128, 8),
    ("waves", (1 << 22) + 3, 102, 1, 0.0, 1e-4, 128, 8),
    ("waves", 1 << 22, 103, 1, 0.0, 1e-2, 32, 16),
    ("uniform", 1 << 21, 104, 0, 0.002, 0.0, 32, 16),
    ("zeroblocks", 3 * (1 << 20) + 7, 105, 1, 0.0, 1e-3, 64, 5),
    ("walk", (1 << 21) + 1, 106, 1, 0.0, 1e-3, 16, 1),
]
