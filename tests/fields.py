"""Deterministic test fields built from integer arithmetic only.

Values are dyadic rationals (integers divided by powers of two), so every
machine produces identical bits without libm; the golden fixtures under
tests/golden/ were written from exactly these arrays.
"""
import numpy as np

PATTERNS = ("smooth", "noise", "constant", "checker", "outliers", "zero", "ramp")


def _lcg(n, seed):
    """32-bit LCG stream (x = x*1664525 + 1013904223) as int64 values of int32."""
    out = np.empty(n, np.int64)
    x = np.uint64(seed & 0xFFFFFFFF)
    # Jump-ahead by blocks keeps this vectorised: x_{i} = A^i x_0 + C (A^i - 1)/(A - 1).
    a, c, m = 1664525, 1013904223, 1 << 32
    block = 4096
    mult = np.empty(block, np.uint64)
    add = np.empty(block, np.uint64)
    am, ac = 1, 0
    for i in range(block):
        am = (am * a) % m
        ac = (ac * a + c) % m
        mult[i], add[i] = am, ac
    for start in range(0, n, block):
        k = min(block, n - start)
        vals = (mult[:k] * x + add[:k]) & np.uint64(0xFFFFFFFF)
        out[start:start + k] = vals.astype(np.int64)
        x = vals[k - 1]
    return np.where(out >= 1 << 31, out - (1 << 32), out)


def field(shape, pattern, dtype=np.float32, seed=1):
    """(nz, ny, nx) field of dtype; all values exactly representable in float32."""
    nz, ny, nx = shape
    n = nz * ny * nx
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    if pattern == "smooth":
        # piecewise-quadratic bumps on a 1/256 grid plus small dyadic noise
        v = ((x * 7 + y * 3) % 64 - 32) * (z % 16 + 4) / 256.0
        v = v + ((x * x + 2 * y * y + 3 * z) % 61) / 64.0
        v = v + (_lcg(n, seed).reshape(shape) >> 20) / 65536.0
    elif pattern == "noise":
        v = (_lcg(n, seed).reshape(shape) >> 8) / float(1 << 22)
    elif pattern == "constant":
        v = np.full(shape, 1.25)
    elif pattern == "checker":
        v = np.where((x + y + z) & 1, -4.0, 4.0)
        v.flat[0] = 32.0
        v.flat[-1] = -32.0
    elif pattern == "outliers":
        v = (_lcg(n, seed).reshape(shape) >> 12) / 16.0  # |v| up to 2^15
        v = np.where((x + 2 * y + 3 * z) % 11 == 0, v * 64.0, v)
    elif pattern == "zero":
        v = np.zeros(shape)
    elif pattern == "ramp":
        v = (x - nx / 2.0) / 8.0 + (y % 5) / 4.0 - (z % 3) / 2.0
    else:
        raise ValueError(pattern)
    return np.ascontiguousarray(v, dtype=dtype)
