"""Seeded synthetic fields for benchmarks (generated on the GPU with torch.fft).

grf(): Gaussian random field with isotropic power spectrum P(k) ~ k**slope on a
periodic grid (white noise -> multiply by sqrt(P) in Fourier space -> inverse
real FFT -> zero mean, unit variance), then `offset` added.
"""
import math

import torch


def grf(shape, slope=-5.0 / 3.0, seed=20261202, offset=0.0, dtype=torch.float32, device="cuda"):
    nz, ny, nx = shape
    g = torch.Generator(device=device).manual_seed(int(seed))
    white = torch.randn((nz, ny, nx), generator=g, device=device, dtype=torch.float32)
    spec = torch.fft.rfftn(white)
    del white
    kz = torch.fft.fftfreq(nz, device=device).view(nz, 1, 1)
    ky = torch.fft.fftfreq(ny, device=device).view(1, ny, 1)
    kx = torch.fft.rfftfreq(nx, device=device).view(1, 1, nx // 2 + 1)
    k2 = kz * kz + ky * ky + kx * kx
    amp = torch.where(k2 > 0, k2.clamp_min(1e-30) ** (slope / 4.0), torch.zeros_like(k2))
    spec *= amp
    del amp, k2
    f = torch.fft.irfftn(spec, s=(nz, ny, nx)).to(torch.float64)
    del spec
    f = (f - f.mean()) / f.std()
    return (f + offset).to(dtype).contiguous()


def sinusoids(shape, count=8, seed=20261202, noise=0.01, dtype=torch.float32, device="cuda"):
    """Sum of `count` random 3-D sinusoids (integer frequencies 1..8, amplitudes
    U[0.1, 1], phases U[0, 2pi)) plus N(0, noise^2) noise."""
    nz, ny, nx = shape
    g = torch.Generator(device="cpu").manual_seed(int(seed))
    z = torch.arange(nz, device=device, dtype=torch.float64).view(nz, 1, 1) / nz
    y = torch.arange(ny, device=device, dtype=torch.float64).view(1, ny, 1) / ny
    x = torch.arange(nx, device=device, dtype=torch.float64).view(1, 1, nx) / nx
    f = torch.zeros((nz, ny, nx), device=device, dtype=torch.float64)
    for _ in range(count):
        fz, fy, fx = (int(v) for v in torch.randint(1, 9, (3,), generator=g))
        a = 0.1 + 0.9 * float(torch.rand((), generator=g))
        ph = 2 * math.pi * float(torch.rand((), generator=g))
        f += a * torch.sin(2 * math.pi * (fz * z + fy * y + fx * x) + ph)
    gn = torch.Generator(device=device).manual_seed(int(seed) + 1)
    f += noise * torch.randn((nz, ny, nx), generator=gn, device=device, dtype=torch.float64)
    return f.to(dtype).contiguous()


def fronts(shape, seed=20261202, count=3, amplitude=40.0, noise=3.0, dtype=torch.float64, device="cuda"):
    """Smooth field with sharp fronts: sum of `count` seeded tanh fronts
    amplitude * tanh((n_i . r - c_i) / w_i) (unit normals n_i, offsets c_i
    inside the box, widths w_i of 2-8 samples) plus `noise` times a unit
    variance k^-3 Gaussian random field; range about [-50, 50]."""
    nz, ny, nx = shape
    g = torch.Generator(device="cpu").manual_seed(int(seed))
    z = torch.arange(nz, device=device, dtype=torch.float64).view(nz, 1, 1)
    y = torch.arange(ny, device=device, dtype=torch.float64).view(1, ny, 1)
    x = torch.arange(nx, device=device, dtype=torch.float64).view(1, 1, nx)
    f = torch.zeros((nz, ny, nx), device=device, dtype=torch.float64)
    for _ in range(count):
        v = torch.randn(3, generator=g, dtype=torch.float64)
        v = v / v.norm()
        c = float(torch.rand((), generator=g, dtype=torch.float64)) * float(v[0] * nx + v[1] * ny + v[2] * nz)
        w = 2.0 + 6.0 * float(torch.rand((), generator=g, dtype=torch.float64))
        f += amplitude / count * torch.tanh((float(v[0]) * x + float(v[1]) * y + float(v[2]) * z - c) / w)
    f += noise * grf(shape, slope=-3.0, seed=int(seed) + 1, device=device).to(torch.float64)
    return f.to(dtype).contiguous()
