"""Synthetic float32 test fields (C order, DIM0 slowest).

Gaussian random fields are white noise shaped in Fourier space so that the
power spectrum is proportional to |K|^exponent, with
K^2 = (s0 k0)^2 + (s1 k1)^2 + (s2 k2)^2 (stretch factors s stretch the
correlation length along each axis).  Generation runs with torch on any
device (CPU for small fields, the GPU for 512^3 and larger).
"""
from __future__ import annotations

import math

import numpy as np
import torch


def grf(shape, exponent, seed, stretch=(1.0, 1.0, 1.0), device="cpu", dtype=torch.float64):
    """Zero-mean Gaussian random field with power spectrum ~ |K|^exponent."""
    g = torch.Generator(device="cpu").manual_seed(int(seed))
    white = torch.randn(shape, generator=g, dtype=dtype).to(device)
    spec = torch.fft.rfftn(white)
    del white
    k0 = torch.fft.fftfreq(shape[0], device=device, dtype=dtype) * stretch[0]
    k1 = torch.fft.fftfreq(shape[1], device=device, dtype=dtype) * stretch[1]
    k2 = torch.fft.rfftfreq(shape[2], device=device, dtype=dtype) * stretch[2]
    k2sum = k0[:, None, None] ** 2 + k1[None, :, None] ** 2 + k2[None, None, :] ** 2
    k2sum[0, 0, 0] = 1.0
    spec *= k2sum ** (exponent / 4.0)  # amplitude = |K|^(exponent/2)
    spec[0, 0, 0] = 0
    del k2sum
    out = torch.fft.irfftn(spec, s=shape)
    return out


def rescale(f, lo, hi):
    mn, mx = f.min(), f.max()
    return (f - mn) / (mx - mn) * (hi - lo) + lo


def config1_field(shape=(100, 500, 500), seed=12345, device="cpu"):
    """|k|^-11/3 GRF rescaled to [-80, 40] plus N(0, (1e-4 * 120)^2) noise."""
    f = rescale(grf(shape, -11.0 / 3.0, seed, device=device), -80.0, 40.0)
    g = torch.Generator(device="cpu").manual_seed(seed + 1)
    f += (torch.randn(shape, generator=g, dtype=f.dtype) * (1e-4 * 120.0)).to(f.device)
    return f.to(torch.float32)


def config2_field(shape=(512, 512, 512), seed=2, device="cpu"):
    """Anisotropic |K|^-11/3 GRF (axis0 4x, axis2 2x stretched) plus an oblique step of 0.3 x range."""
    f = grf(shape, -11.0 / 3.0, seed, stretch=(4.0, 1.0, 2.0), device=device)
    rng = float(f.max() - f.min())
    n = torch.tensor([1.0, 2.0, 3.0], dtype=f.dtype, device=f.device)
    n = n / n.norm()
    c = [(s - 1) / 2.0 for s in shape]
    i0 = torch.arange(shape[0], dtype=f.dtype, device=f.device)[:, None, None] - c[0]
    i1 = torch.arange(shape[1], dtype=f.dtype, device=f.device)[None, :, None] - c[1]
    i2 = torch.arange(shape[2], dtype=f.dtype, device=f.device)[None, None, :] - c[2]
    f += 0.3 * rng * ((n[0] * i0 + n[1] * i1 + n[2] * i2) > 0).to(f.dtype)
    return f.to(torch.float32)


def config3_field(shape=(1024, 1024, 1024), seed=3, sigma=1.0, device="cpu"):
    """exp of a unit-variance |k|^-3 GRF (heavy tailed, density-like), rescaled to [0, 1e3]."""
    g = grf(shape, -3.0, seed, device=device, dtype=torch.float32)
    g = (g - g.mean()) / g.std() * sigma
    g = torch.exp(g)
    return rescale(g, 0.0, 1e3).to(torch.float32)


def config5_field(shape=(512, 512, 512), seed=5, device="cpu"):
    """Axis0 lower half smooth (|k|^-11/3), upper half rough (|k|^-5/3), same range."""
    h = shape[0] // 2
    a = rescale(grf((h,) + tuple(shape[1:]), -11.0 / 3.0, seed, device=device), 0.0, 1.0)
    b = rescale(grf((shape[0] - h,) + tuple(shape[1:]), -5.0 / 3.0, seed + 1, device=device), 0.0, 1.0)
    return torch.cat([a, b], 0).to(torch.float32)


def shifted_series(shape, steps, seed=4, device="cpu"):
    """Time series: step n = base field shifted 0.5 n voxels along axis2, 5% fresh noise blended in."""
    base = grf(shape, -11.0 / 3.0, seed, device=device)
    spec = torch.fft.rfftn(base)
    k2 = torch.fft.rfftfreq(shape[2], device=device, dtype=base.dtype)
    out = []
    for n in range(steps):
        phase = torch.exp(-2j * math.pi * k2 * 0.5 * n)[None, None, :]
        shifted = torch.fft.irfftn(spec * phase, s=shape)
        fresh = grf(shape, -11.0 / 3.0, seed * 1000 + n, device=device)
        out.append((0.95 * shifted + 0.05 * fresh).to(torch.float32))
    return out


def to_numpy(t):
    return np.ascontiguousarray(t.detach().cpu().numpy())
