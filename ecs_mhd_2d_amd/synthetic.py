"""Seeded synthetic initial states for tests and benchmarks.

The reference draws 10 N(0,1) real-space noise (generate_turbulence.py:54-55);
the benchmark workloads instead use a spectrum-shaped generator G(N, seed):
for each field (omega, j), complex N(0,1) coefficients on the rfft half plane,
weighted by |k|^-2 on the shell 1 <= |k| <= 10 and zero elsewhere, made
Hermitian-consistent by rfft2(irfft2(.)), zero mean, 2/3-dealiased, and
scaled to real-space RMS 2.0.  numpy's PCG64 (default_rng) makes it
reproducible on any machine.
"""
from __future__ import annotations

import numpy as np


def spectrum(n: int, seed: int, rms: float = 2.0, kmin: float = 1.0, kmax: float = 10.0) -> np.ndarray:
    """Complex128 [2, N, N/2+1] dealiased spectrum with real-space RMS ``rms`` per field."""
    rng = np.random.default_rng(seed)
    k = np.fft.fftfreq(n, d=1 / n)
    kx = k.reshape(-1, 1)
    ky = k[: n // 2 + 1].reshape(1, -1)
    kk = np.sqrt(kx * kx + ky * ky)
    shell = (kk >= kmin) & (kk <= kmax)
    weight = np.where(shell, 1.0 / np.where(kk > 0, kk, 1.0) ** 2, 0.0)
    mask = (np.abs(kx) < n / 3) & (np.abs(ky) < n / 3)
    mask[0, 0] = False
    out = np.empty((2, n, n // 2 + 1), dtype=np.complex128)
    for c in range(2):
        z = rng.standard_normal((n, n // 2 + 1)) + 1j * rng.standard_normal((n, n // 2 + 1))
        z = z * weight
        z = np.fft.rfft2(np.fft.irfft2(z, s=(n, n)))
        z[0, 0] = 0.0
        z = z * mask
        r = np.fft.irfft2(z, s=(n, n))
        z = z * (rms / np.sqrt(np.mean(r * r)))
        out[c] = z
    return out


def fields(n: int, seed: int, rms: float = 2.0) -> np.ndarray:
    """Real [2, N, N] fields of ``spectrum``."""
    return np.fft.irfft2(spectrum(n, seed, rms), s=(n, n))
