"""Symmetry operators of the RPO search (jax_scripts/lib/symmetry.py:4-25) and
their tangents / transposes, as elementwise ops on spectra."""
from __future__ import annotations

import math

import torch


def shift_x(f, s, param_dict):
    """exp(i kx s) f with masked kx (symmetry.py:4-5)."""
    return torch.exp(1j * param_dict["kx"] * s) * f


def shift_reflect(f, n, param_dict):
    """exp(i ky n pi/4) f; odd n also maps kx -> -kx and flips the sign (symmetry.py:7-19)."""
    n = int(n)
    s = (n * 2 * math.pi) / 8
    f = torch.exp(1j * param_dict["ky"] * s) * f
    if n % 2 == 1:
        f = torch.roll(torch.flip(f, dims=[-2]), shifts=1, dims=-2)
        f = -f
    return f


def shift_reflect_T(g, n, param_dict):
    """Transpose (= inverse) of shift_reflect."""
    n = int(n)
    s = (n * 2 * math.pi) / 8
    if n % 2 == 1:
        g = -g
        g = torch.flip(torch.roll(g, shifts=-1, dims=-2), dims=[-2])
    return torch.conj(torch.exp(1j * param_dict["ky"] * s)) * g


def _sign(f):
    return torch.tensor([1.0, -1.0], dtype=torch.float64, device=f.device).reshape(2, 1, 1)


def rot180(f):
    """conj(f) [1, -1] (symmetry.py:21-25)."""
    return _sign(f) * torch.conj(f)


def apply(f, sx, param_dict):
    """shift_x, shift_reflect(ny), rot180 if rot (loss_functions.py:91-94)."""
    f = shift_x(f, sx, param_dict)
    f = shift_reflect(f, param_dict.get("shift_reflect_ny", 0), param_dict)
    if param_dict.get("rot", False):
        f = rot180(f)
    return f


def apply_jvp(f, df, sx, dsx, param_dict):
    a = torch.exp(1j * param_dict["kx"] * sx)
    dg = a * df + (1j * param_dict["kx"] * dsx) * a * f
    dg = shift_reflect(dg, param_dict.get("shift_reflect_ny", 0), param_dict)
    if param_dict.get("rot", False):
        dg = rot180(dg)
    return dg


def apply_vjp(f, sx, param_dict, gbar):
    """Transpose of apply_jvp: (fbar, sxbar) for the spectral cotangent gbar."""
    if param_dict.get("rot", False):
        gbar = _sign(gbar) * torch.conj(gbar)
    gbar = shift_reflect_T(gbar, param_dict.get("shift_reflect_ny", 0), param_dict)
    a = torch.exp(1j * param_dict["kx"] * sx)
    fbar = torch.conj(a) * gbar
    t = 1j * param_dict["kx"] * a * f
    sxbar = torch.sum(gbar.real * t.real + gbar.imag * t.imag)
    return fbar, sxbar


def find_optimal_symmetry(f1, f2, param_dict, nx):
    """The symmetry g = (sx, ny, rot) minimising mean(|f1 - g(f2)|^2)
    (symmetry.py:28-56): sx = 2 pi i / nx for i < nx, ny in 0..7, rot in
    (True, False), searched in that order with the first minimum kept.  The
    nx shifts of one (ny, rot) are evaluated as one batch on the device."""
    shifts = torch.arange(nx, dtype=torch.float64, device=f2.device) * (2 * math.pi / nx)
    kx = param_dict["kx"]
    errs = torch.empty((nx, 8, 2), dtype=torch.float64, device=f2.device)
    for ny in range(8):
        for ri, rot in enumerate((True, False)):
            f = torch.exp(1j * kx * shifts.reshape(-1, 1, 1, 1)) * f2
            f = shift_reflect(f, ny, param_dict)
            if rot:
                f = rot180(f)
            errs[:, ny, ri] = torch.mean(torch.square(torch.abs(f1 - f)), dim=(-3, -2, -1))
    flat = errs.reshape(-1).cpu().numpy()
    k = int(flat.argmin())     # numpy's argmin returns the first minimum
    i, rem = divmod(k, 16)
    ny, ri = divmod(rem, 2)
    best = (float((2 * math.pi / nx) * i), ny, ri == 0)
    print(f"Best symmetry found is {best}")
    return best
