"""Spectral domain and the MHD right-hand side (jax_scripts/lib/mhd_jax.py).

Tensors are torch CUDA tensors: real fields [..., 2, N, N] float64, spectra
[..., 2, N, N/2+1] complex128 (field 0 = vorticity omega, 1 = current j).
The RHS itself runs in the HIP library; this module builds the same
``param_dict`` tables as the reference so code written against it keeps
working.
"""
from __future__ import annotations

import torch

from . import _runtime

PI_LITERAL = 3.14159265358979323846  # mhd_jax.py:47


def construct_domain(n: int, data_type=torch.float64, device="cuda") -> dict:
    """Grid, wavenumbers, 2/3 mask and uncurl tables (mhd_jax.py:20-94).

    x varies along axis -2 and y along axis -1; kx, ky, to_u, to_v are
    pre-multiplied by the mask, whose (0,0) entry is False.
    """
    assert data_type in (torch.float32, torch.float64)
    grid = torch.arange(n, dtype=data_type, device=device) / n * 2 * PI_LITERAL
    x, y = torch.meshgrid(grid, grid, indexing="ij")
    k = torch.fft.fftfreq(n, d=1 / n, dtype=data_type, device=device)
    kx = k.reshape(-1, 1)
    ky = k[: n // 2 + 1].reshape(1, -1)
    mask = (kx.abs() < n / 3) & (ky.abs() < n / 3)
    mask[0, 0] = False
    k_sq = kx * kx + ky * ky
    k_sq[0, 0] = 1
    inv_lap = -1 / k_sq
    inv_lap[0, 0] = 0
    inv_lap = inv_lap * mask
    to_u = ky / k_sq * mask
    to_v = -kx / k_sq * mask
    return {"x": x, "y": y, "kx": kx * mask, "ky": ky * mask, "mask": mask, "to_u": to_u, "to_v": to_v,
            "inv_lap": inv_lap}


def make_param_dict(n: int, nu: float, eta: float, b0=(0.0, 0.1), forcing_amp: float = -4.0,
                    forcing_k: float = 4.0, device="cuda") -> dict:
    """construct_domain + nu, eta, b0 and the forcing F = amp cos(k y).

    Every shipped solution uses F = -4 cos(4 y) (generate_turbulence.py:49-50).
    """
    pd = construct_domain(n, torch.float64, device)
    pd.update({"nu": float(nu), "eta": float(eta), "b0": [float(b0[0]), float(b0[1])],
               "forcing": forcing_amp * torch.cos(forcing_k * pd["y"])})
    return pd


def context(param_dict: dict, lane: int = 0) -> _runtime.Context:
    """The device context (tables, workspace) for this param_dict's physics;
    concurrent callers on one grid size use distinct lanes."""
    n = param_dict["x"].shape[-1]
    return _runtime.context_for(n, param_dict["nu"], param_dict["eta"], param_dict["b0"], param_dict["forcing"],
                                lane=lane)


def dissipation(param_dict: dict) -> torch.Tensor:
    """L = -(kx^2 + ky^2) [nu; eta] with masked kx, ky (mhd_jax.py:97-102)."""
    k_sq = param_dict["kx"] ** 2 + param_dict["ky"] ** 2
    coeffs = torch.tensor([param_dict["nu"], param_dict["eta"]], dtype=k_sq.dtype,
                          device=k_sq.device).reshape(2, 1, 1)
    return -k_sq * coeffs


def state_vel(fields: torch.Tensor, param_dict: dict, include_dissipation: bool) -> torch.Tensor:
    """State velocity of (omega, j) in Fourier space (mhd_jax.py:105-175), on the GPU."""
    return context(param_dict).state_vel(fields, include_dissipation)


class VelocityFn:
    """The ``v_fn`` the reference's steppers take: lambda f: state_vel(f, pd, diss)
    (loss_functions.py:61,74).  The fused steppers need to know it is the MHD
    RHS, so they accept this object (callable like the lambda)."""

    def __init__(self, param_dict: dict, include_dissipation: bool = False):
        self.param_dict = param_dict
        self.include_dissipation = bool(include_dissipation)

    def __call__(self, f):
        return state_vel(f, self.param_dict, self.include_dissipation)

    def jvp(self, f, df):
        return context(self.param_dict).state_vel_jvp(f, df, self.include_dissipation)

    def vjp(self, f, cot):
        return context(self.param_dict).state_vel_vjp(f, cot, self.include_dissipation)
