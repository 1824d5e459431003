"""Guess generation: a turbulent time series, its recurrence diagram, and an
RPO guess from a near-recurrence (jax_scripts/generate_turbulence.py:22-167,
jax_scripts/lib/utils.py:60-86).

The time series is integrated with Lawson RK6 on the GPU (one fused call
for the transient and one per saved snapshot, generate_turbulence.py:88,105)
and the recurrence diagram is a HIP kernel (ecs_recurrence).  The random
initial state uses numpy's PCG64 generator: the reference draws it with
jax.random (PRNGKey(111111)), which is not available here.
"""
from __future__ import annotations


import numpy as np
import torch

from . import _lib, _runtime, mhd, symmetry

FORCING_STR = "lambda x,y : -4*jnp.cos(4*y)"


def recurrence_diagram(fs):
    """dist[i, j] = || |fs_i| - |fs_j| ||_2 over all modes (generate_turbulence.py:140-151)
    for spectra fs [steps, 2, N, N/2+1] complex128 on the device."""
    if fs.dtype != torch.complex128 or not fs.is_cuda:
        raise TypeError("fs must be a complex128 CUDA tensor")
    fs = fs.contiguous()
    S = fs.shape[0]
    M = fs[0].numel()
    dist = torch.empty((S, S), dtype=torch.float64, device=fs.device)
    lib = _lib.load()
    _lib.check(lib.ecs_recurrence(fs.data_ptr(), S, M, dist.data_ptr(), _runtime._stream()), "ecs_recurrence")
    return dist


def generate_turbulence(n=128, dt=1 / 256, ministeps=64, steps=256, transient_steps=1024 * 2 * 4, nu=1 / 40,
                        eta=1 / 40, b0=(0.0, 0.3), seed=111111, amplitude=10.0, log=None):
    """The DNS of generate_turbulence.py: random fields amplitude * N(0,1) on
    the grid, masked rfft2, a transient of `transient_steps` Lawson-RK6 steps
    of dt, then `steps` snapshots ministeps * dt apart (each block of
    ministeps steps with h = (dt ministeps) / ministeps, :105).
    Returns ({"fs": fs, "dist": dist}, param_dict) with the reference's
    param_dict entries (grid, nu, eta, b0, forcing, forcing_str, dt,
    ministeps, steps)."""
    pd = mhd.make_param_dict(n, nu, eta, b0)
    ctx = mhd.context(pd)
    rng = np.random.default_rng(seed)
    f = torch.as_tensor(amplitude * rng.standard_normal((2, n, n))).cuda()
    f = (pd["mask"] * torch.fft.rfft2(f)).contiguous()
    f = ctx.lawson_rk6(f, dt, int(transient_steps), masked=True)
    if log:
        log(f"transient of {transient_steps} Lawson-RK6 steps done")
    h = (dt * ministeps) / ministeps
    fs = torch.empty((steps, 2, n, n // 2 + 1), dtype=torch.complex128, device="cuda")
    for i in range(steps):
        f = ctx.lawson_rk6(f, h, int(ministeps), masked=True)
        fs[i] = f
    dist = recurrence_diagram(fs)
    pd.update({"forcing_str": FORCING_STR, "dt": dt, "ministeps": ministeps, "steps": steps})
    return {"fs": fs, "dist": dist}, pd


def create_state_from_turb(turb_dict, idx, param_dict, nx=64):
    """RPO guess between snapshots idx[0] and idx[1] (1-based, utils.py:60-86):
    fields = irfft2(fs[idx0-1]), T = dt ministeps (idx1 - idx0),
    steps = ministeps (idx1 - idx0), and the symmetry (sx, ny, rot) that best
    maps fs[idx1-1] onto fs[idx0-1]; param_dict gains steps, shift_reflect_ny
    and rot and loses dt and ministeps."""
    f1 = turb_dict["fs"][idx[0] - 1]
    f2 = turb_dict["fs"][idx[1] - 1]
    g = symmetry.find_optimal_symmetry(f1, f2, param_dict, nx=nx)
    n = f1.shape[-2]
    f = torch.fft.irfft2(f1, s=(n, n))
    dt, ministeps = float(param_dict["dt"]), int(param_dict["ministeps"])
    T = dt * ministeps * (idx[1] - idx[0])
    param_dict.update({"steps": int(ministeps * (idx[1] - idx[0])), "shift_reflect_ny": g[1], "rot": g[2]})
    del param_dict["dt"]
    del param_dict["ministeps"]
    return {"fields": f, "T": T, "sx": g[0]}, param_dict
