"""Time steppers of jax_scripts/lib/timestepping.py on the GPU.

The kernels integrate the MHD RHS itself, so ``v_fn`` must be an
``mhd.VelocityFn``: without dissipation for the Lawson schemes (whose
``L_diag`` must be ``mhd.dissipation(param_dict)``), with dissipation for
classical RK4 and TDRK4, as mismatch_RPO builds them
(loss_functions.py:53,61,68,74,81).  Anything else raises (there is no
generic-callable fallback).
"""
from __future__ import annotations

import torch

from . import mhd


def _check(v_fn, L_diag, dissipative=False):
    if not isinstance(v_fn, mhd.VelocityFn) or v_fn.include_dissipation != dissipative:
        raise TypeError(f"v_fn must be mhd.VelocityFn(param_dict, include_dissipation={dissipative})")
    if L_diag is not None:
        ref = mhd.dissipation(v_fn.param_dict)
        if tuple(L_diag.shape) != tuple(ref.shape) or not torch.equal(L_diag.to(ref.dtype), ref):
            raise ValueError("L_diag must equal mhd.dissipation(param_dict)")
    return mhd.context(v_fn.param_dict)


def lawson_rk4(f, t, steps, v_fn, L_diag, mask=None):
    """Lawson RK4 (timestepping.py:46-81): h = t/steps, e = exp(h/2 L) [* mask].

    ``f`` may carry leading batch dimensions; ``t`` is a float or one value
    per candidate.
    """
    ctx = _check(v_fn, L_diag)
    B = int(torch.tensor(f.shape[:-3]).prod()) if f.dim() > 3 else 1
    t = torch.as_tensor(t, dtype=torch.float64).reshape(-1).expand(B)
    h = (t / steps).tolist()
    return ctx.lawson_rk4(f, h, int(steps), masked=mask is not None)


def _steps_h(f, t, steps):
    B = int(torch.tensor(f.shape[:-3]).prod()) if f.dim() > 3 else 1
    t = torch.as_tensor(t, dtype=torch.float64).reshape(-1).expand(B)
    return (t / steps).tolist()


def rk4(f, t, steps, v_fn):
    """Classical RK4 (timestepping.py:17-44), h = t/steps; v_fn with dissipation."""
    ctx = _check(v_fn, None, dissipative=True)
    return ctx.rk4(f, _steps_h(f, t, steps), int(steps))


def lawson_rk6(f, t, steps, v_fn, L_diag, mask=None):
    """Lawson RK6 (timestepping.py:83-125): h = t/steps, e = exp(h/6 L) [* mask]."""
    ctx = _check(v_fn, L_diag)
    return ctx.lawson_rk6(f, _steps_h(f, t, steps), int(steps), masked=mask is not None)


def tdrk4(f, t, steps, v_fn):
    """Two-derivative RK4 (timestepping.py:127-149); v_fn with dissipation."""
    ctx = _check(v_fn, None, dissipative=True)
    return ctx.tdrk4(f, _steps_h(f, t, steps), int(steps))


def lawson_rk43(x, vel_fn, diag_L, t, h=1e-2, atol=1e-4, max_steps_per_checkpoint=1024, checkpoints=1):
    """Adaptive Lawson RK4(3) (timestepping.py:151-335); returns (x, info).

    info has the reference's keys (completed, s, accepted, rejected, fevals,
    hs) plus 'schedule' (per-attempt accepted flags and count) that the JVP
    replays.  For a batch, info is a list with one dict per candidate.
    """
    ctx = _check(vel_fn, diag_L)
    B = int(torch.tensor(x.shape[:-3]).prod()) if x.dim() > 3 else 1
    T = torch.as_tensor(t, dtype=torch.float64).reshape(-1).expand(B).tolist()
    out, infos = ctx.lawson_rk43(x, T, h, atol, max_steps_per_checkpoint, checkpoints)
    return out, (infos[0] if x.dim() == 3 else infos)
