"""Diagonal estimators of the RPO Jacobian (jax_scripts/lib/preconditioners.py:225-359),
consumers of batched tangents: stochastic (Hutchinson) estimates of the
Jacobian's diagonal in real space (+-1 probes) or in Fourier space
(unit-amplitude random phases), and the diagonal preconditioner v / diag.

The reference multiplies the length-n probe by the length-(n+2) flattened
Jacobian output and reshapes that output to the field shape, which cannot
run as written; here the Jacobian's ``fields`` block is used (the probes
have zero T and sx parts, as the reference's zero padding makes them) and
the preconditioner divides the fields part of a flattened [T, fields, sx]
vector, passing T and sx through.  Probes come from numpy's PCG64
generator (the reference uses jax.random).
"""
from __future__ import annotations

import numpy as np
import torch

from . import loss_functions as LF, symmetry


def batched_fields_jvp(input_dict, param_dict, mode):
    """z [B, 2, N, N] -> the fields block of J z for tangents with zero T and
    sx parts: irfft2(rfft2(z) - g dPhi_T rfft2(z)), one batched call."""
    lin = LF.Linearization(input_dict, param_dict, mode)

    def apply(z):
        B = z.shape[0]
        dz = LF._rfft2(z.contiguous())
        f0 = lin.f0.unsqueeze(0).expand(B, -1, -1, -1).contiguous()
        f, df = LF.integrate_jvp(f0, dz, lin.T, 0.0, param_dict, mode, lin.info)
        dg = symmetry.apply_jvp(f, df, lin.sx, 0.0, param_dict)
        return torch.fft.irfft2(dz - dg, s=(lin.n, lin.n))
    return apply


def _precond(diag):
    d = diag.reshape(-1)

    def M(v, mode=None):
        """Divide the fields part of a flattened [T, fields, sx] vector by the
        diagonal (diagonal matrices are self-adjoint: mode is ignored)."""
        out = v.clone()
        out[1:-1] = v[1:-1] / d
        return out
    return M


def diagonal_preconditioner_spatial(input_dict, param_dict, jac_batched, k=8, batch=4, seed=0):
    """diag = mean over k * batch probes z in {-1, +1} of z * (J z)
    (preconditioners.py:225-288).  jac_batched: [B, 2, N, N] -> [B, 2, N, N]
    (e.g. batched_fields_jvp).  Returns (M, diag)."""
    shape = tuple(input_dict["fields"].shape)
    rng = np.random.default_rng(seed)
    diag = torch.zeros(shape, dtype=torch.float64, device="cuda")
    for _ in range(k):
        z = torch.as_tensor(2.0 * rng.integers(0, 2, size=(batch,) + shape) - 1.0).cuda()
        diag += torch.sum(z * jac_batched(z), dim=0)
    diag = diag / k / batch
    return _precond(diag), diag


def diagonal_preconditioner_fourier(input_dict, jac_batched, k=8, batch=4, seed=0):
    """Fourier-space diagonal (preconditioners.py:291-359): probes z =
    rfft2(w) / |rfft2(w)| for Gaussian fields w, diag_hat = mean conj(z) *
    rfft2(J irfft2(z)), diag = irfft2(diag_hat).  Returns (M, diag, diag_hat)."""
    shape = tuple(input_dict["fields"].shape)
    n = shape[-1]
    rng = np.random.default_rng(seed)
    acc = 0
    for _ in range(k):
        w = torch.as_tensor(rng.standard_normal((batch,) + shape)).cuda()
        z = torch.fft.rfft2(w)
        z = z / torch.abs(z)
        Jz = torch.fft.rfft2(jac_batched(torch.fft.irfft2(z, s=(n, n))))
        acc = acc + torch.sum(torch.conj(z) * Jz, dim=0)
    diag_hat = acc / k / batch
    diag = torch.fft.irfft2(diag_hat, s=(n, n))
    return _precond(diag), diag, diag_hat
