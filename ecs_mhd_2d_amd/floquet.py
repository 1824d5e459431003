"""Floquet spectrum of a relative periodic orbit by block power iteration
(jax_scripts/floquet.py:42-99), a consumer of batched tangents.

The map is f -> irfft2(mask * g Phi_T rfft2(f)) (mismatch_RPO's integration
and symmetry at fixed T and sx, floquet.py:42-48); its Jacobian acts on a
block of tangent fields as one batched forward-mode call of the fused
steppers (every tangent of the block is a candidate of the batch, the
primal repeated), followed by a Householder QR of the block
(torch.linalg.qr, the LAPACK convention of jnp.linalg.qr).  After maxit
iterations R = Q^T J Q estimates the leading Floquet multipliers (its
eigenvalues).  The initial block is Gaussian from numpy's PCG64 generator
(the reference draws it with jax.random, which is not available here).
"""
from __future__ import annotations

import numpy as np
import torch

from . import loss_functions as LF, symmetry


class FloquetMap:
    """The map of floquet.py:42-48 at one orbit and its batched Jacobian."""

    def __init__(self, input_dict, param_dict, mode="TDRK4"):
        self.pd, self.mode = param_dict, mode
        self.fields = LF._fields(input_dict)
        self.n = self.fields.shape[-1]
        self.T, self.sx = LF._f(input_dict["T"]), LF._f(input_dict["sx"])
        self.f0 = LF._rfft2(self.fields)
        self.info = None
        if mode == "Lawson_RK43":
            _, self.info = LF.integrate(self.f0, self.T, param_dict, mode)

    def forward(self):
        f, _ = LF.integrate(self.f0, self.T, self.pd, self.mode)
        g = symmetry.apply(f, self.sx, self.pd)
        return torch.fft.irfft2(g * self.pd["mask"], s=(self.n, self.n))

    def jac(self, tang):
        """J tang for a block of real tangent fields [B, 2, N, N]."""
        B = tang.shape[0]
        dfl = LF._rfft2(tang.contiguous())
        f0 = self.f0.unsqueeze(0).expand(B, -1, -1, -1).contiguous()
        f, df = LF.integrate_jvp(f0, dfl, self.T, 0.0, self.pd, self.mode, self.info)
        dg = symmetry.apply_jvp(f, df, self.sx, 0.0, self.pd)
        return torch.fft.irfft2(dg * self.pd["mask"], s=(self.n, self.n))


def _orthonormalize(tang):
    B = tang.shape[0]
    q, _ = torch.linalg.qr(tang.reshape(B, -1).T.contiguous(), mode="reduced")
    return q.T.contiguous().reshape(tang.shape)


def block_power_iteration(input_dict, param_dict, block_size=32, maxit=256, mode="TDRK4", seed=0, tang=None,
                          log=None):
    """Returns {"R", "Q", "diff", "relative_error"} as floquet.py saves them:
    R = Q J(Q)^T over the block (rows = tangents), Q the final orthonormal
    block [B, 2, N, N], diff = forward(f) - f."""
    fm = FloquetMap(input_dict, param_dict, mode)
    f_out = fm.forward()
    diff = f_out - fm.fields
    rel = float(torch.linalg.vector_norm(diff) / torch.linalg.vector_norm(fm.fields))
    if log:
        log(f"relative error in periodic orbit = {rel:.3e}")
    n = fm.n
    if tang is None:
        rng = np.random.default_rng(seed)
        tang = torch.as_tensor(rng.standard_normal((block_size, 2, n, n))).cuda()
    for i in range(maxit):
        tang = _orthonormalize(fm.jac(tang))
        if log:
            log(f"iteration {i}")
    t = tang.reshape(tang.shape[0], -1)
    jt = fm.jac(tang).reshape(tang.shape[0], -1)
    R = t @ jt.T
    return {"R": R, "Q": tang, "diff": diff, "relative_error": rel}
