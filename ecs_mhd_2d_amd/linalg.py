"""Matrix-free GMRES (jax_scripts/lib/linalg.py:12-98) with device-resident Krylov basis.

The Arnoldi orthogonalisation is the reference's single-pass modified
Gram-Schmidt (no reorthogonalisation, no breakdown guard, no restarts:
``outer`` is accepted and ignored as in the reference).  Each sweep runs in
the HIP library (``ecs_mgs``: deterministic fused dot/axpy kernels); the
(inner+1) x inner least-squares problem is solved on the host with
numpy.linalg.lstsq(rcond=None), the SVD solver the reference's
jnp.linalg.lstsq also uses.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib, _runtime


def _mgs(Q, k, w, hcol):
    lib = _lib.load()
    n = Q.shape[1]
    _lib.check(lib.ecs_mgs(Q.data_ptr(), n, n, k, w.data_ptr(), hcol.data_ptr(), _runtime._stream()), "ecs_mgs")


def gmres(A, b, inner, outer=1, preconditioner_list=(), return_history=False):
    """Solve A x = b over the Krylov space of dimension ``inner``.

    A: callable on flat float64 CUDA vectors; b: flat float64 CUDA vector.
    Returns (x, rel_res) like the reference (rel_res = |M(A x) - b| / |b|
    with one extra matvec, linalg.py:97).  With ``return_history`` also
    returns a dict with H, the per-Krylov-step least-squares residuals
    min_y |H_k y - Q_k^T b| / |b| (k = 1..inner) and the matvec count.
    """
    def precond(v):
        for M in preconditioner_list:
            v = M(v)
        return v

    if b.dtype != torch.float64 or not b.is_cuda or b.dim() != 1:
        raise TypeError("b must be a flat float64 CUDA vector")
    n = b.numel()
    b = precond(b).contiguous()
    Q = torch.zeros((inner + 1, n), dtype=torch.float64, device=b.device)
    Hd = torch.zeros((inner, inner + 2), dtype=torch.float64, device=b.device)
    Q[0] = b / torch.linalg.vector_norm(b)
    for k in range(inner):
        w = precond(A(Q[k])).contiguous().clone()
        _mgs(Q, k, w, Hd[k])
    H = np.zeros((inner + 1, inner))
    Hh = Hd.cpu().numpy()
    for k in range(inner):
        H[: k + 2, k] = Hh[k, : k + 2]
    b2 = torch.mv(Q, b).cpu().numpy()
    y = np.linalg.lstsq(H, b2, rcond=None)[0]
    x = torch.mv(Q[:inner].T.contiguous(), torch.as_tensor(y, device=b.device))
    nb = torch.linalg.vector_norm(b)
    rel_res = float(torch.linalg.vector_norm(precond(A(x)) - b) / nb)
    if not return_history:
        return x, rel_res
    nbh = float(nb)
    hist = []
    for k in range(1, inner + 1):
        yk = np.linalg.lstsq(H[: k + 1, :k], b2[: k + 1], rcond=None)[0]
        hist.append(float(np.linalg.norm(H[: k + 1, :k] @ yk - b2[: k + 1]) / nbh))
    return x, rel_res, {"H": H, "krylov_residuals": hist, "matvecs": inner + 1}
