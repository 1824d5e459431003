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
    returns a dict with H, the matvec count and two per-Krylov-step residual
    histories (k = 1..inner):

    * ``krylov_residuals``: min_y |H_k y - |b| e_1| / |b|, the Arnoldi
      residual estimate.  For modified Gram-Schmidt it follows the true
      residual |A Q_k y - b| / |b| down to rounding level even after the
      basis has lost orthogonality;
    * ``krylov_residuals_qtb``: min_y |H_k y - (Q^T b)_k| / |b| with the
      right-hand side the reference's solve uses (b2 = Q^T b, linalg.py:92).
      Once the basis has lost orthogonality, Q^T b is no longer |b| e_1 and
      this value stops tracking the true residual.

    ``orthogonality_loss`` = |Q^T Q - I|_F of the final basis says how far
    the second form has drifted; ``Q`` is the basis itself, rows q_0..q_inner
    on the device.
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
    return x, rel_res, _history(H, b2, float(nb), Q, inner)


def _history(H, b2, nb, Q, inner):
    e1 = np.zeros(inner + 1)
    e1[0] = nb
    hist, hist_qtb = [], []
    for k in range(1, inner + 1):
        Hk = H[: k + 1, :k]
        for rhs, out in ((e1, hist), (b2, hist_qtb)):
            yk = np.linalg.lstsq(Hk, rhs[: k + 1], rcond=None)[0]
            out.append(float(np.linalg.norm(Hk @ yk - rhs[: k + 1]) / nb))
    gram = Q @ Q.T
    gram.diagonal().sub_(1.0)
    return {"H": H, "Q": Q, "krylov_residuals": hist, "krylov_residuals_qtb": hist_qtb,
            "orthogonality_loss": float(torch.linalg.matrix_norm(gram)), "matvecs": inner + 1}


# ------------------------------------------------------------------ hookstep
def _mgs_reorth(Q, k, w, hcol, passes=2):
    lib = _lib.load()
    n = Q.shape[1]
    _lib.check(lib.ecs_mgs_reorth(Q.data_ptr(), n, n, k, w.data_ptr(), hcol.data_ptr(), int(passes),
                                  _runtime._stream()), "ecs_mgs_reorth")


def _arnoldi_reorth(A, b, m):
    """Arnoldi with one recorded MGS pass and two discarded reorthogonalisation
    passes per step (linalg.py:137-153, :255-268), basis on the device.
    Returns Q [(m+1), n] (rows are basis vectors) and H (m+1, m) on the host."""
    n = b.numel()
    Q = torch.zeros((m + 1, n), dtype=torch.float64, device=b.device)
    Hd = torch.zeros((m, m + 2), dtype=torch.float64, device=b.device)
    Q[0] = b / torch.linalg.vector_norm(b)
    for k in range(m):
        w = A(Q[k]).contiguous().clone()
        _mgs_reorth(Q, k, w, Hd[k])
    H = np.zeros((m + 1, m))
    Hh = Hd.cpu().numpy()
    for k in range(m):
        H[: k + 2, k] = Hh[k, : k + 2]
    return Q, H


def _bisection(g, lower, upper, maxit=128):
    """linalg.py:165-177 / :289-301: t = 0 if g(0) > 0, else the last midpoint."""
    if g(0.0) > 0:
        return 0.0
    t = lower
    for _ in range(maxit):
        t = (lower + upper) / 2
        if g(t) > 0:
            upper = t
        else:
            lower = t
    return t


def _flat_objective(objective, input_dict):
    from . import utils
    n = input_dict["fields"].shape[-1]
    x = utils.flatten(input_dict)
    return x, (lambda v: utils.flatten(objective(utils.unflatten(v, n)))), (lambda v: utils.unflatten(v, n))


def _first_decrease(x, tries, objective, objective_batch, window, nref, unravel):
    """The first (step, s) of ``tries`` (an iterator of the steps a trust-region
    loop tries, in order, with the size s each was made for) whose
    |objective(x - step)| < nref, or None.  With ``objective_batch`` and
    window > 1 the tries are integrated ``window`` at a time as one batch;
    the answer is that of evaluating them one by one."""
    from . import utils
    if objective_batch is None or window <= 1:
        for dx, s in tries:
            if float(torch.linalg.vector_norm(utils.flatten(objective(unravel(x - dx))))) < nref:
                return dx, s
        return None
    while True:
        win = []
        for item in tries:
            win.append(item)
            if len(win) == window:
                break
        if not win:
            return None
        outs = objective_batch([unravel(x - dx) for dx, _ in win])
        for (dx, s), o in zip(win, outs):
            if float(torch.linalg.vector_norm(utils.flatten(o))) < nref:
                return dx, s


def newton_gmres_hookstep_v2(A, b, m, s, objective, input_dict, max_halvings=64, objective_batch=None,
                             window=1):
    """Newton step with a bound on its size (linalg.py:217-337).

    Krylov space of A of dimension m from b, SVD of the (m+1) x m Hessenberg
    matrix, least-squares step y; while |y| > s*s the step is regularised to
    |y| = s by bisection on t in (H^T H + t I) y = H^T Q^T b; the step is kept
    when |objective(x - dx)| < |b| (then s doubles), otherwise s halves.
    Returns (new_input_dict, s, rel_res) with rel_res = |A dx - b| / |b|.
    The reference loops until a step decreases |f|; here more than
    ``max_halvings`` halvings raise RuntimeError.  The steps tried depend only
    on the sequence of sizes, so with ``objective_batch`` (objectives of a
    list of input dicts) and window > 1 they are integrated a window at a
    time, with the same result.
    """
    Q, H = _arnoldi_reorth(A, b, m)
    b2 = torch.mv(Q, b).cpu().numpy()
    U, S, Vh = np.linalg.svd(H, full_matrices=False)
    y0 = Vh.T @ ((U.T @ b2) / S)
    x, f_vec, unravel = _flat_objective(objective, input_dict)
    nb = float(torch.linalg.vector_norm(b))

    def tries():
        y, sk = y0, s
        for _ in range(max_halvings):
            if np.linalg.norm(y) > sk * sk:
                b4 = Vh @ (H.T @ b2)
                g = lambda t: sk - np.linalg.norm(b4 / (S * S + t))
                t = _bisection(g, 0.0, np.sqrt(m) * np.linalg.norm(b4) / sk)
                y = Vh.T @ (b4 / (S * S + t))
            yield torch.mv(Q[:m].T.contiguous(), torch.as_tensor(y, device=b.device)), sk
            sk = sk / 2

    hit = _first_decrease(x, tries(), objective, objective_batch, window, nb, unravel)
    if hit is None:
        raise RuntimeError("hookstep: no step decreased |f|")
    dx, s = hit
    s = s * 2
    rel_res = float(torch.linalg.vector_norm(A(dx) - b)) / nb
    return unravel(x - dx), s, rel_res


def newton_gmres_hookstep(AtA, Atb, m, s, f, f0, J, b, input_dict, max_halvings=64, objective_batch=None,
                          window=1):
    """Hookstep on the normal equations (linalg.py:102-213): Krylov space of
    A^T A from A^T b, eigendecomposition of the m x m Hessenberg matrix (lower
    triangle, as jnp/numpy eigh read it), bisection for a step of norm s on
    [0, |b| / s].  f maps input dicts to output dicts, f0 is the flattened
    objective at input_dict.  Returns (new_input_dict, s, rel_res_1,
    rel_res_2) = (..., |A^T A step - A^T b| / |A^T b|, |J step + b| / |b|).
    With ``objective_batch`` and window > 1 the tries are integrated a window
    at a time (the same result)."""
    Q, Hf = _arnoldi_reorth(AtA, Atb, m)
    H = Hf[:m, :m]
    D, V = np.linalg.eigh(H)
    b2 = V.T @ torch.mv(Q[:m], Atb).cpu().numpy()
    x0, f_vec, unravel = _flat_objective(f, input_dict)
    nf0 = float(torch.linalg.vector_norm(f0))
    nb = float(torch.linalg.vector_norm(b))

    def tries():
        sk = s
        for _ in range(max_halvings):
            g = lambda t: sk - np.linalg.norm(b2 / (D + t))
            t = 0.0 if g(0.0) > 0 else _bisection(g, 0.0, nb / sk)
            yield torch.mv(Q[:m].T.contiguous(), torch.as_tensor(V @ (b2 / (D + t)), device=Atb.device)), sk
            sk = sk / 2

    hit = _first_decrease(x0, tries(), f, objective_batch, window, nf0, unravel)
    if hit is None:
        raise RuntimeError("hookstep: no step decreased |f|")
    step, s = hit
    s = s * 2
    rr1 = float(torch.linalg.vector_norm(AtA(step) - Atb) / torch.linalg.vector_norm(Atb))
    rr2 = float(torch.linalg.vector_norm(J(step) + b)) / nb
    return unravel(x0 - step), s, rr1, rr2
