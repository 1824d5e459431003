"""Newton-GMRES and adjoint-descent iterations (jax_scripts/newton.py:151-215,
jax_scripts/adjoint_descent.py:93-123) on the fused GPU path."""
from __future__ import annotations

import time

import torch

from . import adam, linalg, loss_functions, utils


def dealias(fields, param_dict):
    """irfft2(mask * rfft2(fields)) (newton.py:211-214, adjoint_descent.py:111-114)."""
    n = fields.shape[-1]
    return torch.fft.irfft2(param_dict["mask"] * torch.fft.rfft2(fields), s=(n, n))


def newton_gmres_update(input_dict, param_dict, mode, inner=64, outer=1, do_line_search=True,
                        default_damp=0.1, use_transpose=False, line_search_window=None):
    """One Newton-GMRES iteration (newton.py:151-215).

    Returns (new_input_dict, log) with the reference's logged quantities
    (rel_err = |f| / |x_new.fields| as newton.py:195 computes it, |f|,
    fwall, gmreswall, gmres_rel_res, damp, T, sx) plus H, the per-Krylov-step
    residual histories of linalg.gmres, the basis's loss of orthogonality and
    whether the adaptive integration completed (the reference goes on either
    way, newton.py:159; so does this function).  fwall covers the objective
    and the primal data of the linearisation, one integration for both.

    The line search's tries are integrated ``line_search_window`` at a time as
    one batch (default 8 for N <= 128, 4 above; 1 integrates them one by one):
    the same damp and new input as the reference's one-at-a-time search.
    """
    n = input_dict["fields"].shape[-1]
    t0 = time.perf_counter()
    # one integration serves the objective and the schedule the JVPs replay
    lin = loss_functions.Linearization(input_dict, param_dict, mode)
    f = lin.objective()
    b = utils.flatten(f)
    torch.cuda.synchronize()
    fwall = time.perf_counter() - t0
    A = lambda v: utils.flatten(lin.jvp(utils.unflatten(v, n)))
    precond = []
    if use_transpose:
        if mode not in loss_functions.REVERSE_MODES:
            raise NotImplementedError(f"use_transpose needs the reverse pass (modes {loss_functions.REVERSE_MODES})")
        # newton.py:168-169: one vjp function of the primal, applied per Krylov step
        A_T = lambda v: utils.flatten(lin.vjp(utils.unflatten(v, n)))
        precond = [A_T]
    t1 = time.perf_counter()
    step, gres, hist = linalg.gmres(A, b, inner, outer, preconditioner_list=precond, return_history=True)
    torch.cuda.synchronize()
    gwall = time.perf_counter() - t1
    x = utils.flatten(input_dict)
    obj = lambda d: loss_functions.objective_RPO(d, param_dict, mode)
    unravel = lambda v: utils.unflatten(v, n)
    t2 = time.perf_counter()
    window = 1
    if do_line_search:
        window = line_search_window if line_search_window is not None else (8 if n <= 128 else 4)
        obj_batch = lambda ds: loss_functions.objective_RPO_batch(ds, param_dict, mode)
        x, damp = utils.line_search_unravel(x, step, obj, unravel, b, max_iters=20, obj_batch=obj_batch,
                                            window=window)
    else:
        damp = default_damp
        x = x - damp * step
    torch.cuda.synchronize()
    lswall = time.perf_counter() - t2
    new = unravel(x)
    rel_err = float(torch.linalg.vector_norm(f["fields"]) / torch.linalg.vector_norm(new["fields"]))
    new = {"fields": dealias(new["fields"].contiguous(), param_dict), "T": new["T"].clone(), "sx": new["sx"].clone()}
    log = {"rel_err": rel_err, "f_norm": float(torch.linalg.vector_norm(b)), "fwall": fwall, "gmreswall": gwall,
           "gmres_rel_res": gres, "damp": damp, "T": float(new["T"]), "sx": float(new["sx"]),
           "lswall": lswall, "line_search_window": window,
           "krylov_residuals": hist["krylov_residuals"], "krylov_residuals_qtb": hist["krylov_residuals_qtb"],
           "orthogonality_loss": hist["orthogonality_loss"], "H": hist["H"], "matvecs": hist["matvecs"],
           "completed": True if lin.info is None else bool(lin.info["completed"])}
    return new, log


def newton_hookstep_update(input_dict, param_dict, mode, s, inner=64, window=None):
    """One iteration of newton_hookstep_v2.py:95-121: objective, Jacobian
    matvec, newton_gmres_hookstep_v2 with trust-region size s, dealias.
    The trust-region tries are integrated ``window`` at a time as one batch
    (default 8 for N <= 128, 4 above; 1: one by one), with the same result.
    Returns (new_input_dict, s, log)."""
    n = input_dict["fields"].shape[-1]
    t0 = time.perf_counter()
    # one integration serves the objective and the schedule the JVPs replay
    lin = loss_functions.Linearization(input_dict, param_dict, mode)
    f = lin.objective()
    b = utils.flatten(f)
    torch.cuda.synchronize()
    fwall = time.perf_counter() - t0
    A = lambda v: utils.flatten(lin.jvp(utils.unflatten(v, n)))
    obj = lambda d: loss_functions.objective_RPO(d, param_dict, mode)
    t1 = time.perf_counter()
    window = window if window is not None else (8 if n <= 128 else 4)
    obj_batch = lambda ds: loss_functions.objective_RPO_batch(ds, param_dict, mode)
    new, s, rel_res = linalg.newton_gmres_hookstep_v2(A, b, inner, s, obj, input_dict, objective_batch=obj_batch,
                                                      window=window)
    torch.cuda.synchronize()
    gwall = time.perf_counter() - t1
    rel_err = float(torch.linalg.vector_norm(f["fields"]) / torch.linalg.vector_norm(new["fields"]))
    new = {"fields": dealias(new["fields"].contiguous(), param_dict), "T": new["T"].clone(), "sx": new["sx"].clone()}
    log = {"rel_err": rel_err, "f_norm": float(torch.linalg.vector_norm(b)), "fwall": fwall, "gmreswall": gwall,
           "rel_res": rel_res, "s": s, "T": float(new["T"]), "sx": float(new["sx"])}
    return new, s, log


class AdjointDescent:
    """ADAM on the loss_RPO gradient (adjoint_descent.py:79-123): lr 1e-2,
    betas (0.9, 0.999), eps 1e-6, fields dealiased after every step."""

    def __init__(self, input_dict, param_dict, lr=1e-2, beta1=0.9, beta2=0.999, eps=1e-6, mode="Lawson_RK4"):
        self.params = {"fields": loss_functions._fields(input_dict),
                       "T": torch.as_tensor(input_dict["T"], dtype=torch.float64, device="cuda").reshape(()),
                       "sx": torch.as_tensor(input_dict["sx"], dtype=torch.float64, device="cuda").reshape(())}
        self.pd = param_dict
        self.m, self.v = adam.init_adam(self.params)
        self.t = 0
        self.hyper = (lr, beta1, beta2, eps)
        self.mode = mode   # adjoint_descent.py:49 uses "Lawson_RK4"

    def step(self):
        (loss, _), grad = loss_functions.loss_RPO_value_and_grad(self.params, self.pd, self.mode)
        lr, b1, b2, eps = self.hyper
        self.params, self.m, self.v = adam.adam_update(self.params, grad, self.m, self.v, self.t + 1, lr=lr,
                                                       beta1=b1, beta2=b2, eps=eps)
        self.params["fields"] = dealias(self.params["fields"], self.pd)
        self.t += 1
        return float(loss)
