"""Flattening, JVP compilation and line search (jax_scripts/lib/utils.py, newton.py)."""
from __future__ import annotations

import time

import torch

from . import loss_functions


def flatten(d: dict) -> torch.Tensor:
    """ravel_pytree order of the reference: sorted keys -> [T, fields (C-order), sx] (newton.py:106-110)."""
    f = d["fields"]
    dev = f.device if isinstance(f, torch.Tensor) else "cuda"
    T = torch.as_tensor(d["T"], dtype=torch.float64, device=dev).reshape(1)
    sx = torch.as_tensor(d["sx"], dtype=torch.float64, device=dev).reshape(1)
    return torch.cat([T, torch.as_tensor(f, dtype=torch.float64, device=dev).reshape(-1), sx])


def unflatten(x: torch.Tensor, n: int) -> dict:
    return {"T": x[0], "fields": x[1:-1].reshape(2, n, n), "sx": x[-1]}


def compile_objective_and_Jacobian(input_dict, param_dict, obj_mode):
    """(objective, jac) with jac(primal, tangent) -> dict (utils.py:93-119).

    ``obj_mode`` is the integration mode of objective_RPO.  The linearisation
    data of a primal (adaptive schedule, phase vectors) is computed once and
    reused while the same primal is passed again (one Newton iteration).  The
    cache key holds the fields tensor's storage and version counter (an
    in-place update of the fields bumps it) with T and sx, and the cached
    entry keeps a reference to that tensor, so a freed and reused address
    cannot match either; fields that are not tensors are never cached.
    """
    mode = obj_mode

    def objective(inp):
        return loss_functions.objective_RPO(inp, param_dict, mode)

    cache = {}

    def jac(primal, tangent):
        fl = primal["fields"]
        if not isinstance(fl, torch.Tensor):
            return loss_functions.Linearization(primal, param_dict, mode).jvp(tangent)
        key = (fl.data_ptr(), fl._version, tuple(fl.shape), float(primal["T"]), float(primal["sx"]))
        if cache.get("fields") is not fl:
            cache.pop("key", None)
        cache["fields"] = fl
        lin = cache.get("lin")
        if lin is None or cache.get("key") != key:
            lin = loss_functions.Linearization(primal, param_dict, mode)
            cache["lin"], cache["key"] = lin, key
        return lin.jvp(tangent)

    t0 = time.time()
    f = objective(input_dict)
    torch.cuda.synchronize()
    t1 = time.time()
    jac(input_dict, input_dict)
    torch.cuda.synchronize()
    t2 = time.time()
    print(f"Evaluating objective: {t1 - t0:.3} seconds")
    print(f"Evaluating Jacobian: {t2 - t1:.3} seconds")
    return objective, jac


def line_search_unravel(x, step, obj, unravel_fn, f_vec, max_iters=20, obj_batch=None, window=1):
    """Halving line search (utils.py:17-58): the first damp in 1, 1/2, 1/4, ...
    (at most max_iters tries) with |f(x - damp step)| < |f(x)|; returns (x, damp).
    If none improves, x is returned unchanged and damp = 2^-max_iters.

    With ``obj_batch`` (a function of a list of input dicts returning their
    objectives, each the same as ``obj`` of that input) and ``window`` > 1, the
    tries are evaluated ``window`` damps at a time, in order, and the first
    improving one is taken: the same (x, damp) as the one-at-a-time search,
    with the latency of one batched evaluation per window."""
    norm_orig = torch.linalg.vector_norm(f_vec)
    if obj_batch is None or window <= 1:
        damp = 1.0
        for _ in range(max_iters):
            xt = x - damp * step
            ft = flatten(obj(unravel_fn(xt)))
            if bool(torch.linalg.vector_norm(ft) < norm_orig):
                return xt, damp
            damp = damp / 2.0
        return x, damp
    damps = [1.0]
    for _ in range(max_iters - 1):
        damps.append(damps[-1] / 2.0)
    for w in range(0, max_iters, window):
        trial = damps[w:w + window]
        xts = [x - d * step for d in trial]
        fts = obj_batch([unravel_fn(xt) for xt in xts])
        for d, xt, ft in zip(trial, xts, fts):
            if bool(torch.linalg.vector_norm(flatten(ft)) < norm_orig):
                return xt, d
    return x, damps[-1] / 2.0
