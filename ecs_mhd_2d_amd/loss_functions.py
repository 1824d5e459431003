"""RPO objective, loss and their derivatives (jax_scripts/lib/loss_functions.py:9-148).

Inputs follow the reference: ``input_dict = {"fields": real [2, N, N],
"T": scalar, "sx": scalar}`` (float64 CUDA tensors or Python floats) and a
``param_dict`` from ``mhd.make_param_dict`` plus the mode items
(``num_checkpoints``/``ministeps`` for "Lawson_RK4", ``adaptive_dict`` for
"Lawson_RK43", and ``shift_reflect_ny``, ``rot``).  Time integration runs in
the fused HIP steppers; the spectral transforms at the boundary of an
objective evaluation (rfft2 of the input fields, irfft2 of the output) use
torch.fft on the GPU.

Derivatives replace jax.jvp / jax.vjp / jax.value_and_grad of these
functions: ``objective_RPO_jvp`` (utils.py:108), ``objective_RPO_vjp``
(newton.py:168-169) and ``RPOLoss`` / ``loss_RPO_value_and_grad``
(adjoint_descent.py:90-96), the last a torch.autograd.Function whose backward
is the hand-written adjoint kernel with checkpoint/recompute every 64 steps.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib, _runtime, mhd, symmetry

MODES = ("RK4", "Lawson_RK4", "Lawson_RK6", "Lawson_RK43", "TDRK4")
CKPT = 64  # steps per reverse-pass segment


def _f(v) -> float:
    return float(v.item() if isinstance(v, torch.Tensor) else v)


def _fields(inp) -> torch.Tensor:
    f = inp["fields"]
    if not isinstance(f, torch.Tensor):
        f = torch.as_tensor(f)
    if not f.is_cuda:
        f = f.cuda()
    return f.to(torch.float64).contiguous()


def _rfft2(x):
    return torch.fft.rfft2(x).contiguous()


def _irfft2(z, n):
    return torch.fft.irfft2(z, s=(n, n))


def _col_weight(n, inner, edge, device):
    w = torch.full((1, n // 2 + 1), inner, dtype=torch.float64, device=device)
    w[0, 0] = edge
    w[0, -1] = edge
    return w


def irfft2_T(g, n):
    """Real transpose of irfft2: (w'/N^2) rfft2(g), w' = 1 on ky = 0, N/2 and 2 elsewhere."""
    return _col_weight(n, 2.0, 1.0, g.device) / (n * n) * _rfft2(g)


def rfft2_T(c, n):
    """Real transpose of rfft2: N^2 irfft2(w c), w = 1 on ky = 0, N/2 and 1/2 elsewhere."""
    return n * n * _irfft2(_col_weight(n, 0.5, 1.0, c.device) * c, n)


def _rk4_steps(pd):
    nc, ms = int(pd["num_checkpoints"]), int(pd["ministeps"])
    return nc, ms, nc * ms


def _rk4_h(T, pd):
    nc, ms, _ = _rk4_steps(pd)
    return (T / nc) / ms  # loss_functions.py:63 then timestepping.py:69


def integrate(f0, T, param_dict, mode):
    """The time integration of mismatch_RPO (loss_functions.py:48-88)."""
    ctx = mhd.context(param_dict)
    if mode == "Lawson_RK4":
        _, _, steps = _rk4_steps(param_dict)
        return ctx.lawson_rk4(f0, _rk4_h(T, param_dict), steps, masked=True), None
    if mode == "Lawson_RK43":
        ad = param_dict["adaptive_dict"]
        f, infos = ctx.lawson_rk43(f0, T, 1e-2, ad["atol"], ad["max_steps_per_checkpoint"], ad["checkpoints"])
        return f, infos[0]
    if mode == "RK4":
        _, _, steps = _rk4_steps(param_dict)
        return ctx.rk4(f0, _rk4_h(T, param_dict), steps), None
    if mode == "TDRK4":
        _, _, steps = _rk4_steps(param_dict)
        return ctx.tdrk4(f0, _rk4_h(T, param_dict), steps), None
    if mode == "Lawson_RK6":
        steps = int(param_dict["steps"])
        return ctx.lawson_rk6(f0, T / steps, steps, masked=True), None
    raise ValueError(f"You selected mode = {mode}, which does not exist.")


def integrate_jvp(f0, df0, T, dT, param_dict, mode, info=None):
    """Forward-mode derivative of the integration of mismatch_RPO for spectra
    with leading batch dimensions (all candidates share T and dT; mode
    "Lawson_RK43" replays ``info``, the schedule of the primal, for every
    candidate).  Returns (f, df)."""
    ctx = mhd.context(param_dict)
    if mode in ("Lawson_RK4", "RK4", "TDRK4"):
        _, _, steps = _rk4_steps(param_dict)
        h, dh = _rk4_h(T, param_dict), _rk4_h(dT, param_dict)
        if mode == "Lawson_RK4":
            return ctx.lawson_rk4_jvp(f0, df0, h, dh, steps, True)
        if mode == "RK4":
            return ctx.rk4_jvp(f0, df0, h, dh, steps)
        return ctx.tdrk4_jvp(f0, df0, h, dh, steps)
    if mode == "Lawson_RK6":
        steps = int(param_dict["steps"])
        return ctx.lawson_rk6_jvp(f0, df0, T / steps, dT / steps, steps, True)
    if mode == "Lawson_RK43":
        B = int(np.prod(f0.shape[:-3])) if f0.dim() > 3 else 1
        return ctx.lawson_rk43_jvp(f0, df0, T, dT, [info] * B)
    raise ValueError(f"You selected mode = {mode}, which does not exist.")


def mismatch_RPO(input_dict, param_dict, mode):
    """(f0, f, info): f0 = rfft2(fields) (unmasked), f = g(Phi_T f0) (loss_functions.py:9-96)."""
    f0 = _rfft2(_fields(input_dict))
    f, info = integrate(f0, _f(input_dict["T"]), param_dict, mode)
    f = symmetry.apply(f, _f(input_dict["sx"]), param_dict)
    return f0, f, info


def loss_RPO(input_dict, param_dict, mode):
    """(mean(irfft2(f - f0)^2), info)  (loss_functions.py:99-111)."""
    f0, f, info = mismatch_RPO(input_dict, param_dict, mode)
    n = f0.shape[-2]
    return torch.mean(torch.square(_irfft2(f - f0, n))), info


def phase_vectors(f0, param_dict):
    """Stop-gradient phase vectors vt = irfft2(state_vel(f0, diss)), vx = irfft2(i kx f0)
    (loss_functions.py:115-122)."""
    n = f0.shape[-2]
    vt = _irfft2(mhd.state_vel(f0, param_dict, True), n)
    vx = _irfft2(1j * param_dict["kx"] * f0, n)
    return vt, vx


def phase_conditions(f0, param_dict):
    """(pt, px): exactly zero in value (x - stop_gradient(x), loss_functions.py:127-128)."""
    z = torch.zeros((), dtype=torch.float64, device=f0.device)
    return z, z.clone()


def objective_RPO(input_dict, param_dict, mode):
    """{"fields": irfft2(f0 - f), "T": pt, "sx": px}  (loss_functions.py:132-148)."""
    f0, f, _ = mismatch_RPO(input_dict, param_dict, mode)
    pt, px = phase_conditions(f0, param_dict)
    return {"fields": _irfft2(f0 - f, f0.shape[-2]), "T": pt, "sx": px}


def integrate_batch(f0s, Ts, param_dict, mode):
    """The integration of mismatch_RPO for several inputs of one grid size at
    once (each with its own period), as one batched call of the fused
    steppers.  Every candidate's result is bitwise its solo result (the kernels
    treat candidates independently).  Returns (f [B, 2, N, N/2+1], infos)."""
    ctx = mhd.context(param_dict)
    f0 = torch.stack(f0s).contiguous()
    if mode == "Lawson_RK43":
        ad = param_dict["adaptive_dict"]
        return ctx.lawson_rk43(f0, list(Ts), 1e-2, ad["atol"], ad["max_steps_per_checkpoint"], ad["checkpoints"])
    if mode in ("Lawson_RK4", "RK4", "TDRK4"):
        _, _, steps = _rk4_steps(param_dict)
        hs = [_rk4_h(T, param_dict) for T in Ts]
        if mode == "Lawson_RK4":
            return ctx.lawson_rk4(f0, hs, steps, masked=True), [None] * len(Ts)
        if mode == "RK4":
            return ctx.rk4(f0, hs, steps), [None] * len(Ts)
        return ctx.tdrk4(f0, hs, steps), [None] * len(Ts)
    if mode == "Lawson_RK6":
        steps = int(param_dict["steps"])
        return ctx.lawson_rk6(f0, [T / steps for T in Ts], steps, masked=True), [None] * len(Ts)
    raise ValueError(f"You selected mode = {mode}, which does not exist.")


def objective_RPO_batch(input_dicts, param_dict, mode):
    """objective_RPO of several inputs of one grid size, their integrations run
    as one batch; returns the list of objective dicts, each bitwise the one
    objective_RPO returns for that input alone (the transforms, the symmetry
    and the difference are formed per input, with the shapes of the solo
    path)."""
    f0s = [_rfft2(_fields(d)) for d in input_dicts]
    f, _ = integrate_batch(f0s, [_f(d["T"]) for d in input_dicts], param_dict, mode)
    out = []
    for b, d in enumerate(input_dicts):
        fb = symmetry.apply(f[b], _f(d["sx"]), param_dict)
        pt, px = phase_conditions(f0s[b], param_dict)
        out.append({"fields": _irfft2(f0s[b] - fb, f0s[b].shape[-2]), "T": pt, "sx": px})
    return out


class Linearization:
    """Primal data shared by every JVP at one input (one Newton iteration):
    the adaptive schedule (decisions follow the primal) and the phase vectors."""

    def __init__(self, input_dict, param_dict, mode):
        if mode not in MODES:
            raise ValueError(f"You selected mode = {mode}, which does not exist.")
        self.pd, self.mode = param_dict, mode
        self.fields = _fields(input_dict)
        self.T, self.sx = _f(input_dict["T"]), _f(input_dict["sx"])
        self.n = self.fields.shape[-1]
        self.f0 = _rfft2(self.fields)
        self.info = None
        self.fT = None
        self._rev = {}   # the reverse passes' forward pass of this primal (vjp), once computed
        if mode == "Lawson_RK43":
            self.fT, self.info = integrate(self.f0, self.T, param_dict, mode)
        self.vt, self.vx = phase_vectors(self.f0, param_dict)

    def objective(self):
        """objective_RPO at this input, from the integration the linearisation
        already holds (mode Lawson_RK43: the same solve whose schedule the
        JVPs replay) -- the same operations as objective_RPO, so the same bits."""
        if self.fT is None:
            self.fT, _ = integrate(self.f0, self.T, self.pd, self.mode)
        g = symmetry.apply(self.fT, self.sx, self.pd)
        z = torch.zeros((), dtype=torch.float64, device=self.f0.device)
        return {"fields": _irfft2(self.f0 - g, self.n), "T": z, "sx": z.clone()}

    def jvp(self, tangent):
        """J(x) t for the tangent dict t (utils.py:108)."""
        dfields = _fields(tangent)
        dT, dsx = _f(tangent["T"]), _f(tangent["sx"])
        df0 = _rfft2(dfields)
        f, df = integrate_jvp(self.f0, df0, self.T, dT, self.pd, self.mode, self.info)
        dg = symmetry.apply_jvp(f, df, self.sx, dsx, self.pd)
        return {"fields": _irfft2(df0 - dg, self.n),
                "T": torch.mean(self.vt * dfields),
                "sx": torch.mean(self.vx * dfields)}


    def vjp(self, cot):
        """J(x)^T c for the cotangent dict c: the function jax.vjp(obj, x)
        returns, applied to c (newton.py:168-169); the adaptive schedule and
        final state of the primal are the ones computed once for this input,
        and so are the reverse pass's checkpoints (the first call runs the
        forward pass, later calls only the reverse pass: the same bits)."""
        return _objective_vjp(self.f0, self.T, self.sx, self.n, cot, self.pd, self.mode, self.vt, self.vx,
                              self.info, self.fT, self._rev)


def objective_RPO_jvp(primal, tangent, param_dict, mode):
    """jax.jvp(objective_RPO)(primal, tangent)[1] (utils.py:108)."""
    return Linearization(primal, param_dict, mode).jvp(tangent)


def _rk4_vjp_from(ctx, f0, fbar, h, steps, pd, cache=None):
    """Reverse pass of the Lawson RK4 integration from f0 (two-phase C-ABI);
    cache: a dict that keeps the forward pass (final state, checkpoints) for
    later cotangents at the same primal."""
    lib = _lib.load()
    hp = _lib.dptr([h])
    if cache is not None and "ck" in cache:
        fw, ck = cache["fT"], cache["ck"]
    else:
        nd = lib.ecs_ckpt_doubles(ctx.handle, 1, steps, CKPT)
        ck = torch.empty(int(nd), dtype=torch.float64, device=f0.device)
        fw = f0.clone().contiguous()
        _lib.check(lib.ecs_lawson_rk4_fwd_ckpt(ctx.handle, fw.data_ptr(), 1, hp, steps, 1, CKPT, ck.data_ptr(),
                                               None, _runtime._stream()), "ecs_lawson_rk4_fwd_ckpt")
        if cache is not None:
            cache.update(fT=fw, ck=ck)
    g = fbar.clone().contiguous()
    hb = (ctypes.c_double * 1)()
    _lib.check(lib.ecs_lawson_rk4_vjp_ckpt(ctx.handle, ck.data_ptr(), g.data_ptr(), hb, 1, hp, steps, 1, CKPT,
                                           None, _runtime._stream()), "ecs_lawson_rk4_vjp_ckpt")
    return fw, g, hb[0]


REVERSE_MODES = ("Lawson_RK4", "RK4", "TDRK4", "Lawson_RK6", "Lawson_RK43")
_GEN_MODES = {"RK4": "rk4", "TDRK4": "tdrk4", "Lawson_RK6": "lawson_rk6"}


def _gen_args(T, pd, mode):
    """(scheme, h, steps) of a fixed-step mode other than Lawson_RK4 (loss_functions.py:49-84)."""
    if mode == "Lawson_RK6":
        steps = int(pd["steps"])
        return _GEN_MODES[mode], T / steps, steps
    _, _, steps = _rk4_steps(pd)
    return _GEN_MODES[mode], _rk4_h(T, pd), steps


def _gen_forward(ctx, f0, T, pd, mode, cache):
    """The integration of a mode of _GEN_MODES as the reverse pass's forward
    pass: cache gets the final state (the bits of integrate) and the reverse
    pass's checkpoints; returns cache."""
    scheme, h, steps = _gen_args(T, pd, mode)
    fT, ck = ctx.gen_fwd_ckpt(scheme, f0, h, steps, CKPT, True)
    cache.update(fT=fT, ck=ck)
    return cache


def _integrate_vjp(ctx, f0, gbar, T, param_dict, mode, info=None, fT=None, cache=None):
    """Reverse pass of mismatch_RPO's integration: (fT, f0bar, Tbar) for the
    cotangent gbar of the integrated spectra (h = T / steps, so Tbar = hbar / steps).
    Lawson_RK43: info, fT of the primal solve when already known.  cache: a
    dict that keeps the forward pass of this primal (its final state and the
    reverse pass's checkpoints), so that later cotangents at the same primal
    (Linearization.vjp) run only the reverse pass."""
    if mode == "Lawson_RK4":
        _, _, steps = _rk4_steps(param_dict)
        fT, f0bar, hbar = _rk4_vjp_from(ctx, f0, gbar, _rk4_h(T, param_dict), steps, param_dict, cache)
        return fT, f0bar, hbar / steps
    if mode in _GEN_MODES:
        scheme, h, steps = _gen_args(T, param_dict, mode)
        if cache is None or "ck" not in cache:
            cache = _gen_forward(ctx, f0, T, param_dict, mode, {} if cache is None else cache)
        f0bar, hb = ctx.gen_vjp_ckpt(scheme, cache["ck"], gbar, h, steps, CKPT, True)
        return cache["fT"], f0bar, hb[0] / steps
    if mode == "Lawson_RK43":
        if info is None:
            fT, info = integrate(f0, T, param_dict, mode)
        if cache is None:
            f0bar, Tb = ctx.lawson_rk43_vjp(f0, gbar, T, [info], CKPT)
            return fT, f0bar, Tb[0]
        if "ck" not in cache:
            cache["ck"] = ctx.lawson_rk43_fwd_ckpt(f0, T, [info], CKPT)
        f0bar, Tb = ctx.lawson_rk43_vjp_ckpt(cache["ck"], gbar, T, [info], CKPT)
        return fT, f0bar, Tb[0]
    raise NotImplementedError(f"the reverse pass is built for modes {REVERSE_MODES}, not {mode!r}")


def objective_RPO_vjp(primal, cot, param_dict, mode="Lawson_RK4"):
    """J(x)^T c for the cotangent dict c (newton.py:168-169), modes of REVERSE_MODES.
    For many cotangents at one primal, Linearization(primal, ...).vjp reuses the
    primal's schedule."""
    if mode not in MODES:
        raise ValueError(f"You selected mode = {mode}, which does not exist.")
    fields = _fields(primal)
    f0 = _rfft2(fields)
    vt, vx = phase_vectors(f0, param_dict)
    return _objective_vjp(f0, _f(primal["T"]), _f(primal["sx"]), fields.shape[-1], cot, param_dict, mode, vt, vx)


def _objective_vjp(f0, T, sx, n, cot, param_dict, mode, vt, vx, info=None, fT=None, cache=None):
    ctx = mhd.context(param_dict)
    cbar = irfft2_T(_fields(cot), n)                 # cotangent of (f0 - g)
    # g = sym(fT): the symmetry transpose of -cbar does not need fT; sxbar does
    fT, f0bar, Tbar = _integrate_vjp(ctx, f0, _sym_T(-cbar, sx, param_dict), T, param_dict, mode, info, fT,
                                     cache)
    _, sxbar = symmetry.apply_vjp(fT, sx, param_dict, -cbar)
    n2 = 2 * n * n
    gfields = rfft2_T(cbar + f0bar, n) + (_f(cot["T"]) * vt + _f(cot["sx"]) * vx) / n2
    return {"fields": gfields, "T": torch.tensor(Tbar, dtype=torch.float64, device=f0.device), "sx": sxbar}


def _sym_T(gbar, sx, pd):
    """fbar part of symmetry.apply_vjp (independent of the state)."""
    if pd.get("rot", False):
        gbar = symmetry._sign(gbar) * torch.conj(gbar)
    gbar = symmetry.shift_reflect_T(gbar, pd.get("shift_reflect_ny", 0), pd)
    return torch.conj(torch.exp(1j * pd["kx"] * sx)) * gbar


class RPOLoss(torch.autograd.Function):
    """loss_RPO in mode "Lawson_RK4" with a hand-written backward.

    forward: one fused trajectory that also stores the state every 64 steps;
    backward: the adjoint kernels recompute each 64-step segment from its
    checkpoint (keeping the segment's stage data on the device) and run the
    transposed stages in reverse.
    """

    @staticmethod
    def forward(ctx_, fields, T, sx, param_dict):
        ctx = mhd.context(param_dict)
        n = fields.shape[-1]
        nc, ms, steps = _rk4_steps(param_dict)
        Tf, sxf = float(T), float(sx)
        h = _rk4_h(Tf, param_dict)
        f0 = _rfft2(fields.detach().to(torch.float64))
        lib = _lib.load()
        nd = lib.ecs_ckpt_doubles(ctx.handle, 1, steps, CKPT)
        ck = torch.empty(int(nd), dtype=torch.float64, device=f0.device)
        fT = f0.clone().contiguous()
        _lib.check(lib.ecs_lawson_rk4_fwd_ckpt(ctx.handle, fT.data_ptr(), 1, _lib.dptr([h]), steps, 1, CKPT,
                                               ck.data_ptr(), None, _runtime._stream()), "ecs_lawson_rk4_fwd_ckpt")
        g = symmetry.apply(fT, sxf, param_dict)
        diff = _irfft2(g - f0, n)
        ctx_.save_for_backward(ck, fT, diff)
        ctx_.pd, ctx_.h, ctx_.steps, ctx_.sx, ctx_.n = param_dict, h, steps, sxf, n
        return torch.mean(diff * diff)

    @staticmethod
    def backward(ctx_, gl):
        ck, fT, diff = ctx_.saved_tensors
        pd, n = ctx_.pd, ctx_.n
        ctx = mhd.context(pd)
        dbar = irfft2_T(diff * (gl / (n * n)), n)         # cotangent of g - f0
        fTbar, sxbar = symmetry.apply_vjp(fT, ctx_.sx, pd, dbar)
        g = fTbar.clone().contiguous()
        hb = (ctypes.c_double * 1)()
        lib = _lib.load()
        _lib.check(lib.ecs_lawson_rk4_vjp_ckpt(ctx.handle, ck.data_ptr(), g.data_ptr(), hb, 1,
                                               _lib.dptr([ctx_.h]), ctx_.steps, 1, CKPT, None, _runtime._stream()),
                   "ecs_lawson_rk4_vjp_ckpt")
        gfields = rfft2_T(g - dbar, n)
        Tbar = torch.tensor(hb[0] / ctx_.steps, dtype=torch.float64, device=diff.device)
        return gfields, Tbar, sxbar, None


class RPOLossModes(torch.autograd.Function):
    """loss_RPO in modes "RK4", "TDRK4", "Lawson_RK6" and "Lawson_RK43" with
    the hand-written reverse pass of their integrators (checkpoint every 64
    steps, each segment run forward again into a step tape before it is
    transposed; RK4(3) along the primal's schedule) as the backward."""

    @staticmethod
    def forward(ctx_, fields, T, sx, param_dict, mode):
        n = fields.shape[-1]
        f0 = _rfft2(fields.detach().to(torch.float64))
        Tf, sxf = float(T), float(sx)
        # the integration the backward reuses: the fixed-step modes' forward
        # pass with its checkpoints (the bits of integrate), RK4(3)'s schedule
        rev, info = {}, None
        if mode in _GEN_MODES:
            fT = _gen_forward(mhd.context(param_dict), f0, Tf, param_dict, mode, rev)["fT"]
        else:
            fT, info = integrate(f0, Tf, param_dict, mode)
        g = symmetry.apply(fT, sxf, param_dict)
        diff = _irfft2(g - f0, n)
        ctx_.save_for_backward(f0, fT, diff)
        ctx_.pd, ctx_.mode, ctx_.T, ctx_.sx, ctx_.n = param_dict, mode, Tf, sxf, n
        ctx_.rev, ctx_.info = rev, info
        return torch.mean(diff * diff)

    @staticmethod
    def backward(ctx_, gl):
        f0, fT, diff = ctx_.saved_tensors
        pd, n = ctx_.pd, ctx_.n
        dbar = irfft2_T(diff * (gl / (n * n)), n)         # cotangent of g - f0
        fTbar, sxbar = symmetry.apply_vjp(fT, ctx_.sx, pd, dbar)
        _, f0bar, Tbar = _integrate_vjp(mhd.context(pd), f0, fTbar.contiguous(), ctx_.T, pd, ctx_.mode,
                                        ctx_.info, fT, ctx_.rev)
        gfields = rfft2_T(f0bar - dbar, n)
        return gfields, torch.tensor(Tbar, dtype=torch.float64, device=diff.device), sxbar, None, None


def loss_RPO_value_and_grad(input_dict, param_dict, mode="Lawson_RK4"):
    """jax.value_and_grad(loss_RPO, has_aux=True): returns ((loss, None),
    grad_dict)  (adjoint_descent.py:90-96), modes of REVERSE_MODES."""
    if mode not in REVERSE_MODES:
        raise NotImplementedError(f"the reverse pass is built for modes {REVERSE_MODES}, not {mode!r}")
    # a leaf of its own: the caller's tensor keeps its requires_grad flag
    fields = _fields(input_dict).detach().requires_grad_(True)
    T = torch.tensor(_f(input_dict["T"]), dtype=torch.float64, device=fields.device, requires_grad=True)
    sx = torch.tensor(_f(input_dict["sx"]), dtype=torch.float64, device=fields.device, requires_grad=True)
    if mode == "Lawson_RK4":
        loss = RPOLoss.apply(fields, T, sx, param_dict)
    else:
        loss = RPOLossModes.apply(fields, T, sx, param_dict, mode)
    gf, gT, gs = torch.autograd.grad(loss, (fields, T, sx))
    return (loss.detach(), None), {"fields": gf, "T": gT, "sx": gs}
