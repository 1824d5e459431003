"""Device runtime: one ``Context`` per (grid size, physics), wrapping ecs_ctx.

Spectral arrays are torch.complex128 CUDA tensors of shape [..., 2, N, N/2+1]
(the layout of jnp.fft.rfft2 of real [..., 2, N, N] fields, field 0 = omega,
field 1 = j).  Leading dimensions are flattened into the kernels' batch.
Every method launches on torch's current stream.
"""
from __future__ import annotations

import ctypes
import hashlib

import numpy as np
import torch

from . import _lib


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _as_spec(f: torch.Tensor, n: int) -> torch.Tensor:
    if f.dtype != torch.complex128 or not f.is_cuda:
        raise TypeError("spectral state must be a complex128 CUDA tensor")
    if tuple(f.shape[-3:]) != (2, n, n // 2 + 1):
        raise ValueError(f"spectral state must end in (2, {n}, {n // 2 + 1}), got {tuple(f.shape)}")
    return f.contiguous()


def _batch(f: torch.Tensor) -> int:
    return int(np.prod(f.shape[:-3])) if f.dim() > 3 else 1


def _per_batch(v, B: int) -> list[float]:
    if isinstance(v, torch.Tensor):
        v = v.detach().cpu().reshape(-1).tolist()
    elif np.ndim(v) == 0:
        v = [float(v)] * B
    v = [float(x) for x in v]
    if len(v) != B:
        raise ValueError(f"expected {B} per-candidate values, got {len(v)}")
    return v


class Context:
    """Grid tables, forcing spectrum and device workspace of one (N, nu, eta, b0, forcing)."""

    def __init__(self, n: int, nu: float, eta: float, b0, forcing: torch.Tensor | np.ndarray, device: int = 0):
        self.lib = _lib.load()
        self.n = int(n)
        self.nu, self.eta = float(nu), float(eta)
        self.b0 = (float(b0[0]), float(b0[1]))
        self.device = int(device)
        forc = torch.as_tensor(np.asarray(forcing.detach().cpu() if isinstance(forcing, torch.Tensor) else forcing),
                               dtype=torch.float64)
        if tuple(forc.shape) != (self.n, self.n):
            raise ValueError("forcing must be a real [N, N] field")
        fhat = torch.fft.rfft2(forc).numpy()
        buf = np.ascontiguousarray(np.stack([fhat.real, fhat.imag], axis=-1)).ravel()
        self._fhat = buf
        handle = ctypes.c_void_p()
        _lib.check(self.lib.ecs_ctx_create(ctypes.byref(handle), self.n, self.nu, self.eta, self.b0[0], self.b0[1],
                                           buf.ctypes.data_as(_lib.P_double), self.device), "ecs_ctx_create")
        self.handle = handle

    def __del__(self):
        try:
            if getattr(self, "handle", None) is not None and self.handle.value:
                self.lib.ecs_ctx_destroy(self.handle)
                self.handle = None
        except Exception:
            pass

    def set_rev_overlap(self, enable: bool) -> None:
        """Reverse passes on this context (every stepper's VJP): recompute the next
        segment on a second stream while the caller's stream sweeps the current one
        backward (default), or one stream (the same bits either way)."""
        _lib.check(self.lib.ecs_ctx_set_rev_overlap(self.handle, int(bool(enable))), "ecs_ctx_set_rev_overlap")

    # ------------------------------------------------------------ RHS
    def state_vel(self, f: torch.Tensor, include_dissipation: bool) -> torch.Tensor:
        f = _as_spec(f, self.n)
        out = torch.empty_like(f)
        _lib.check(self.lib.ecs_state_vel(self.handle, f.data_ptr(), out.data_ptr(), _batch(f),
                                          int(bool(include_dissipation)), _stream()), "ecs_state_vel")
        return out

    def state_vel_jvp(self, f, df, include_dissipation: bool):
        f, df = _as_spec(f, self.n), _as_spec(df, self.n)
        out, dout = torch.empty_like(f), torch.empty_like(f)
        _lib.check(self.lib.ecs_state_vel_jvp(self.handle, f.data_ptr(), df.data_ptr(), out.data_ptr(),
                                              dout.data_ptr(), _batch(f), int(bool(include_dissipation)),
                                              _stream()), "ecs_state_vel_jvp")
        return out, dout

    def state_vel_vjp(self, f, cot, include_dissipation: bool):
        f, cot = _as_spec(f, self.n), _as_spec(cot, self.n)
        grad = torch.empty_like(f)
        _lib.check(self.lib.ecs_state_vel_vjp(self.handle, f.data_ptr(), cot.data_ptr(), grad.data_ptr(),
                                              _batch(f), int(bool(include_dissipation)), _stream()),
                   "ecs_state_vel_vjp")
        return grad

    # ------------------------------------------------------------ Lawson RK4
    def lawson_rk4(self, f, dt, steps: int, masked: bool = True) -> torch.Tensor:
        f = _as_spec(f, self.n).clone()
        B = _batch(f)
        _lib.check(self.lib.ecs_lawson_rk4(self.handle, f.data_ptr(), B, _lib.dptr(_per_batch(dt, B)), int(steps),
                                           int(bool(masked)), _stream()), "ecs_lawson_rk4")
        return f

    def lawson_rk4_jvp(self, f, df, dt, ddt, steps: int, masked: bool = True):
        f, df = _as_spec(f, self.n).clone(), _as_spec(df, self.n).clone()
        B = _batch(f)
        _lib.check(self.lib.ecs_lawson_rk4_jvp(self.handle, f.data_ptr(), df.data_ptr(), B,
                                               _lib.dptr(_per_batch(dt, B)), _lib.dptr(_per_batch(ddt, B)),
                                               int(steps), int(bool(masked)), _stream()), "ecs_lawson_rk4_jvp")
        return f, df

    def lawson_rk4_vjp(self, f0, fbar, dt, steps: int, masked: bool = True, ckpt: int = 64):
        """Returns (f0bar, hbar[B]) for the cotangent fbar of the final state."""
        f0 = _as_spec(f0, self.n)
        g = _as_spec(fbar, self.n).clone()
        B = _batch(f0)
        hbar = (ctypes.c_double * B)()
        _lib.check(self.lib.ecs_lawson_rk4_vjp(self.handle, f0.data_ptr(), g.data_ptr(), hbar, B,
                                               _lib.dptr(_per_batch(dt, B)), int(steps), int(bool(masked)),
                                               int(ckpt), _stream()), "ecs_lawson_rk4_vjp")
        return g, list(hbar)

    # ------------------------------------------------------------ RK4, Lawson RK6, TDRK4
    def _fixed(self, name, f, dt, steps, *extra):
        f = _as_spec(f, self.n).clone()
        B = _batch(f)
        _lib.check(getattr(self.lib, name)(self.handle, f.data_ptr(), B, _lib.dptr(_per_batch(dt, B)), int(steps),
                                           *extra, _stream()), name)
        return f

    def _fixed_jvp(self, name, f, df, dt, ddt, steps, *extra):
        f, df = _as_spec(f, self.n).clone(), _as_spec(df, self.n).clone()
        B = _batch(f)
        _lib.check(getattr(self.lib, name)(self.handle, f.data_ptr(), df.data_ptr(), B, _lib.dptr(_per_batch(dt, B)),
                                           _lib.dptr(_per_batch(ddt, B)), int(steps), *extra, _stream()), name)
        return f, df

    def rk4(self, f, dt, steps: int) -> torch.Tensor:
        """Classical RK4 on the dissipative RHS, step dt (per candidate)."""
        return self._fixed("ecs_rk4", f, dt, steps)

    def rk4_jvp(self, f, df, dt, ddt, steps: int):
        return self._fixed_jvp("ecs_rk4_jvp", f, df, dt, ddt, steps)

    def lawson_rk6(self, f, dt, steps: int, masked: bool = True) -> torch.Tensor:
        """Lawson RK6, e = exp(dt/6 L) [* mask]."""
        return self._fixed("ecs_lawson_rk6", f, dt, steps, int(bool(masked)))

    def lawson_rk6_jvp(self, f, df, dt, ddt, steps: int, masked: bool = True):
        return self._fixed_jvp("ecs_lawson_rk6_jvp", f, df, dt, ddt, steps, int(bool(masked)))

    def tdrk4(self, f, dt, steps: int) -> torch.Tensor:
        """Two-derivative RK4 on the dissipative RHS."""
        return self._fixed("ecs_tdrk4", f, dt, steps)

    def tdrk4_jvp(self, f, df, dt, ddt, steps: int):
        return self._fixed_jvp("ecs_tdrk4_jvp", f, df, dt, ddt, steps)

    def _fixed_vjp(self, name, f0, fbar, dt, steps, ckpt, *extra):
        f0 = _as_spec(f0, self.n)
        g = _as_spec(fbar, self.n).clone()
        B = _batch(f0)
        hbar = (ctypes.c_double * B)()
        args = [self.handle, f0.data_ptr(), g.data_ptr(), hbar, B, _lib.dptr(_per_batch(dt, B)), int(steps)]
        _lib.check(getattr(self.lib, name)(*args, *extra, int(ckpt), _stream()), name)
        return g, list(hbar)

    def rk4_vjp(self, f0, fbar, dt, steps: int, ckpt: int = 64):
        """(f0bar, hbar[B]) of classical RK4 for the cotangent fbar of the final state."""
        return self._fixed_vjp("ecs_rk4_vjp", f0, fbar, dt, steps, ckpt)

    def lawson_rk6_vjp(self, f0, fbar, dt, steps: int, masked: bool = True, ckpt: int = 64):
        return self._fixed_vjp("ecs_lawson_rk6_vjp", f0, fbar, dt, steps, ckpt, int(bool(masked)))

    def tdrk4_vjp(self, f0, fbar, dt, steps: int, ckpt: int = 64):
        return self._fixed_vjp("ecs_tdrk4_vjp", f0, fbar, dt, steps, ckpt)

    # two-phase reverse passes (one forward pass, then a reverse pass per cotangent)
    SCHEMES = {"rk4": 0, "lawson_rk6": 1, "tdrk4": 2}

    def gen_fwd_ckpt(self, scheme: str, f0, dt, steps: int, ckpt: int = 64, masked: bool = True):
        """Forward pass of "rk4", "lawson_rk6" or "tdrk4" that keeps the checkpoints of its
        reverse pass: (final spectra, the bits of rk4 / lawson_rk6 / tdrk4; checkpoints)."""
        f = _as_spec(f0, self.n).clone()
        B = _batch(f)
        ck = torch.empty(int(self.lib.ecs_ckpt_doubles(self.handle, B, int(steps), int(ckpt))),
                         dtype=torch.float64, device=f.device)
        _lib.check(self.lib.ecs_gen_fwd_ckpt(self.handle, f.data_ptr(), B, _lib.dptr(_per_batch(dt, B)),
                                             int(steps), self.SCHEMES[scheme], int(bool(masked)), int(ckpt),
                                             ck.data_ptr(), _stream()), "ecs_gen_fwd_ckpt")
        return f, ck

    def gen_vjp_ckpt(self, scheme: str, ck, fbar, dt, steps: int, ckpt: int = 64, masked: bool = True):
        """(f0bar, hbar[B]) from the checkpoints of gen_fwd_ckpt (the same arguments)."""
        g = _as_spec(fbar, self.n).clone()
        B = _batch(g)
        if ck.numel() != int(self.lib.ecs_ckpt_doubles(self.handle, B, int(steps), int(ckpt))):
            raise ValueError("gen_vjp_ckpt: checkpoints of another batch, step count or segment length")
        hbar = (ctypes.c_double * B)()
        _lib.check(self.lib.ecs_gen_vjp_ckpt(self.handle, ck.data_ptr(), g.data_ptr(), hbar, B,
                                             _lib.dptr(_per_batch(dt, B)), int(steps), self.SCHEMES[scheme],
                                             int(bool(masked)), int(ckpt), _stream()), "ecs_gen_vjp_ckpt")
        return g, list(hbar)

    # ------------------------------------------------------------ Lawson RK4(3)
    def lawson_rk43(self, x, T, h0: float, atol: float, max_steps_per_checkpoint: int, checkpoints: int):
        """Returns (x_T, infos) with one info dict per candidate (keys of timestepping.py:327-333
        plus 'schedule' = (hs, accepted flags, attempts) for the tangent replay)."""
        x = _as_spec(x, self.n).clone()
        B = _batch(x)
        budget = int(max_steps_per_checkpoint) * int(checkpoints)
        info = (_lib.RK43Info * B)()
        hs = np.zeros((B, budget), dtype=np.float64)
        acc = np.zeros((B, budget), dtype=np.int8)
        _lib.check(self.lib.ecs_lawson_rk43(self.handle, x.data_ptr(), B, _lib.dptr(_per_batch(T, B)), float(h0),
                                            float(atol), int(max_steps_per_checkpoint), int(checkpoints), info,
                                            hs.ctypes.data_as(_lib.P_double), acc.ctypes.data_as(_lib.P_int8),
                                            _stream()), "ecs_lawson_rk43")
        infos = []
        for b in range(B):
            i = info[b]
            infos.append({"completed": bool(i.completed), "s": i.s, "accepted": i.accepted, "rejected": i.rejected,
                          "fevals": i.fevals, "hs": hs[b].copy(), "attempts": i.attempts, "acc": acc[b].copy()})
        return x, infos

    def lawson_rk43_jvp(self, x, dx, T, dT, infos):
        x, dx = _as_spec(x, self.n).clone(), _as_spec(dx, self.n).clone()
        B = _batch(x)
        budget = len(infos[0]["hs"])
        hs = np.ascontiguousarray(np.stack([i["hs"] for i in infos]).astype(np.float64))
        acc = np.ascontiguousarray(np.stack([i["acc"] for i in infos]).astype(np.int8))
        att = np.ascontiguousarray(np.array([i["attempts"] for i in infos], dtype=np.int32))
        _lib.check(self.lib.ecs_lawson_rk43_jvp(self.handle, x.data_ptr(), dx.data_ptr(), B,
                                                _lib.dptr(_per_batch(T, B)), _lib.dptr(_per_batch(dT, B)),
                                                hs.ctypes.data_as(_lib.P_double), acc.ctypes.data_as(_lib.P_int8),
                                                att.ctypes.data_as(_lib.P_int), budget, _stream()),
                   "ecs_lawson_rk43_jvp")
        return x, dx


    @staticmethod
    def _schedule(infos):
        budget = len(infos[0]["hs"])
        hs = np.ascontiguousarray(np.stack([i["hs"] for i in infos]).astype(np.float64))
        acc = np.ascontiguousarray(np.stack([i["acc"] for i in infos]).astype(np.int8))
        att = np.ascontiguousarray(np.array([i["attempts"] for i in infos], dtype=np.int32))
        return budget, hs, acc, att

    def lawson_rk43_vjp(self, x0, xbar, T, infos, ckpt: int = 64):
        """(x0bar, Tbar[B]) along the schedules ``infos`` of a lawson_rk43 call."""
        x0 = _as_spec(x0, self.n)
        g = _as_spec(xbar, self.n).clone()
        B = _batch(x0)
        budget, hs, acc, att = self._schedule(infos)
        Tbar = (ctypes.c_double * B)()
        _lib.check(self.lib.ecs_lawson_rk43_vjp(self.handle, x0.data_ptr(), g.data_ptr(), Tbar, B,
                                                _lib.dptr(_per_batch(T, B)), hs.ctypes.data_as(_lib.P_double),
                                                acc.ctypes.data_as(_lib.P_int8), att.ctypes.data_as(_lib.P_int),
                                                budget, int(ckpt), _stream()), "ecs_lawson_rk43_vjp")
        return g, list(Tbar)

    def _rk43_ckpt_doubles(self, B, infos, ckpt):
        m = max(1, max(int(i["accepted"]) for i in infos))
        return int(self.lib.ecs_ckpt_doubles(self.handle, B, m, int(ckpt)))

    def lawson_rk43_fwd_ckpt(self, x0, T, infos, ckpt: int = 64):
        """Checkpoints of the reverse pass along the schedules ``infos`` (x0 not modified)."""
        x0 = _as_spec(x0, self.n)
        B = _batch(x0)
        budget, hs, acc, att = self._schedule(infos)
        ck = torch.empty(self._rk43_ckpt_doubles(B, infos, ckpt), dtype=torch.float64, device=x0.device)
        _lib.check(self.lib.ecs_lawson_rk43_fwd_ckpt(self.handle, x0.data_ptr(), B, _lib.dptr(_per_batch(T, B)),
                                                     hs.ctypes.data_as(_lib.P_double),
                                                     acc.ctypes.data_as(_lib.P_int8),
                                                     att.ctypes.data_as(_lib.P_int), budget, int(ckpt),
                                                     ck.data_ptr(), _stream()), "ecs_lawson_rk43_fwd_ckpt")
        return ck

    def lawson_rk43_vjp_ckpt(self, ck, xbar, T, infos, ckpt: int = 64):
        """(x0bar, Tbar[B]) from the checkpoints of lawson_rk43_fwd_ckpt (the same schedules)."""
        g = _as_spec(xbar, self.n).clone()
        B = _batch(g)
        if ck.numel() != self._rk43_ckpt_doubles(B, infos, ckpt):
            raise ValueError("lawson_rk43_vjp_ckpt: checkpoints of another batch, schedule or segment length")
        budget, hs, acc, att = self._schedule(infos)
        Tbar = (ctypes.c_double * B)()
        _lib.check(self.lib.ecs_lawson_rk43_vjp_ckpt(self.handle, ck.data_ptr(), g.data_ptr(), Tbar, B,
                                                     _lib.dptr(_per_batch(T, B)), hs.ctypes.data_as(_lib.P_double),
                                                     acc.ctypes.data_as(_lib.P_int8),
                                                     att.ctypes.data_as(_lib.P_int), budget, int(ckpt),
                                                     _stream()), "ecs_lawson_rk43_vjp_ckpt")
        return g, list(Tbar)


_CACHE: dict = {}


def context_for(n, nu, eta, b0, forcing, device=None, lane: int = 0) -> Context:
    """Cached Context for these physics (one device workspace per key).

    A context serves one call at a time (the library refuses concurrent calls
    on one context); callers that run several operations concurrently on one
    grid size use different ``lane`` numbers, each with its own workspace."""
    if device is None:
        device = torch.cuda.current_device()
    farr = np.ascontiguousarray(np.asarray(forcing.detach().cpu() if isinstance(forcing, torch.Tensor) else forcing,
                                           dtype=np.float64))
    key = (int(n), float(nu), float(eta), float(b0[0]), float(b0[1]), hashlib.sha1(farr.tobytes()).hexdigest(),
           int(device), int(lane))
    ctx = _CACHE.get(key)
    if ctx is None:
        ctx = Context(n, nu, eta, b0, farr, device)
        _CACHE[key] = ctx
    return ctx
