"""Continuous-batching server for Newton-GMRES and adjoint-descent jobs on one GPU.

Jobs arrive over time (e.g. Poisson), are admitted while their estimated
device memory fits an HBM budget, and then advance together: every job is a
generator that yields the next device operation it needs; each scheduler tick
gathers the pending operations of all admitted jobs, groups the compatible
ones (same kind and grid size; same step count for the fixed-step gradient)
into one batched call of the fused kernels, runs the groups concurrently on
their own streams, and sends each job its slice of the result.  Krylov bases
and reverse-pass checkpoints stay on the device for the life of the job.

A job's arithmetic does not depend on the batch it rides in (every kernel
treats candidates independently and the Krylov reductions are per job), so
its result is the same alone and in the mix.  Latency = completion time minus
arrival time; the server reports throughput and p50/p99 latency.
"""
from __future__ import annotations

import math
import threading
import time

import numpy as np
import torch

from . import adam, loss_functions as LF, mhd, symmetry
from .synthetic import fields as synth_fields


# ------------------------------------------------------------------ operations
class Op:
    """A request from a job: kind in {"grad", "objective", "linearize", "jvp"}."""

    def __init__(self, kind, n, **kw):
        self.kind, self.n, self.kw = kind, n, kw

    def group_key(self):
        if self.kind == "grad":
            # step counts that are whole 64-step segments may differ within a
            # batch (each candidate runs its own count); others batch alike only
            steps = self.kw["steps"]
            return ("grad", self.n) if steps % 64 == 0 else ("grad", self.n, steps)
        if self.kind in ("objective", "linearize"):
            ad = self.kw["adaptive"]
            return (self.kind, self.n, ad["atol"], ad["max_steps_per_checkpoint"], ad["checkpoints"])
        if self.kind == "jvp":
            return ("jvp", self.n, len(self.kw["lin"].info["hs"]))
        return (self.kind, self.n)


def _pd(n, nu, eta, b0):
    return mhd.make_param_dict(n, nu, eta, b0)


def _batch_grad(ops, pd, lane):
    """loss_RPO value and gradient (mode Lawson_RK4) for a batch; each candidate
    takes its own step count (a multiple of the 64-step segment)."""
    import ctypes
    from . import _lib, _runtime
    ctx = mhd.context(pd, lane)
    n = ops[0].n
    nst = [int(o.kw["steps"]) for o in ops]
    steps = max(nst)
    nsteps = (ctypes.c_int * len(ops))(*nst) if len(set(nst)) > 1 else None
    B = len(ops)
    fields = torch.stack([o.kw["fields"] for o in ops]).contiguous()
    T = [float(o.kw["T"]) for o in ops]
    sx = [float(o.kw["sx"]) for o in ops]
    h = [t / k for t, k in zip(T, nst)]
    f0 = torch.fft.rfft2(fields).contiguous()
    lib = _lib.load()
    ck = torch.empty(int(lib.ecs_ckpt_doubles(ctx.handle, B, steps, LF.CKPT)), dtype=torch.float64, device="cuda")
    fT = f0.clone()
    _lib.check(lib.ecs_lawson_rk4_fwd_ckpt(ctx.handle, fT.data_ptr(), B, _lib.dptr(h), steps, 1, LF.CKPT,
                                           ck.data_ptr(), nsteps, _runtime._stream()), "ecs_lawson_rk4_fwd_ckpt")
    out = []
    gbar = torch.empty_like(f0)
    sxb = []
    diffs = []
    for b in range(B):
        pdb = dict(pd, shift_reflect_ny=ops[b].kw.get("ny", 0), rot=ops[b].kw.get("rot", False))
        g = symmetry.apply(fT[b], sx[b], pdb)
        diff = torch.fft.irfft2(g - f0[b], s=(n, n))
        diffs.append(diff)
        dbar = LF.irfft2_T(diff / (n * n), n)
        fb, sb = symmetry.apply_vjp(fT[b], sx[b], pdb, dbar)
        gbar[b] = fb
        sxb.append(sb)
        out.append(dbar)
    hb = (ctypes.c_double * B)()
    _lib.check(lib.ecs_lawson_rk4_vjp_ckpt(ctx.handle, ck.data_ptr(), gbar.data_ptr(), hb, B, _lib.dptr(h), steps,
                                           1, LF.CKPT, nsteps, _runtime._stream()), "ecs_lawson_rk4_vjp_ckpt")
    res = []
    for b in range(B):
        loss = torch.mean(diffs[b] * diffs[b])
        gf = LF.rfft2_T(gbar[b] - out[b], n)
        res.append((loss, {"fields": gf, "T": torch.tensor(hb[b] / nst[b], dtype=torch.float64, device="cuda"),
                           "sx": sxb[b]}))
    return res


def _batch_objective(ops, pd, lane):
    """objective_RPO in mode Lawson_RK43 for a batch (adaptive, per-candidate T)."""
    ctx = mhd.context(pd, lane)
    n = ops[0].n
    fields = torch.stack([o.kw["fields"] for o in ops]).contiguous()
    f0 = torch.fft.rfft2(fields).contiguous()
    ad = ops[0].kw["adaptive"]
    fT, infos = ctx.lawson_rk43(f0, [float(o.kw["T"]) for o in ops], 1e-2, ad["atol"],
                                ad["max_steps_per_checkpoint"], ad["checkpoints"])
    res = []
    for b, o in enumerate(ops):
        g = symmetry.apply(fT[b], float(o.kw["sx"]), pd)
        res.append(({"fields": torch.fft.irfft2(f0[b] - g, s=(n, n)),
                     "T": torch.zeros((), dtype=torch.float64, device="cuda"),
                     "sx": torch.zeros((), dtype=torch.float64, device="cuda")}, infos[b]))
    return res


class _Lin:
    """Primal data of one Newton iteration: adaptive schedule and phase vectors
    (the content of loss_functions.Linearization, built on a given lane)."""

    def __init__(self, fields, T, sx, adaptive, pd, lane):
        ctx = mhd.context(pd, lane)
        n = fields.shape[-1]
        self.T, self.sx = float(T), float(sx)
        self.f0 = torch.fft.rfft2(fields).contiguous()
        _, infos = ctx.lawson_rk43(self.f0, self.T, 1e-2, adaptive["atol"], adaptive["max_steps_per_checkpoint"],
                                   adaptive["checkpoints"])
        self.info = infos[0]
        self.vt = torch.fft.irfft2(ctx.state_vel(self.f0, True), s=(n, n))
        self.vx = torch.fft.irfft2(1j * pd["kx"] * self.f0, s=(n, n))


def _batch_jvp(ops, pd, lane):
    """J(x_b) t_b for Newton jobs at their own primal (schedule from 'linearize')."""
    ctx = mhd.context(pd, lane)
    n = ops[0].n
    lins = [o.kw["lin"] for o in ops]
    tans = [o.kw["tangent"] for o in ops]
    f0 = torch.stack([l.f0 for l in lins]).contiguous()
    dfields = torch.stack([t["fields"] for t in tans]).contiguous()
    df0 = torch.fft.rfft2(dfields).contiguous()
    f, df = ctx.lawson_rk43_jvp(f0, df0, [l.T for l in lins], [float(t["T"]) for t in tans], [l.info for l in lins])
    res = []
    for b, (l, t) in enumerate(zip(lins, tans)):
        dg = symmetry.apply_jvp(f[b], df[b], l.sx, float(t["sx"]), pd)
        res.append({"fields": torch.fft.irfft2(df0[b] - dg, s=(n, n)),
                    "T": torch.mean(l.vt * dfields[b]), "sx": torch.mean(l.vx * dfields[b])})
    return res


# ------------------------------------------------------------------ jobs
def _flat(d):
    return torch.cat([torch.as_tensor(d["T"], dtype=torch.float64, device="cuda").reshape(1),
                      d["fields"].reshape(-1),
                      torch.as_tensor(d["sx"], dtype=torch.float64, device="cuda").reshape(1)])


def _unflat(v, n):
    return {"T": v[0], "fields": v[1:-1].reshape(2, n, n), "sx": v[-1]}


def adjoint_descent_job(fields, T, sx, n, steps, iters, lr=1e-2):
    """ADAM on the loss_RPO gradient (adjoint_descent.py:93-123), `iters` steps."""
    params = {"fields": fields, "T": torch.tensor(T, dtype=torch.float64, device="cuda"),
              "sx": torch.tensor(sx, dtype=torch.float64, device="cuda")}
    m, v = adam.init_adam(params)
    pd_mask = None
    losses = []
    for t in range(iters):
        loss, grad = yield Op("grad", n, fields=params["fields"], T=params["T"], sx=params["sx"], steps=steps)
        losses.append(float(loss))
        params, m, v = adam.adam_update(params, grad, m, v, t + 1, lr=lr, beta1=0.9, beta2=0.999, eps=1e-6)
        if pd_mask is None:
            pd_mask = mhd.construct_domain(n)["mask"]
        params["fields"] = torch.fft.irfft2(pd_mask * torch.fft.rfft2(params["fields"]), s=(n, n))
    return {"params": params, "losses": losses}


def newton_job(fields, T, sx, n, adaptive, inner, iters):
    """Newton-GMRES (newton.py:151-215) with per-job Krylov basis on the device;
    matvecs are yielded so they batch with other jobs'."""
    from . import linalg  # noqa: F401  (same algorithm, written inline to yield matvecs)
    x = {"fields": fields, "T": T, "sx": sx}
    logs = []
    for it in range(iters):
        f, _ = yield Op("objective", n, fields=x["fields"], T=x["T"], sx=x["sx"], adaptive=adaptive)
        b = _flat(f)
        lin = yield Op("linearize", n, fields=x["fields"], T=x["T"], sx=x["sx"], adaptive=adaptive)
        nb = torch.linalg.vector_norm(b)
        Q = torch.zeros((inner + 1, b.numel()), dtype=torch.float64, device="cuda")
        Hd = torch.zeros((inner, inner + 2), dtype=torch.float64, device="cuda")
        Q[0] = b / nb
        for k in range(inner):
            w = yield Op("jvp", n, lin=lin, tangent=_unflat(Q[k], n))
            w = _flat(w).contiguous().clone()
            if not bool(torch.isfinite(w).all()):
                raise FloatingPointError(f"non-finite matvec in Newton iteration {it}, Krylov column {k} "
                                         f"(|b| = {float(nb):.3e}, T = {float(x['T']):.6g})")
            linalg._mgs(Q, k, w, Hd[k])
        H = np.zeros((inner + 1, inner))
        Hh = Hd.cpu().numpy()
        for k in range(inner):
            H[: k + 2, k] = Hh[k, : k + 2]
        b2 = torch.mv(Q, b).cpu().numpy()
        y = np.linalg.lstsq(H, b2, rcond=None)[0]
        step = torch.mv(Q[:inner].T.contiguous(), torch.as_tensor(y, device="cuda"))
        Ax = yield Op("jvp", n, lin=lin, tangent=_unflat(step, n))
        rel_res = float(torch.linalg.vector_norm(_flat(Ax) - b) / nb)
        xv = _flat(x)
        damp = 1.0
        for _ in range(20):   # line search (utils.py:17-58)
            xt = xv - damp * step
            ft, _ = yield Op("objective", n, fields=xt[1:-1].reshape(2, n, n).contiguous(), T=float(xt[0]),
                             sx=float(xt[-1]), adaptive=adaptive)
            if float(torch.linalg.vector_norm(_flat(ft))) < float(nb):
                xv = xt
                break
            damp /= 2.0
        mask = mhd.construct_domain(n)["mask"]
        x = {"fields": torch.fft.irfft2(mask * torch.fft.rfft2(xv[1:-1].reshape(2, n, n)), s=(n, n)),
             "T": float(xv[0]), "sx": float(xv[-1])}
        logs.append({"f_norm": float(nb), "gmres_rel_res": rel_res, "damp": damp})
    return {"x": x, "logs": logs}


class Job:
    def __init__(self, jid, kind, arrival, gen, mem_bytes, n):
        self.id, self.kind, self.arrival, self.gen, self.mem, self.n = jid, kind, arrival, gen, mem_bytes, n
        self.pending = None
        self.result = None
        self.error = None
        self.start = None
        self.finish = None


def job_memory(kind, n, inner=64, steps=0, budget=4096):
    """Device bytes a job keeps resident (Krylov basis, checkpoints, tapes)."""
    vec = (2 * n * n + 2) * 8
    kmax = (n - 1) // 3
    comp = 2 * (2 * kmax + 1) * (kmax + 1) * 16
    kp = (kmax + 1 + 7) // 8 * 8  # padded row stride of the x-space tape
    if kind == "newton":
        return (inner + 1) * vec + 8 * comp + budget * 64
    nseg = math.ceil(steps / 64)
    return nseg * comp + 64 * 5 * comp + 64 * 4 * 8 * n * kp * 16 + 6 * vec


# ------------------------------------------------------------------ scheduler
class Server:
    """``executor(key, ops, lane) -> results`` runs one group of compatible ops;
    the default runs them on the GPU through the fused kernels."""

    def __init__(self, hbm_budget=200e9, nu=1 / 40, eta=1 / 40, b0=(0.0, 0.1), max_batch=64, executor=None):
        self.executor = executor
        self.budget = hbm_budget
        self.nu, self.eta, self.b0 = nu, eta, b0
        self.max_batch = max_batch
        self._pds = {}
        self.ticks = 0
        self.batched_ops = 0

    def pd(self, n):
        if n not in self._pds:
            self._pds[n] = _pd(n, self.nu, self.eta, self.b0)
        return self._pds[n]

    def _run_group(self, key, ops, lane):
        pd = self.pd(ops[0].n)
        kind = key[0]
        if kind == "grad":
            return _batch_grad(ops, pd, lane)
        if kind == "objective":
            return _batch_objective(ops, pd, lane)
        if kind == "linearize":
            return [_Lin(o.kw["fields"], o.kw["T"], o.kw["sx"], o.kw["adaptive"], pd, lane) for o in ops]
        if kind == "jvp":
            return _batch_jvp(ops, pd, lane)
        raise ValueError(kind)

    def run(self, arrivals, concurrent=True, log=None):
        """arrivals: list of (t_arrival_seconds, kind, make_generator, mem_bytes, n).
        Returns the finished jobs and summary statistics; `log(msg)`, if given,
        hears of every finished job and, every 10 s, of the ticks run so far."""
        t0 = time.perf_counter()
        waiting = [Job(i, k, t, g(), m, n) for i, (t, k, g, m, n) in enumerate(sorted(arrivals, key=lambda a: a[0]))]
        active, done = [], []
        used = 0.0
        peak = 0.0
        last_log = 0.0
        while waiting or active:
            now = time.perf_counter() - t0
            while waiting and waiting[0].arrival <= now and used + waiting[0].mem <= self.budget:
                j = waiting.pop(0)
                j.start = now
                used += j.mem
                peak = max(peak, used)
                try:
                    j.pending = next(j.gen)
                except StopIteration as e:
                    j.result, j.finish = e.value, now
                    done.append(j)
                    used -= j.mem
                    continue
                except Exception as e:
                    j.error, j.finish = repr(e), now
                    done.append(j)
                    used -= j.mem
                    continue
                active.append(j)
            if not active:
                if waiting:
                    time.sleep(max(0.0, min(0.01, waiting[0].arrival - now)))
                continue
            groups = {}
            for j in active:
                groups.setdefault(j.pending.group_key(), []).append(j)
            tasks = []
            lanes = []
            per_n = {}
            for key, js in groups.items():
                for i in range(0, len(js), self.max_batch):
                    n = js[0].n
                    lanes.append(per_n.get(n, 0))   # concurrent tasks on one N get their own context
                    per_n[n] = per_n.get(n, 0) + 1
                    tasks.append((key, js[i:i + self.max_batch]))
            outs = [None] * len(tasks)
            errs = []

            def work(ti):
                key, js = tasks[ti]
                try:
                    if self.executor is not None:
                        outs[ti] = self.executor(key, [j.pending for j in js], lanes[ti])
                        return
                    s = torch.cuda.Stream()
                    with torch.cuda.stream(s):
                        outs[ti] = self._run_group(key, [j.pending for j in js], lanes[ti])
                        s.synchronize()
                except Exception as e:
                    errs.append(e)

            if concurrent and len(tasks) > 1:
                ths = [threading.Thread(target=work, args=(ti,)) for ti in range(len(tasks))]
                for t in ths:
                    t.start()
                for t in ths:
                    t.join()
            else:
                for ti in range(len(tasks)):
                    work(ti)
            if errs:
                raise errs[0]
            if self.executor is None:
                torch.cuda.synchronize()   # results cross from the worker streams to the jobs' stream
            self.ticks += 1
            self.batched_ops += sum(len(js) for _, js in tasks)
            now = time.perf_counter() - t0
            if log is not None and now - last_log >= 10.0:
                last_log = now
                kinds = ", ".join(f"{k[0]} N={k[1]} x{len(js)}" for k, js in tasks)
                log(f"{now:.1f} s: tick {self.ticks}, {len(active)} active jobs, {len(waiting)} waiting; groups: {kinds}")
            still = []
            for (key, js), res in zip(tasks, outs):
                for j, r in zip(js, res):
                    try:
                        j.pending = j.gen.send(r)
                        still.append(j)
                    except StopIteration as e:
                        j.result, j.finish = e.value, now
                        done.append(j)
                        used -= j.mem
                        if log is not None:
                            log(f"job {j.id} ({j.kind}, N={j.n}) done at {now:.1f} s, latency {now - j.arrival:.1f} s")
                    except Exception as e:  # the job's own failure ends that job, not the server
                        j.error, j.finish = repr(e), now
                        done.append(j)
                        used -= j.mem
                        if log is not None:
                            log(f"job {j.id} ({j.kind}, N={j.n}) FAILED at {now:.1f} s: {j.error}")
            active = still
        wall = time.perf_counter() - t0
        self.max_concurrent = getattr(self, "max_concurrent", 0)
        ok = [j for j in done if j.error is None]
        lat = np.array([j.finish - j.arrival for j in ok]) if ok else np.array([np.nan])
        return done, {"jobs": len(done), "failed_jobs": len(done) - len(ok), "wall_s": wall,
                      "throughput_jobs_per_s": len(ok) / wall,
                      "p50_latency_s": float(np.percentile(lat, 50)), "p99_latency_s": float(np.percentile(lat, 99)),
                      "ticks": self.ticks, "mean_batch": self.batched_ops / max(1, self.ticks),
                      "peak_resident_bytes": peak, "hbm_budget_bytes": self.budget}


def make_workload(count, rate=4.0, seed=7, sizes=(128, 256), T_range=(2.0, 8.0), adjoint_frac=0.7,
                  descent_iters=64, newton_iters=5, inner=64, dt=0.004, atol=1e-8, adaptive_budget=(64, 64)):
    """Poisson arrivals at `rate` per second of mixed jobs (the serving workload)."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    adaptive = {"atol": atol, "max_steps_per_checkpoint": adaptive_budget[0], "checkpoints": adaptive_budget[1]}
    for i in range(count):
        t += rng.exponential(1.0 / rate)
        n = int(rng.choice(sizes))
        T = float(rng.uniform(*T_range))
        kind = "adjoint" if rng.uniform() < adjoint_frac else "newton"
        f = torch.as_tensor(synth_fields(n, 1000 + i)).cuda()
        if kind == "adjoint":
            steps = int(math.ceil(round(T / dt) / 64) * 64)
            gen = (lambda f=f, T=T, n=n, steps=steps: adjoint_descent_job(f, T, 0.0, n, steps, descent_iters))
            mem = job_memory("adjoint", n, steps=steps)
        else:
            gen = (lambda f=f, T=T, n=n: newton_job(f, T, 0.0, n, adaptive, inner, newton_iters))
            mem = job_memory("newton", n, inner=inner)
        out.append((t, kind, gen, mem, n))
    return out
