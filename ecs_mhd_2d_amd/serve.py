"""Continuous-batching server for Newton-GMRES and adjoint-descent jobs on one GPU.

Jobs arrive over time (e.g. Poisson) and are admitted while their estimated
device memory fits an HBM budget (a job larger than the whole budget is
rejected; one that does not fit yet lets smaller later jobs pass for a
while).  Every job is a generator that yields the next device operation it
needs.  Whenever a group slot is free, the scheduler gathers the pending
operations of the admitted jobs that are not running, groups the compatible
ones (same kind and grid size; gradients of whole 64-step segments across
step counts) into one batched call of the fused kernels, and runs it on a
context lane of its own with that lane's stream; when a group finishes, its
jobs advance at once, without waiting for the other groups.  Krylov bases
and reverse-pass checkpoints stay on the device for the life of the job.

A job's arithmetic does not depend on the batch it rides in (every kernel
treats candidates independently and the Krylov reductions are per job), so
its result is the same alone and in the mix.  Latency = completion time minus
arrival time; the server reports throughput and p50/p99 latency.
"""
from __future__ import annotations

import math
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

    def length(self):
        """Relative cost of the operation within its group key: a batch runs as
        long as its longest member, so members of similar length batch best
        (stepper steps; the period for adaptive solves, whose steps are not
        known in advance)."""
        if self.kind == "grad":
            return float(self.kw["steps"])
        if self.kind == "jvp":
            return float(self.kw["lin"].info["accepted"])
        if self.kind == "objective":
            return max(float(t["T"]) for t in self.kw["trials"])
        return float(self.kw.get("T", 0.0))


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
    """objective_RPO in mode Lawson_RK43 (adaptive, per-candidate T) of every
    trial input of a batch of operations (an operation holds the tries of one
    window of a job's line search), as one batched solve; each operation gets
    the list of its trials' (objective, info)."""
    ctx = mhd.context(pd, lane)
    n = ops[0].n
    trials = [t for o in ops for t in o.kw["trials"]]
    fields = torch.stack([t["fields"] for t in trials]).contiguous()
    f0 = torch.fft.rfft2(fields).contiguous()
    ad = ops[0].kw["adaptive"]
    fT, infos = ctx.lawson_rk43(f0, [float(t["T"]) for t in trials], 1e-2, ad["atol"],
                                ad["max_steps_per_checkpoint"], ad["checkpoints"])
    res, i = [], 0
    for o in ops:
        out = []
        for t in o.kw["trials"]:
            g = symmetry.apply(fT[i], float(t["sx"]), pd)
            out.append(({"fields": torch.fft.irfft2(f0[i] - g, s=(n, n)),
                         "T": torch.zeros((), dtype=torch.float64, device="cuda"),
                         "sx": torch.zeros((), dtype=torch.float64, device="cuda")}, infos[i]))
            i += 1
        res.append(out)
    return res


class _Lin:
    """Primal data of one Newton iteration from one adaptive solve: the
    objective (loss_functions.py:132-148), the schedule the JVPs replay and
    the phase vectors -- the content of loss_functions.Linearization."""

    def __init__(self, f0, fT, info, T, sx, pd, ctx):
        n = f0.shape[-2]
        self.T, self.sx = float(T), float(sx)
        self.f0 = f0
        self.info = info
        g = symmetry.apply(fT, self.sx, pd)
        self.objective = {"fields": torch.fft.irfft2(f0 - g, s=(n, n)),
                          "T": torch.zeros((), dtype=torch.float64, device="cuda"),
                          "sx": torch.zeros((), dtype=torch.float64, device="cuda")}
        self.vt = torch.fft.irfft2(ctx.state_vel(f0, True), s=(n, n))
        self.vx = torch.fft.irfft2(1j * pd["kx"] * f0, s=(n, n))


def _batch_linearize(ops, pd, lane):
    """One batched adaptive solve per Newton iteration: each job's objective
    and the schedule its JVPs replay come from the same integration."""
    ctx = mhd.context(pd, lane)
    fields = torch.stack([o.kw["fields"] for o in ops]).contiguous()
    f0 = torch.fft.rfft2(fields).contiguous()
    ad = ops[0].kw["adaptive"]
    fT, infos = ctx.lawson_rk43(f0, [float(o.kw["T"]) for o in ops], 1e-2, ad["atol"],
                                ad["max_steps_per_checkpoint"], ad["checkpoints"])
    return [_Lin(f0[b].contiguous(), fT[b], infos[b], o.kw["T"], o.kw["sx"], pd, ctx) for b, o in enumerate(ops)]


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
    dev = d["fields"].device
    return torch.cat([torch.as_tensor(d["T"], dtype=torch.float64, device=dev).reshape(1),
                      d["fields"].reshape(-1),
                      torch.as_tensor(d["sx"], dtype=torch.float64, device=dev).reshape(1)])


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


def _dump(dump_dir, tag, why, x, lin=None, **extra):
    """Write the input of a failed Newton iteration (fields, T, sx, and the
    adaptive schedule when there is one) so it can be replayed on the CPU."""
    if not dump_dir:
        return None
    import os
    os.makedirs(dump_dir, exist_ok=True)
    path = os.path.join(dump_dir, f"newton_failure_{tag}_{os.getpid()}.npz")
    data = {"fields": torch.as_tensor(x["fields"]).detach().cpu().numpy(), "T": float(x["T"]),
            "sx": float(x["sx"]), "why": why}
    if lin is not None:
        data.update(hs=lin.info["hs"], acc=lin.info["acc"], attempts=lin.info["attempts"])
    data.update({k: np.asarray(v) for k, v in extra.items()})
    np.savez(path, **data)
    return path


def newton_job(fields, T, sx, n, adaptive, inner, iters, dump_dir=None, tag="job", window=None):
    """Newton-GMRES (newton.py:151-215) with per-job Krylov basis on the device;
    matvecs are yielded so they batch with other jobs'.  The objective and
    the schedule of the JVPs come from one adaptive solve per iteration.  An
    adaptive solve that does not complete within its attempt budget, or a
    non-finite matvec or least-squares problem, ends the job with an error
    (after writing its input to ``dump_dir``, when given).  The line search's
    tries go out ``window`` at a time (default 8 for N <= 128, 4 above), each
    window one operation; the damp taken is the one-at-a-time search's."""
    from . import linalg
    window = window if window is not None else (8 if n <= 128 else 4)
    x = {"fields": fields, "T": T, "sx": sx}
    logs = []
    for it in range(iters):
        lin = yield Op("linearize", n, fields=x["fields"], T=x["T"], sx=x["sx"], adaptive=adaptive)
        if not lin.info["completed"]:
            path = _dump(dump_dir, tag, "incomplete", x, lin)
            raise RuntimeError(f"Newton iteration {it}: the adaptive integration stopped at s = {lin.info['s']:.6g} "
                               f"after {lin.info['attempts']} attempts (dump: {path})")
        b = _flat(lin.objective)
        nb = torch.linalg.vector_norm(b)
        Q = torch.zeros((inner + 1, b.numel()), dtype=torch.float64, device="cuda")
        Hd = torch.zeros((inner, inner + 2), dtype=torch.float64, device="cuda")
        Q[0] = b / nb
        for k in range(inner):
            w = yield Op("jvp", n, lin=lin, tangent=_unflat(Q[k], n))
            w = _flat(w).contiguous().clone()
            if not bool(torch.isfinite(w).all()):
                path = _dump(dump_dir, tag, "nonfinite_matvec", x, lin, tangent=Q[k].cpu().numpy(), it=it, k=k)
                raise FloatingPointError(f"non-finite matvec in Newton iteration {it}, Krylov column {k} "
                                         f"(|b| = {float(nb):.3e}, T = {float(x['T']):.6g}, dump: {path})")
            linalg._mgs(Q, k, w, Hd[k])
        H = np.zeros((inner + 1, inner))
        Hh = Hd.cpu().numpy()
        for k in range(inner):
            H[: k + 2, k] = Hh[k, : k + 2]
        b2 = torch.mv(Q, b).cpu().numpy()
        if not (np.isfinite(H).all() and np.isfinite(b2).all()):
            path = _dump(dump_dir, tag, "nonfinite_hessenberg", x, lin, H=H, b2=b2, it=it)
            raise FloatingPointError(f"non-finite Hessenberg matrix in Newton iteration {it} (dump: {path})")
        y = np.linalg.lstsq(H, b2, rcond=None)[0]
        step = torch.mv(Q[:inner].T.contiguous(), torch.as_tensor(y, device="cuda"))
        Ax = yield Op("jvp", n, lin=lin, tangent=_unflat(step, n))
        rel_res = float(torch.linalg.vector_norm(_flat(Ax) - b) / nb)
        xv = _flat(x)
        # line search (utils.py:17-58): damps 1, 1/2, ... (20 tries), the first that
        # lowers |f| taken; the tries are integrated a window at a time as one batch
        damps = [1.0]
        for _ in range(19):
            damps.append(damps[-1] / 2.0)
        damp = damps[-1] / 2.0
        found = None
        for w0 in range(0, 20, window):
            tried = damps[w0:w0 + window]
            xts = [xv - d * step for d in tried]
            outs = yield Op("objective", n, adaptive=adaptive, f_norm=float(nb),
                            trials=[{"fields": xt[1:-1].reshape(2, n, n).contiguous(), "T": float(xt[0]),
                                     "sx": float(xt[-1])} for xt in xts])
            for d, xt, (ft, info) in zip(tried, xts, outs):
                if info["completed"] and float(torch.linalg.vector_norm(_flat(ft))) < float(nb):
                    found = (xt, d)
                    break
            if found is not None:
                xv, damp = found
                break
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
        self.in_flight = False
        self.op_counts = {}   # kind -> operations run
        self.op_steps = {}    # kind -> stepper steps of those operations
        self.first_op = {}    # kind -> (operation, its steps): the sample timed alone afterwards


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


def useful_trials(op, result):
    """Tries of a line-search window that the one-at-a-time search evaluates:
    up to and including the first completed one that lowers |f| below the
    job's f_norm (all of them when none does)."""
    for i, (ft, info) in enumerate(result):
        if info["completed"] and float(torch.linalg.vector_norm(_flat(ft))) < op.kw["f_norm"]:
            return i + 1
    return len(result)


def op_rk_steps(op, result):
    """Stepper steps one operation integrates (the forward and the reverse
    sweep of a gradient each count its steps; an adaptive solve counts its
    attempts; a tangent replay its accepted steps).  A line-search window
    counts the tries the one-at-a-time search would make, not the ones the
    window evaluated past the first success."""
    if op.kind == "grad":
        return 2 * int(op.kw["steps"])
    if op.kind == "objective":
        return sum(int(info["attempts"]) for _, info in result[:useful_trials(op, result)])
    if op.kind == "linearize":
        return int(result.info["attempts"])
    if op.kind == "jvp":
        return int(op.kw["lin"].info["accepted"])
    return 0


def _own(obj, stream):
    """Mark every CUDA tensor reachable from ``obj`` as used on ``stream``, so
    the caching allocator does not hand its memory to the stream that made it
    while work queued on ``stream`` may still read it."""
    if isinstance(obj, torch.Tensor):
        if obj.is_cuda:
            obj.record_stream(stream)
    elif isinstance(obj, dict):
        for v in obj.values():
            _own(v, stream)
    elif isinstance(obj, (list, tuple)):
        for v in obj:
            _own(v, stream)
    elif isinstance(obj, _Lin):
        _own(vars(obj), stream)


# ------------------------------------------------------------------ scheduler
class Server:
    """Continuous batching of job operations on one GPU.

    ``executor(key, ops, lane) -> results`` runs one group of compatible ops;
    the default runs them on the GPU through the fused kernels.  Up to
    ``max_groups`` groups run at once, each in a slot (lane) of its own: the
    lane's context of the group's grid size and the lane's stream, so groups
    in different lanes overlap on the device.  A group's jobs advance as soon
    as it finishes, so a slow group delays only its own jobs.

    Grouping: whenever a lane is free, the ready jobs are grouped by operation
    key, each key's jobs sorted by operation length (a batch runs as long as
    its longest member) and cut into batches of at most ``max_batch``; while
    there are more free lanes than batches, the largest batch is halved (into
    its shorter and its longer jobs), so every free lane runs a group.  The
    oldest batches start first.

    Stream order: a job's generator code runs on the caller's current stream;
    a group's stream first waits for an event recorded on that stream after
    the group was formed (so it sees everything the jobs queued, such as the
    Gram-Schmidt sweep that wrote the next Krylov vector), and the caller's
    stream waits for the group's completion event before the jobs read the
    results, whose memory is then marked as in use on the caller's stream.
    """

    def __init__(self, hbm_budget=200e9, nu=1 / 40, eta=1 / 40, b0=(0.0, 0.1), max_batch=64, executor=None,
                 max_groups=4, bypass_wait_s=30.0, keep_samples=False, rev_overlap=False):
        self.executor = executor
        # reverse passes of the lanes' gradients on one stream: the lanes already
        # run concurrently, and a second stream per lane measured slower
        self.rev_overlap = rev_overlap
        self.budget = hbm_budget
        self.nu, self.eta, self.b0 = nu, eta, b0
        self.max_batch = max_batch
        self.max_groups = max_groups
        self.bypass_wait_s = bypass_wait_s
        self.keep_samples = keep_samples   # keep each job's first operation of each kind (unbatched_estimate)
        self._pds = {}
        self._streams = {}
        self.ticks = 0
        self.batched_ops = 0
        self.rk_steps = 0

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
            return _batch_linearize(ops, pd, lane)
        if kind == "jvp":
            return _batch_jvp(ops, pd, lane)
        raise ValueError(kind)

    def _work(self, key, ops, lane, ready):
        """Runs in a worker thread; returns (results, completion event)."""
        if self.executor is not None:
            return self.executor(key, ops, lane), None
        s = self._streams[lane]
        mhd.context(self.pd(ops[0].n), lane).set_rev_overlap(self.rev_overlap)
        with torch.cuda.stream(s):
            s.wait_event(ready)
            out = self._run_group(key, ops, lane)
            done = torch.cuda.Event()
            done.record(s)
        return out, done

    def _admit(self, waiting, now, used):
        """Jobs to admit now: every arrived job whose memory fits, in arrival
        order; a job that does not fit lets later ones pass unless it has
        waited more than bypass_wait_s (then it holds the queue until memory
        frees up).  Returns (admitted, rejected)."""
        admitted, rejected = [], []
        for j in list(waiting):
            if j.arrival > now:
                break
            if j.mem > self.budget:
                rejected.append(j)
                waiting.remove(j)
                continue
            if used + j.mem <= self.budget:
                admitted.append(j)
                waiting.remove(j)
                used += j.mem
            elif now - j.arrival > self.bypass_wait_s:
                break
        return admitted, rejected

    def _batches(self, ready, free):
        """(key, jobs) batches for the ``free`` lanes, oldest first."""
        if free <= 0 or not ready:
            return []
        groups = {}
        for j in ready:
            groups.setdefault(j.pending.group_key(), []).append(j)
        parts = []
        for key, js in groups.items():
            js = sorted(js, key=lambda j: (j.pending.length(), j.arrival))
            parts += [(key, js[i:i + self.max_batch]) for i in range(0, len(js), self.max_batch)]
        while len(parts) < free:
            big = max(parts, key=lambda p: len(p[1]))
            if len(big[1]) < 2:
                break
            parts.remove(big)
            h = len(big[1]) // 2
            parts += [(big[0], big[1][:h]), (big[0], big[1][h:])]
        parts.sort(key=lambda p: min(j.arrival for j in p[1]))
        return parts[:free]

    def run(self, arrivals, concurrent=True, log=None):
        """arrivals: list of (t_arrival_seconds, kind, make_generator, mem_bytes, n).
        Returns the finished jobs and summary statistics; `log(msg)`, if given,
        hears of every finished job and, every 10 s, of the groups run so far.
        A job whose memory estimate exceeds the budget is rejected at once
        (error "exceeds the HBM budget")."""
        from concurrent.futures import FIRST_COMPLETED, ThreadPoolExecutor, wait
        t0 = time.perf_counter()
        waiting = [Job(i, k, t, g, m, n) for i, (t, k, g, m, n) in enumerate(sorted(arrivals, key=lambda a: a[0]))]
        active, done = [], []
        used, peak, last_log = 0.0, 0.0, 0.0
        slots = self.max_groups if concurrent else 1
        main = torch.cuda.current_stream() if self.executor is None else None
        inflight = {}
        busy = set()
        pool = ThreadPoolExecutor(max_workers=slots)

        def finish(j, now, result=None, error=None):
            nonlocal used
            j.result, j.error, j.finish = result, error, now
            done.append(j)
            used -= j.mem
            if j in active:
                active.remove(j)
            if log is not None:
                state = "done" if error is None else f"FAILED: {error}"
                log(f"job {j.id} ({j.kind}, N={j.n}) {state} at {now:.1f} s, latency {now - j.arrival:.1f} s")

        def advance(j, value, now, first=False):
            try:
                j.pending = next(j.gen) if first else j.gen.send(value)
            except StopIteration as e:
                finish(j, now, result=e.value)
            except Exception as e:   # the job's own failure ends that job, not the server
                finish(j, now, error=repr(e))

        try:
            while waiting or active:
                now = time.perf_counter() - t0
                adm, rej = self._admit(waiting, now, used)
                for j in rej:
                    j.start = now
                    j.gen = None
                    used += j.mem
                    finish(j, now, error=f"job {j.id} needs {j.mem:.3g} bytes, which exceeds the HBM budget "
                                         f"of {self.budget:.3g} bytes")
                for j in adm:
                    j.start = now
                    j.gen = j.gen()
                    used += j.mem
                    peak = max(peak, used)
                    active.append(j)
                    advance(j, None, now, first=True)
                # form groups from the jobs whose next operation is not running
                for key, part in self._batches([j for j in active if not j.in_flight], slots - len(inflight)):
                    lane = 0
                    while lane in busy:
                        lane += 1
                    busy.add(lane)
                    ev = None
                    if self.executor is None:
                        if lane not in self._streams:
                            self._streams[lane] = torch.cuda.Stream()
                        ev = torch.cuda.Event()
                        ev.record(main)
                    for j in part:
                        j.in_flight = True
                    fut = pool.submit(self._work, key, [j.pending for j in part], lane, ev)
                    inflight[fut] = (key, part, lane)
                    self.ticks += 1
                    self.batched_ops += len(part)
                # the next arrival still in the future (an arrived job that does not fit
                # waits for a group to finish, which frees no memory by itself but ends jobs)
                now = time.perf_counter() - t0
                upcoming = [j.arrival for j in waiting if j.arrival > now]
                if not inflight:
                    if waiting:
                        time.sleep(max(0.0, min(0.01, (upcoming[0] if upcoming else now) - now)))
                    continue
                timeout = max(0.0, upcoming[0] - now) if upcoming else None
                fin, _ = wait(list(inflight), timeout=timeout, return_when=FIRST_COMPLETED)
                now = time.perf_counter() - t0
                for fut in fin:
                    key, part, ln = inflight.pop(fut)
                    busy.discard(ln)
                    for j in part:
                        j.in_flight = False
                    try:
                        res, ev = fut.result()
                    except Exception as e:   # a failed group ends its jobs
                        for j in part:
                            finish(j, now, error=f"group {key}: {e!r}")
                        continue
                    if ev is not None:
                        main.wait_event(ev)
                        _own(res, main)
                    for j, r in zip(part, res):
                        st = op_rk_steps(j.pending, r)
                        self.rk_steps += st
                        k = j.pending.kind
                        j.op_counts[k] = j.op_counts.get(k, 0) + 1
                        j.op_steps[k] = j.op_steps.get(k, 0) + st
                        if self.keep_samples and k not in j.first_op:
                            j.first_op[k] = (j.pending, st)
                        advance(j, r, now)
                if log is not None and now - last_log >= 10.0:
                    last_log = now
                    log(f"{now:.1f} s: {self.ticks} groups run, {len(active)} active jobs, {len(waiting)} waiting")
        finally:
            pool.shutdown(wait=True)
        if self.executor is None:
            torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        ok = [j for j in done if j.error is None]
        lat = np.array([j.finish - j.arrival for j in ok]) if ok else np.array([np.nan])
        return done, {"jobs": len(done), "failed_jobs": len(done) - len(ok), "wall_s": wall,
                      "throughput_jobs_per_s": len(ok) / wall,
                      "p50_latency_s": float(np.percentile(lat, 50)), "p99_latency_s": float(np.percentile(lat, 99)),
                      "ticks": self.ticks, "mean_batch": self.batched_ops / max(1, self.ticks),
                      "rk_steps": self.rk_steps, "rk_steps_per_s": self.rk_steps / wall,
                      "peak_resident_bytes": peak, "hbm_budget_bytes": self.budget}


def unbatched_estimate(server, done):
    """The unbatched GPU path for the same jobs: each job's operations run one
    at a time, alone on the GPU.  Each job's first operation of each kind
    (kept by a Server made with keep_samples=True) is run again alone and
    timed, scaled by the stepper steps of all its operations of that kind.
    Returns (seconds, stepper steps) summed over the jobs that finished; each
    such job's own share is left in ``job.unbatched_s``.  A job's operations
    run one after another however they are served, so the served wall time is
    at least the longest job's share: the served / unbatched ratio cannot
    exceed (seconds) / max(unbatched_s)."""
    total_s, total_steps = 0.0, 0
    for j in done:
        if j.error is not None:
            continue
        j.unbatched_s = 0.0
        for kind, (op, st) in j.first_op.items():
            # one at a time, an operation takes the library's fastest form (reverse
            # passes with the next segment recomputed on a second stream)
            mhd.context(server.pd(op.n), 0).set_rev_overlap(True)
            # a line-search window runs as the one-at-a-time search would: its
            # tries one by one, up to the first that lowers |f|
            if kind == "objective":
                res = server._run_group(op.group_key(), [op], 0)[0]
                singles = [Op("objective", op.n, adaptive=op.kw["adaptive"], f_norm=op.kw["f_norm"], trials=[t])
                           for t in op.kw["trials"][:useful_trials(op, res)]]
            else:
                singles = [op]
            server._run_group(op.group_key(), [singles[0]], 0)   # warm: workspace of a batch of one
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for o in singles:
                server._run_group(op.group_key(), [o], 0)
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            j.unbatched_s += dt * (j.op_steps[kind] / max(st, 1))
            total_steps += j.op_steps[kind]
        total_s += j.unbatched_s
    return total_s, total_steps


def make_workload(count, rate=4.0, seed=7, sizes=(128, 256), T_range=(2.0, 8.0), adjoint_frac=0.7,
                  descent_iters=64, newton_iters=5, inner=64, dt=0.004, atol=1e-8, adaptive_budget=(64, 256),
                  dump_dir=None):
    """Poisson arrivals at `rate` per second of mixed jobs (the serving workload).

    ``adaptive_budget`` = (max_steps_per_checkpoint, checkpoints) of the Newton
    jobs' RK4(3) solves.  The seeded N=256 states take steps of ~1e-4 T at
    atol 1e-8, so a period up to 8 needs up to ~9000 attempts: the default
    budget of 16384 attempts completes them, where the reference's 128 x 32 =
    4096 (newton.py:75-79) stops early."""
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
            gen = (lambda f=f, T=T, n=n, i=i: newton_job(f, T, 0.0, n, adaptive, inner, newton_iters,
                                                         dump_dir=dump_dir, tag=f"job{i}"))
            mem = job_memory("newton", n, inner=inner, budget=adaptive_budget[0] * adaptive_budget[1])
        out.append((t, kind, gen, mem, n))
    return out
