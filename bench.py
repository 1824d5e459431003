"""Benchmark of the periodic-orbit hot path on one MI355X (FP64).

Headline workload (the Newton-GMRES configuration): N=128, nu=eta=1/40,
b0=(0,0.1), F=-4cos(4y); guess = synthetic spectrum G(128, seed 0) evolved
8192 Lawson-RK4 steps of h=0.005; T=4.3, sx=0.1; adaptive Lawson RK4(3) with
atol=1e-8, h0=1e-2 (budget 64 x 64 attempts).  One "step" is one Arnoldi step
of GMRES with Krylov dimension 64: the JVP of objective_RPO through the whole
adaptive trajectory (the matvec) and the Gram-Schmidt sweep of its result.
value = GMRES matvecs/s (Arnoldi steps/s); the JVP alone is reported beside it.

Also reported: one whole Newton iteration, Lawson RK4 steps/s at N=128 on four
seeds (batch 4 and 1, with the oracle's rate beside it) and at N=256 (batch 1
and batch 64), the roofline of the dominant kernel of the matvec (the tangent
row kernel of the replay), measured live with HIP events, the batched
candidates, the other steppers, the gradient, the first arrivals of the
serving workload at their full per-job sizes, and the CPU oracle's matvec
rate on a bounded sample.

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# HBM bytes per launch of the two roofline kernels, from rocprofv3 --pmc
# FETCH_SIZE and WRITE_SIZE passes over `bench.py --probe-roofline[-batched]`
# (KiB per dispatch scaled by a 64 MiB device copy: FETCH x 2047.3, WRITE x
# 1024); the runs are summarised in profiles/pmc_traffic_tangent_row_kernel_n128_b1.md
# (the tangent replay's instance of the row kernel) and
# profiles/pmc_batched_row_kernel_n256_b64.md.
TRAFFIC_ROW_T_N128_B1 = 848.0 * 2047.3 + 352.0 * 1024
TRAFFIC_ROW_N256_B64 = 90205.2 * 2047.3 + 44032.0 * 1024
METRIC = "FP64 RK steps/s at N=256 + GMRES matvecs/s; p50/p99 RPO solve latency"


def zbytes(n):
    """Z(N): bytes of one complex half-spectrum field."""
    return n * (n // 2 + 1) * 16


def build_guess(n, seed, steps, h, pd):
    import torch
    from ecs_mhd_2d_amd import mhd
    from ecs_mhd_2d_amd.synthetic import spectrum
    ctx = mhd.context(pd)
    f = torch.as_tensor(spectrum(n, seed)).cuda()
    f = ctx.lawson_rk4(f, h, steps)
    return torch.fft.irfft2(f, s=(n, n)).contiguous()


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline_matvec(fields_np, T, sx, lin_info, sample_attempts=64, runs=5):
    """The oracle's JVP (the CPU restatement of the reference's lines, FP64) on
    the first ``sample_attempts`` attempts of the same N=128 schedule, on the
    host's cores: numpy's single-threaded FFT and scipy.fft with one worker
    per core available to this process; each timed ``runs`` times after one
    warm-up run, median kept, scaled to the whole trajectory.  The faster
    configuration is the baseline; both are reported."""
    from oracle import mhd_oracle as O
    n = fields_np.shape[-1]
    pd = O.make_params(n, 1 / 40, 1 / 40)
    acc = lin_info["acc"][: lin_info["attempts"]]
    hs = lin_info["hs"][: lin_info["attempts"]]
    sched = [(float(h), bool(a)) for h, a in zip(hs, acc)]
    S = min(sample_attempts, len(sched))
    f0 = np.fft.rfft2(fields_np)
    rng = np.random.default_rng(5)
    df0 = np.fft.rfft2(rng.standard_normal((2, n, n)))
    cores = len(os.sched_getaffinity(0))
    lim = os.environ.get("OMP_NUM_THREADS")
    if lim and lim.isdigit():
        cores = min(cores, int(lim))
    res = {}
    for workers in sorted({1, cores}):
        O.set_fft_workers(workers)
        try:
            O.lawson_rk43_jvp(f0, df0, pd, T, 0.1, sched[:S])
            ts = []
            for _ in range(runs):
                t0 = time.perf_counter()
                O.lawson_rk43_jvp(f0, df0, pd, T, 0.1, sched[:S])
                ts.append(time.perf_counter() - t0)
        finally:
            O.set_fft_workers(1)
        res[workers] = float(np.median(ts)) * len(sched) / S
    best = min(res, key=res.get)
    return {"value": 1.0 / res[best], "unit": "matvecs/s", "cores": best, "kind": "port",
            "sample": (f"oracle JVP (numpy/scipy FP64 restatement) through the first {S} of {len(sched)} attempts "
                       f"of the same N=128 schedule, median of {runs} runs after a warm-up, scaled to the whole "
                       f"trajectory; {'numpy FFT, 1 thread' if best == 1 else f'scipy.fft, {best} workers'} "
                       f"was the faster: " + ", ".join(f"{w} thread(s) {1.0 / t:.4f}/s" for w, t in res.items())
                       + f"; host {_cpu_model()}, {len(os.sched_getaffinity(0))} cores available")}


def probe():
    """Short fixed workload for kernel profiles: Lawson RK4 at N=128 (batch 1),
    N=256 (batch 1 and 64), and one RK4(3) JVP at N=128 (T=0.3, atol=1e-8)."""
    import torch
    from ecs_mhd_2d_amd import loss_functions as LF, mhd, utils
    from ecs_mhd_2d_amd.synthetic import spectrum
    for n, B, steps in ((128, 1, 32), (256, 1, 16), (256, 64, 8)):
        pd = mhd.make_param_dict(n, 1 / 40, 1 / 40)
        f = torch.as_tensor(np.stack([spectrum(n, s) for s in range(B)])).cuda()
        mhd.context(pd).lawson_rk4(f, 0.004, steps)
        torch.cuda.synchronize()
        print("probe rk4", n, B, steps, flush=True)
    n = 128
    pd = mhd.make_param_dict(n, 1 / 40, 1 / 40)
    pd.update({"shift_reflect_ny": 0, "rot": False,
               "adaptive_dict": {"atol": 1e-8, "checkpoints": 8, "max_steps_per_checkpoint": 64}})
    inp = {"fields": torch.as_tensor(np.fft.irfft2(spectrum(n, 0), s=(n, n))).cuda(), "T": 0.3, "sx": 0.1}
    lin = LF.Linearization(inp, pd, "Lawson_RK43")
    v = torch.randn(2 * n * n + 2, dtype=torch.float64, device="cuda")
    lin.jvp(utils.unflatten(v, n))
    torch.cuda.synchronize()
    print("probe rk43 jvp", lin.info["accepted"], lin.info["rejected"], flush=True)


def config0_leg(cpu=True):
    """Lawson RK4 at N=128 on G(128, s), s = 0..3, h = 0.005, 1024 steps, as
    one batch of 4 and as one trajectory; with the oracle's rate on a bounded
    sample of one trajectory (1 thread) beside it."""
    import torch
    from ecs_mhd_2d_amd import mhd
    from ecs_mhd_2d_amd.synthetic import spectrum
    n, h, steps = 128, 0.005, 1024
    pd = mhd.make_param_dict(n, 1 / 40, 1 / 40, (0.0, 0.1))
    ctx = mhd.context(pd)
    f4 = torch.as_tensor(np.stack([spectrum(n, s) for s in range(4)])).cuda()
    ctx.lawson_rk4(f4, h, 8)
    ctx.lawson_rk4(f4[:1], h, 8)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ctx.lawson_rk4(f4, h, steps)
    torch.cuda.synchronize()
    t4 = time.perf_counter() - t0
    t0 = time.perf_counter()
    ctx.lawson_rk4(f4[:1], h, steps)
    torch.cuda.synchronize()
    t1 = time.perf_counter() - t0
    out = {"workload": "N=128, G(128, s) s=0..3, Lawson RK4 h=0.005, 1024 steps (T=5.12)",
           "batch4_steps_per_s": 4 * steps / t4, "batch4_rhs_per_s": 16 * steps / t4,
           "batch1_steps_per_s": steps / t1, "batch1_rhs_per_s": 4 * steps / t1}
    if cpu:
        from oracle import mhd_oracle as O
        opd = O.make_params(n, 1 / 40, 1 / 40)
        f = spectrum(n, 0)
        O.lawson_rk4(f, 2 * h, 2, opd)  # warm-up
        k, t = 0, 0.0
        while t < 8.0 and k < steps:
            t0 = time.perf_counter()
            f = O.lawson_rk4(f, 16 * h, 16, opd)
            t += time.perf_counter() - t0
            k += 16
        out["cpu_oracle_steps_per_s"] = k / t
        out["cpu_sample"] = f"oracle Lawson RK4 on G(128, 0), the first {k} of 1024 steps ({t:.1f} s), 1 thread"
    return out


def steppers_leg():
    """The other integrators of mismatch_RPO (classical RK4, Lawson RK6,
    two-derivative RK4), primal, JVP and VJP (reverse pass with checkpoints
    every 64 steps), at N=128 and N=256, batch 1, 256 steps of h = 1/256 on
    G(N, 0)."""
    import torch
    from ecs_mhd_2d_amd import mhd
    from ecs_mhd_2d_amd.synthetic import spectrum
    out = {"workload": "batch 1, 256 steps of h=1/256 on G(N, 0); JVP along G(N, 1), dT=1e-3; "
                       "VJP of the final state along G(N, 2), checkpoint every 64"}
    for n in (128, 256):
        ctx = mhd.context(mhd.make_param_dict(n, 1 / 40, 1 / 40, (0.0, 0.1)))
        f = torch.as_tensor(spectrum(n, 0))[None].cuda()
        df = torch.as_tensor(spectrum(n, 1))[None].cuda()
        fb = torch.as_tensor(spectrum(n, 2))[None].cuda()
        runs = {"rk4": lambda s: ctx.rk4(f, 1 / 256, s), "rk4_jvp": lambda s: ctx.rk4_jvp(f, df, 1 / 256, 1e-3, s),
                "lawson_rk6": lambda s: ctx.lawson_rk6(f, 1 / 256, s),
                "lawson_rk6_jvp": lambda s: ctx.lawson_rk6_jvp(f, df, 1 / 256, 1e-3, s),
                "tdrk4": lambda s: ctx.tdrk4(f, 1 / 256, s), "tdrk4_jvp": lambda s: ctx.tdrk4_jvp(f, df, 1 / 256, 1e-3, s),
                "rk4_vjp": lambda s: ctx.rk4_vjp(f, fb, 1 / 256, s, 64),
                "lawson_rk6_vjp": lambda s: ctx.lawson_rk6_vjp(f, fb, 1 / 256, s, True, 64),
                "tdrk4_vjp": lambda s: ctx.tdrk4_vjp(f, fb, 1 / 256, s, 64)}
        for name, fn in runs.items():
            fn(4)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            fn(256)
            torch.cuda.synchronize()
            out[f"{name}_steps_per_s_n{n}"] = 256 / (time.perf_counter() - t0)
    return out


def transpose_leg(inp, pd, cots):
    """The transpose matvec of newton.py with use_transpose=True (newton.py:167-170):
    the VJP of objective_RPO through the headline's RK4(3) trajectory, as the
    vjp function of one primal applied per Krylov step (Linearization.vjp:
    the checkpointed reverse pass along the stored schedule), and as a
    standalone objective_RPO_vjp call (which solves the primal first)."""
    import torch
    from ecs_mhd_2d_amd import loss_functions as LF
    lin = LF.Linearization(inp, pd, "Lawson_RK43")
    lin.vjp(cots[0])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for c in cots:
        lin.vjp(c)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / len(cots)
    t0 = time.perf_counter()
    LF.objective_RPO_vjp(inp, cots[0], pd, "Lawson_RK43")
    torch.cuda.synchronize()
    standalone = time.perf_counter() - t0
    return {"workload": "VJP of objective_RPO (mode Lawson_RK43) through the headline trajectory, checkpoint every 64",
            "transpose_matvecs_per_s": 1.0 / dt, "seconds_per_vjp": dt, "standalone_objective_vjp_s": standalone}


def newton_leg(inp, pd):
    """One full Newton-GMRES iteration on the headline guess (Krylov dimension 64:
    objective, linearisation, 65 matvecs, line search, dealias)."""
    import torch
    from ecs_mhd_2d_amd import solvers
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    new, log = solvers.newton_gmres_update(inp, pd, "Lawson_RK43", inner=64)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return {"workload": "one Newton-GMRES iteration, inner 64, same guess as the headline", "wall_s": wall,
            "fwall_s": log["fwall"], "gmreswall_s": log["gmreswall"], "line_search_s": log["lswall"],
            "line_search_window": log["line_search_window"],
            "gmres_matvecs_per_s": log["matvecs"] / log["gmreswall"],
            "f_norm": log["f_norm"], "gmres_rel_res": log["gmres_rel_res"], "damp": log["damp"],
            "rel_err": log["rel_err"],
            "arnoldi_residual_estimate_last": log["krylov_residuals"][-1],
            "qtb_residual_last": log["krylov_residuals_qtb"][-1],
            "basis_orthogonality_loss": log["orthogonality_loss"]}


def newton_default_leg(inp, pd):
    """One Newton-GMRES iteration in the reference's default Newton setting
    (newton.py:70,84-87: mode TDRK4, use_transpose=True, inner 16, line
    search), whose one GMRES solve the reference times at 16.1 s on an H100
    (benchmarking.m:43; its trajectory length is not recorded): here on the
    headline guess (N=128, T=4.3, sx=0.1) with 32 checkpoints of 32 TDRK4 steps
    (newton.py:52-55), each Krylov step one JVP and one VJP of the same
    linearisation (normal_matvecs_per_s counts those pairs, the extra one for
    the final residual included)."""
    import torch
    from ecs_mhd_2d_amd import solvers
    pdt = dict(pd)
    pdt.update({"ministeps": 32, "num_checkpoints": 32})
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    new, log = solvers.newton_gmres_update(inp, pdt, "TDRK4", inner=16, use_transpose=True)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return {"workload": "one Newton-GMRES iteration, mode TDRK4, A^T A form, inner 16, 1024 steps, "
                        "same guess as the headline", "wall_s": wall, "fwall_s": log["fwall"],
            "gmreswall_s": log["gmreswall"], "line_search_s": log["lswall"],
            "normal_matvecs_per_s": log["matvecs"] / log["gmreswall"], "f_norm": log["f_norm"],
            "gmres_rel_res": log["gmres_rel_res"], "damp": log["damp"],
            "reference_published_h100_gmres_s": 16.1}


def hookstep_leg(inp, pd):
    """One iteration of the hookstep Newton solver (newton_hookstep_v2.py:95-121:
    mode Lawson_RK4, inner 64, trust region s = 100) on the headline guess, with
    32 checkpoints of 32 Lawson-RK4 steps (newton_hookstep_v2.py:49-58): 64
    JVPs with reorthogonalised Arnoldi, the SVD trust-region solve and its tries."""
    import torch
    from ecs_mhd_2d_amd import solvers
    pdh = dict(pd)
    pdh.update({"ministeps": 32, "num_checkpoints": 32})
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    new, s_new, log = solvers.newton_hookstep_update(inp, pdh, "Lawson_RK4", 100.0, inner=64)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return {"workload": "one hookstep Newton iteration, mode Lawson_RK4, inner 64, s = 100, 1024 steps, "
                        "same guess as the headline", "wall_s": wall, "fwall_s": log["fwall"],
            "gmreswall_s": log["gmreswall"], "f_norm": log["f_norm"], "rel_res": log["rel_res"],
            "s": s_new, "rel_err": log["rel_err"]}


def floquet_leg(inp, pd, iters=2):
    """Block power iteration for Floquet multipliers (floquet.py:57-99: mode TDRK4,
    32 tangents per block) on the headline guess with 32 checkpoints of 32 TDRK4
    steps: each iteration one batched JVP of the 32 tangents and a QR of the block."""
    import torch
    from ecs_mhd_2d_amd import floquet
    pdf = dict(pd)
    pdf.update({"ministeps": 32, "num_checkpoints": 32})
    floquet.block_power_iteration(inp, pdf, block_size=32, maxit=1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    floquet.block_power_iteration(inp, pdf, block_size=32, maxit=iters)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    # the call also integrates the orbit once and applies the block once more for R
    return {"workload": f"Floquet block power iteration, 32 tangents, TDRK4, 1024 steps, {iters} iterations "
                        "(+ the orbit and the final R = Q^T J Q), same guess as the headline",
            "wall_s": wall, "block_jvps": iters + 1, "tangent_steps_per_s": 32 * 1024 * (iters + 1) / wall}


def guess_generation_leg():
    """The reference's published guess-generation workload (generate_turbulence.py:22-151,
    timings in benchmarking.m:11-59): from 10 N(0,1) random fields at N=128
    (b0 = (0, 0.3)), a transient of 8192 Lawson-RK6 steps of dt = 1/256, then 256
    snapshots 64 steps apart (16384 steps), then the 256 x 256 recurrence diagram."""
    import torch
    from ecs_mhd_2d_amd import mhd, turbulence
    n = 128
    pd = mhd.make_param_dict(n, 1 / 40, 1 / 40, (0.0, 0.3))
    ctx = mhd.context(pd)
    rng = np.random.default_rng(111111)
    f = torch.as_tensor(10.0 * rng.standard_normal((2, n, n))).cuda()
    f = (pd["mask"] * torch.fft.rfft2(f)).contiguous()
    ctx.lawson_rk6(f, 1 / 256, 64, masked=True)      # warm-up (autotuner, workspace)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    f = ctx.lawson_rk6(f, 1 / 256, 8192, masked=True)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    fs = torch.empty((256, 2, n, n // 2 + 1), dtype=torch.complex128, device="cuda")
    for i in range(256):
        f = ctx.lawson_rk6(f, (64 / 256) / 64, 64, masked=True)
        fs[i] = f
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    dist = turbulence.recurrence_diagram(fs)
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    assert torch.isfinite(dist).all()
    return {"workload": "N=128: 8192 Lawson-RK6 transient steps (dt 1/256), 256 snapshots x 64 steps, "
                        "256 x 256 recurrence diagram",
            "transient_s": t1 - t0, "transient_steps_per_s": 8192 / (t1 - t0),
            "turbulence_s": t2 - t1, "turbulence_steps_per_s": 16384 / (t2 - t1), "recurrence_s": t3 - t2,
            "reference_published_h100_s": {"transient": 3.13, "turbulence": 5.67, "recurrence": 0.669}}


def gradient_leg(reps=3):
    """Adjoint-descent gradient: loss_RPO value and gradient at N=256, T=4.3, 1088
    Lawson-RK4 steps (17 x 64), reverse pass with checkpoint/recompute every 64."""
    import torch
    from ecs_mhd_2d_amd import loss_functions as LF, mhd
    from ecs_mhd_2d_amd.synthetic import fields as synth_fields
    n = 256
    pd = mhd.make_param_dict(n, 1 / 40, 1 / 40)
    pd.update({"num_checkpoints": 17, "ministeps": 64, "shift_reflect_ny": 0, "rot": False})
    inp = {"fields": torch.as_tensor(synth_fields(n, 2)).cuda(), "T": 4.3, "sx": 0.1}
    LF.loss_RPO_value_and_grad(inp, pd)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        LF.loss_RPO_value_and_grad(inp, pd)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    return {"workload": "loss_RPO value+grad, N=256, T=4.3, 1088 steps, checkpoint every 64",
            "seconds_per_gradient": dt, "gradient_evals_per_s": 1.0 / dt,
            "rk4_steps_per_s_fwd_plus_bwd": 1088 / dt}


def batched_leg():
    """64 RPO candidates of uneven size (N in {64,128,256} at 16:32:16, T ~ U[2,8],
    sx ~ U[-0.5,0.5], spectra G(N, i)), adaptive RK4(3) atol 1e-8 over each period:
    accepted RK steps/s batched (grouped by N, groups concurrent) vs one at a time."""
    import torch
    from ecs_mhd_2d_amd import batch
    cands = batch.make_candidates()
    warm = [dict(c, T=0.05) for c in cands]
    batch.run_batched(warm)
    batch.run_unbatched(warm[:3])
    rb, tb = batch.run_batched(cands)
    ru, tu = batch.run_unbatched(cands)
    sb, su = batch.accepted_steps(rb) / tb, batch.accepted_steps(ru) / tu
    same = all(torch.equal(a[0], b[0]) for a, b in zip(rb, ru))
    fevals = int(sum(info["fevals"] for _, info in rb))   # RHS evaluations, 1 + 4 per attempt
    return {"workload": "64 candidates, N in {64,128,256} (16/32/16), T~U[2,8], RK4(3) atol 1e-8",
            "accepted_steps": batch.accepted_steps(rb), "rhs_evals": fevals, "batched_s": tb, "unbatched_s": tu,
            "batched_steps_per_s": sb, "unbatched_steps_per_s": su, "speedup": sb / su,
            "batched_rhs_per_s": fevals / tb, "unbatched_rhs_per_s": fevals / tu,
            "bitwise_equal_to_unbatched": bool(same)}


def serving_leg(jobs, dump_dir=None, log_to_stderr=False):
    """The serving workload at its stated per-job size: Poisson arrivals at 4 req/s
    (rng seed 7), 70% adjoint descent (64 ADAM steps, Lawson RK4 with steps =
    round(T/0.004) rounded up to a multiple of 64) / 30% Newton-GMRES (Krylov 64,
    5 iterations, RK4(3) atol 1e-8 with up to 64 x 256 attempts), N in {128, 256},
    T ~ U[2, 8], 200 GB budget;
    the first ``jobs`` arrivals of that stream.  Beside the served throughput,
    the unbatched GPU path for the same jobs (each job's operations one at a time,
    alone on the GPU) estimated by timing each job's first operation of each kind
    alone, scaled by its stepper steps (serve.unbatched_estimate)."""
    from ecs_mhd_2d_amd import serve
    w = serve.make_workload(4, rate=100.0, seed=1, sizes=(128, 256), T_range=(0.3, 0.4), descent_iters=1,
                            newton_iters=1, inner=2, atol=1e-6, adaptive_budget=(64, 16))
    serve.Server().run(w)   # warm-up: contexts, workspaces, autotuner
    wl = serve.make_workload(jobs, rate=4.0, seed=7, sizes=(128, 256), T_range=(2.0, 8.0), descent_iters=64,
                             newton_iters=5, inner=64, atol=1e-8, adaptive_budget=(64, 256), dump_dir=dump_dir)
    log = (lambda m: print(m, file=sys.stderr, flush=True)) if log_to_stderr else None
    srv = serve.Server(keep_samples=True)
    done, stats = srv.run(wl, log=log)
    kinds = [j.kind for j in done]
    ub_s, ub_steps = serve.unbatched_estimate(srv, done)
    stats["workload"] = (f"first {jobs} arrivals: Poisson 4/s, 70% adjoint descent (64 ADAM steps, dt 0.004) / "
                         "30% Newton (5 iterations, Krylov 64, RK4(3) atol 1e-8), N in {128,256}, T~U[2,8], "
                         "200 GB budget")
    stats["adjoint_jobs"] = kinds.count("adjoint")
    stats["newton_jobs"] = kinds.count("newton")
    stats["jobs_detail"] = [{"id": j.id, "kind": j.kind, "n": j.n, "latency_s": (j.finish - j.arrival),
                             "unbatched_s": getattr(j, "unbatched_s", None), "error": j.error} for j in done]
    stats["unbatched_estimate_s"] = ub_s
    stats["unbatched_rk_steps_per_s"] = ub_steps / ub_s if ub_s > 0 else None
    stats["served_over_unbatched"] = (stats["rk_steps_per_s"] / (ub_steps / ub_s)) if ub_s > 0 else None
    # a job's operations run one after another, so the wall time is at least the longest
    # job's own unbatched time: the ratio above is bounded by this
    longest = max((getattr(j, "unbatched_s", 0.0) for j in done if j.error is None), default=0.0)
    stats["longest_job_unbatched_s"] = longest
    stats["served_over_unbatched_bound"] = ub_s / longest if longest > 0 else None
    return stats


def probe_reverse():
    """Short fixed workload for a kernel trace of the reverse passes: the RK4(3)
    VJP at N=128, batch 1, along the schedule of T = 1 on G(128, 0) (atol 1e-8,
    checkpoints every 64 accepted steps), twice with the reverse pass on one
    stream, then twice with the next segment's forward pass on the context's
    second stream; prints each call's wall time."""
    import torch
    from ecs_mhd_2d_amd import mhd
    from ecs_mhd_2d_amd.synthetic import spectrum
    n = 128
    pd = mhd.make_param_dict(n, 1 / 40, 1 / 40)
    ctx = mhd.context(pd)
    x0 = torch.as_tensor(spectrum(n, 0)).cuda()[None]
    g = torch.as_tensor(spectrum(n, 2)).cuda()[None]
    _, infos = ctx.lawson_rk43(x0, [1.0], 0.01, 1e-8, 256, 64)
    for ov in (False, True):
        ctx.set_rev_overlap(ov)
        for rep in range(2):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            ctx.lawson_rk43_vjp(x0, g, [1.0], infos, ckpt=64)
            torch.cuda.synchronize()
            print(f"probe-reverse overlap={int(ov)} rep={rep} accepted={infos[0]['accepted']} "
                  f"wall_ms={1e3 * (time.perf_counter() - t0):.3f}", flush=True)


def probe_roofline(batched=False, kernel=None):
    """The matvec's dominant kernel (tangent row kernel of the adaptive tangent
    replay, N=128, batch 1) or,
    with ``batched``, the batched roofline kernel (row kernel, N=256, batch
    64), or ``kernel`` = (N, batch, which) with which as in ecs_time_kernel
    (0 column, 1 row, 2 tangent column, 3 tangent row kernel, 4 tangent row
    kernel of the tangent replay): 50 launches
    through ecs_time_kernel; then a plain device copy of 64 MiB (read 64 MiB,
    write 64 MiB), whose counter readings scale FETCH_SIZE / WRITE_SIZE to
    bytes."""
    import torch
    from ecs_mhd_2d_amd import _lib, _runtime, mhd
    n, B, which = kernel if kernel else ((256, 64, 1) if batched else (128, 1, 4))
    pd = mhd.make_param_dict(n, 1 / 40, 1 / 40)
    ctx = mhd.context(pd)
    ms = ctypes.c_double()
    lib = _lib.load()
    _lib.check(lib.ecs_time_kernel(ctx.handle, B, which, 50, ctypes.byref(ms), _runtime._stream()),
               "ecs_time_kernel")
    a = torch.ones(8 * 1024 * 1024, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print("probe-roofline", ms.value, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true", help="skip the batched-candidates and serving legs")
    ap.add_argument("--serving-full", type=int, default=0, metavar="JOBS",
                    help="run only the serving workload at its full size with JOBS jobs and print its statistics")
    ap.add_argument("--serving-jobs", type=int, default=4, metavar="JOBS",
                    help="first arrivals of the serving workload run by the default run's serving leg "
                         "(0: the leg is skipped)")
    ap.add_argument("--dump-dir", default=None,
                    help="where a failing Newton job of the serving workload writes its input")
    ap.add_argument("--probe-roofline", action="store_true",
                    help="launch only the matvec's dominant kernel (50x) and a 64 MiB copy, for rocprofv3 --pmc")
    ap.add_argument("--probe-roofline-batched", action="store_true",
                    help="the same for the batched row kernel (N=256, batch 64)")
    ap.add_argument("--probe-kernel", default="", metavar="N,BATCH,WHICH",
                    help="the same for one kernel family (WHICH: 0 column, 1 row, 2 tangent column, 3 tangent row, "
                         "4 tangent row of the tangent replay)")
    ap.add_argument("--probe", action="store_true",
                    help="run a short fixed kernel workload for rocprofv3 and exit")
    ap.add_argument("--probe-reverse", action="store_true",
                    help="run the RK4(3) reverse pass on one and on two streams for a kernel trace and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (this build targets a single MI355X); got --gpus {args.gpus}",
              file=sys.stderr)
        sys.exit(2)
    import torch
    from ecs_mhd_2d_amd import _lib, _runtime, loss_functions as LF, mhd, utils

    torch.cuda.init()
    if args.probe:
        probe()
        return
    if args.probe_reverse:
        probe_reverse()
        return
    if args.probe_roofline or args.probe_roofline_batched or args.probe_kernel:
        kernel = tuple(int(t) for t in args.probe_kernel.split(",")) if args.probe_kernel else None
        probe_roofline(batched=args.probe_roofline_batched, kernel=kernel)
        return
    if args.serving_full:
        stats = serving_leg(args.serving_full, dump_dir=args.dump_dir, log_to_stderr=True)
        print(json.dumps({"metric": "p50/p99 RPO solve latency (serving, full size)", "unit": "s",
                          "value": stats["p99_latency_s"], "higher_is_better": False, "n_gpus": 1,
                          "dtype": "f64", "data": "synthetic", "serving": stats}))
        return
    n = 128
    pd = mhd.make_param_dict(n, 1 / 40, 1 / 40, (0.0, 0.1))
    pd.update({"shift_reflect_ny": 0, "rot": False,
               "adaptive_dict": {"atol": 1e-8, "checkpoints": 64, "max_steps_per_checkpoint": 64}})
    fields = build_guess(n, 0, 8192, 0.005, pd)
    inp = {"fields": fields, "T": 4.3, "sx": 0.1}
    lin = LF.Linearization(inp, pd, "Lawson_RK43")
    info = lin.info
    gen = torch.Generator(device="cuda").manual_seed(1)

    def tangent():
        v = torch.randn(2 * n * n + 2, generator=gen, device="cuda", dtype=torch.float64)
        return utils.unflatten(v / torch.linalg.vector_norm(v), n)

    # one step = one Arnoldi step of GMRES (linalg.py:66-85): the JVP matvec through
    # the whole adaptive trajectory and the Gram-Schmidt sweep of its result,
    # on the Krylov basis started from b = flatten(objective) (Krylov dimension 64;
    # past 64 steps a new basis is started from b again)
    from ecs_mhd_2d_amd import linalg
    b = utils.flatten(lin.objective())
    A = lambda v: utils.flatten(lin.jvp(utils.unflatten(v, n)))
    inner = 64
    Q = torch.zeros((inner + 1, b.numel()), dtype=torch.float64, device="cuda")
    Hd = torch.zeros((inner, inner + 2), dtype=torch.float64, device="cuda")
    kstep = [0]

    def arnoldi_step():
        k = kstep[0] % inner
        if k == 0:
            Q.zero_()
            Q[0] = b / torch.linalg.vector_norm(b)
        w = A(Q[k]).contiguous().clone()
        linalg._mgs(Q, k, w, Hd[k])
        kstep[0] += 1

    for _ in range(args.warmup):
        arnoldi_step()
    lib = _lib.load()
    ctx = mhd.context(pd)
    torch.cuda.synchronize()
    # HIP events around the first tangent row pass of every 8th accepted step, inside the
    # timed matvecs, on their stream (up to 1024 samples)
    _lib.check(lib.ecs_replay_probe(ctx.handle, 1, 8, None, None), "ecs_replay_probe")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        arnoldi_step()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    value = args.steps / dt
    probe_ms, probe_n = ctypes.c_double(), ctypes.c_int()
    _lib.check(lib.ecs_replay_probe(ctx.handle, 0, 0, ctypes.byref(probe_ms), ctypes.byref(probe_n)),
               "ecs_replay_probe")
    # the isolated JVP (objective_RPO's Jacobian along a dict tangent), as the reference's
    # benchmarking.m times it
    tans = [tangent() for _ in range(4)]
    lin.jvp(tans[0])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for t in tans:
        lin.jvp(t)
    torch.cuda.synchronize()
    jvp_rate = len(tans) / (time.perf_counter() - t0)

    # Lawson RK4 steps/s at N=256
    from ecs_mhd_2d_amd.synthetic import spectrum
    pd256 = mhd.make_param_dict(256, 1 / 40, 1 / 40, (0.0, 0.1))
    ctx256 = mhd.context(pd256)
    # (median of three timed runs of 256 steps at batch 1, 64 steps at batch 64)
    rk4 = {}
    for B, steps in ((1, 256), (64, 64)):
        f = torch.as_tensor(np.stack([spectrum(256, s) for s in range(B)])).cuda()
        ctx256.lawson_rk4(f, 0.004, 4)
        torch.cuda.synchronize()
        rates = []
        for _ in range(3):
            t1 = time.perf_counter()
            ctx256.lawson_rk4(f, 0.004, steps)
            torch.cuda.synchronize()
            rates.append(B * steps / (time.perf_counter() - t1))
        rk4[B] = float(np.median(rates))

    # live roofline of the dominant kernel of the matvec: the tangent row kernel of the
    # adaptive tangent replay (N=128, batch 1), the instance the matvec launches, timed
    # over 200 back-to-back launches (agrees with its rocprofv3 kernel-trace duration
    # inside the matvec); the event pairs sampled inside the timed matvecs above also
    # hold the kernel's dispatch latency and the marker packets, an upper bound
    ms = ctypes.c_double()

    def timed(c, B, which, reps):
        """median over three rounds of `reps` back-to-back launches"""
        vals = []
        for _ in range(3):
            _lib.check(lib.ecs_time_kernel(c.handle, B, which, reps, ctypes.byref(ms), _runtime._stream()),
                       "ecs_time_kernel")
            vals.append(ms.value)
        return float(np.median(vals))

    row_t_ms = timed(ctx, 1, 4, 200)
    bytes_row_t = 20 * zbytes(n)   # reads 16 Z (8 primal + 8 tangent derived fields), writes 4 Z
    ach = bytes_row_t / (row_t_ms * 1e-3) / 1e9
    # the launch floor beside it: an empty kernel with the same grid and argument block
    floor_ms = timed(ctx, 1, 5, 200)
    # the same kernel family at full occupancy: primal row kernel, N=256, batch 64
    row_b_ms = timed(ctx256, 64, 1, 50)
    ach_b = 64 * 10 * zbytes(256) / (row_b_ms * 1e-3) / 1e9

    tiles = {}
    for nn, bb in ((128, 1), (256, 64)):
        arr = (ctypes.c_int * 4)()
        lib.ecs_tuned_tiles(nn, bb, arr)
        plans = (ctypes.c_int * 4)()
        lib.ecs_tuned_plans(nn, plans)
        splits = (ctypes.c_int * 2)()
        lib.ecs_tuned_splits(nn, bb, splits)
        variants = (ctypes.c_int * 4)()
        lib.ecs_tuned_variants(nn, bb, variants)
        rvariants = (ctypes.c_int * 2)()
        lib.ecs_tuned_replay_variants(nn, bb, rvariants)
        tiles[f"n{nn}_b{bb}"] = {"col_ty": arr[0], "col_tan_ty": arr[1], "row_tx": arr[2], "row_tan_tx": arr[3],
                                 "col_split": splits[0], "col_tan_split": splits[1],
                                 "variants": list(variants), "replay_variants": list(rvariants),
                                 "fft_plan_col": plans[0], "fft_plan_col_tan": plans[1], "fft_plan_row": plans[2],
                                 "fft_plan_row_tan": plans[3]}

    extras = {"config0_lawson_rk4_n128": config0_leg(cpu=not args.no_cpu)}
    if not args.no_extras:
        extras["transpose_matvec"] = transpose_leg(inp, pd, [tans[i] for i in range(3)])
        extras["newton_iteration"] = newton_leg(inp, pd)
        extras["newton_default_setting"] = newton_default_leg(inp, pd)
        extras["hookstep_iteration"] = hookstep_leg(inp, pd)
        extras["floquet"] = floquet_leg(inp, pd)
        extras["other_steppers"] = steppers_leg()
        extras["adjoint_gradient"] = gradient_leg()
        extras["guess_generation"] = guess_generation_leg()
        extras["batched_candidates"] = batched_leg()
        if args.serving_jobs:
            extras["serving"] = serving_leg(args.serving_jobs)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline_matvec(fields.cpu().numpy(), 4.3, 0.1, info)

    line = {
        "metric": METRIC, "value": value, "unit": "matvecs/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * dt / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "jvp_matvecs_per_s": jvp_rate,
        "config": {"workload": "GMRES matvec: one Arnoldi step of GMRES (JVP of objective_RPO through the "
                               "adaptive Lawson RK4(3) trajectory + Gram-Schmidt sweep) on one RPO guess, "
                               "N=128, T=4.3, sx=0.1, atol=1e-8, Krylov dim 64",
                   "n": 128, "attempts": info["attempts"], "accepted": info["accepted"],
                   "rejected": info["rejected"], "parallelism": "single GPU"},
        "rk4_steps_per_s_n256_b1": rk4[1], "rk4_steps_per_s_n256_b64": rk4[64],
        "roofline": {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": ach / HBM_PEAK_GBS, "traffic": TRAFFIC_ROW_T_N128_B1,
                     "traffic_source": "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs of bench.py --probe-roofline, "
                                       "profiles/pmc_traffic_tangent_row_kernel_n128_b1.md",
                     "kernel": f"row_kernel<128,TX={1 << (tiles['n128_b1']['replay_variants'][1] % 3)},tangent,replay> (batch 1)",
                     "avg_ms": row_t_ms,
                     "timing": "HIP events on the launching stream around 200 back-to-back launches of the "
                               "replay instance, median of three rounds",
                     "avg_ms_in_matvec_with_dispatch": probe_ms.value, "in_matvec_samples": probe_n.value,
                     "algorithmic_bytes": bytes_row_t,
                     "launch_floor_ms": floor_ms,
                     "launch_floor": "an empty kernel with the same grid and argument block, 200 back-to-back "
                                     "launches timed the same way: what a launch costs before any work"},
        "roofline_batched": {"bound": "hbm", "achieved": ach_b, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                             "frac": ach_b / HBM_PEAK_GBS, "traffic": TRAFFIC_ROW_N256_B64,
                             "traffic_source": "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs of bench.py "
                                               "--probe-roofline-batched, profiles/pmc_batched_row_kernel_n256_b64.md",
                             "kernel": f"row_kernel<256,TX={tiles['n256_b64']['row_tx']}> (batch 64)",
                             "avg_ms": row_b_ms, "algorithmic_bytes": 64 * 10 * zbytes(256)},
        "cpu_baseline": cpu,
        "autotuned_tiles": tiles,
    }
    line.update(extras)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
