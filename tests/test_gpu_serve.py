"""Serving loop: every job's result is the same alone and in the mix (the
kernels treat candidates independently, so bitwise here), and the loop
reports latency percentiles."""
import numpy as np
import pytest
import torch

from ecs_mhd_2d_amd import serve
from ecs_mhd_2d_amd.synthetic import fields as synth_fields

pytestmark = pytest.mark.gpu


def jobs(gap=0.0):
    """Five small jobs of two kinds on two grid sizes; with gap = 0 they all
    arrive together, so same-kind jobs share their first launches."""
    adaptive = {"atol": 1e-6, "max_steps_per_checkpoint": 32, "checkpoints": 8}
    specs = []
    for i, (kind, n, T) in enumerate([("adjoint", 32, 0.4), ("newton", 32, 0.3), ("adjoint", 32, 0.5),
                                      ("newton", 16, 0.35), ("adjoint", 16, 0.4)]):
        f = torch.as_tensor(synth_fields(n, 50 + i)).cuda()
        if kind == "adjoint":
            steps = 64 if T < 0.45 else 128
            gen = (lambda f=f, T=T, n=n, steps=steps: serve.adjoint_descent_job(f, T, 0.1, n, steps, 3))
            mem = serve.job_memory("adjoint", n, steps=steps)
        else:
            gen = (lambda f=f, T=T, n=n: serve.newton_job(f, T, 0.1, n, adaptive, 6, 2))
            mem = serve.job_memory("newton", n, inner=6)
        specs.append((gap * i, kind, gen, mem, n))
    return specs


def key(j):
    r = j.result
    if j.kind == "adjoint":
        return r["params"]["fields"], r["losses"]
    return r["x"]["fields"], [l["gmres_rel_res"] for l in r["logs"]]


def test_each_job_same_alone_and_in_the_mix():
    specs = jobs()
    done, stats = serve.Server().run(specs)
    mixed = {j.id: key(j) for j in done}
    assert stats["jobs"] == len(specs)
    assert stats["mean_batch"] > 1.0            # work of several jobs shared launches
    assert 0 < stats["p50_latency_s"] <= stats["p99_latency_s"]
    for i, spec in enumerate(sorted(specs, key=lambda a: a[0])):
        solo, _ = serve.Server().run([(0.0,) + spec[1:]])
        a, b = key(solo[0]), mixed[i]
        assert torch.equal(a[0], b[0]), i
        assert np.array_equal(np.asarray(a[1]), np.asarray(b[1])), i


def test_admission_respects_the_memory_budget():
    specs = jobs()
    budget = max(s[3] for s in specs) * 1.2       # the largest job fits, two large ones do not
    done, stats = serve.Server(hbm_budget=budget).run(specs)
    assert stats["jobs"] == len(specs)
    assert stats["peak_resident_bytes"] <= budget
    free, _ = serve.Server().run(specs)
    assert stats["wall_s"] >= 0.0 and len(free) == len(specs)


def config4_jobs():
    """Three jobs at the serving workload's per-job sizes of grid and period
    (N in {128, 256}, T in [2, 8], atol 1e-8, gradient steps = round(T/0.004)
    rounded up to whole 64-step segments), with fewer iterations each."""
    adaptive = {"atol": 1e-8, "max_steps_per_checkpoint": 64, "checkpoints": 64}
    specs = []
    for i, (kind, n, T) in enumerate([("adjoint", 256, 2.3), ("newton", 128, 2.1), ("adjoint", 128, 4.8)]):
        f = torch.as_tensor(synth_fields(n, 70 + i)).cuda()
        if kind == "adjoint":
            steps = int(np.ceil(round(T / 0.004) / 64) * 64)
            gen = (lambda f=f, T=T, n=n, steps=steps: serve.adjoint_descent_job(f, T, 0.0, n, steps, 2))
            mem = serve.job_memory("adjoint", n, steps=steps)
        else:
            gen = (lambda f=f, T=T, n=n: serve.newton_job(f, T, 0.0, n, adaptive, 4, 1))
            mem = serve.job_memory("newton", n, inner=4)
        specs.append((0.1 * i, kind, gen, mem, n))
    return specs


def test_serving_sizes_each_job_same_alone_and_in_the_mix():
    print("serving, 3 jobs at N in {128, 256}, T in [2, 8]", flush=True)
    specs = config4_jobs()
    done, stats = serve.Server().run(specs)
    print(stats, flush=True)
    assert stats["jobs"] == 3 and stats["failed_jobs"] == 0
    # later arrivals were admitted while earlier jobs were on the GPU
    assert any(a.start < b.finish and b.start < a.finish for a in done for b in done if a.id < b.id)
    mixed = {j.id: key(j) for j in done}
    for i, spec in enumerate(sorted(specs, key=lambda a: a[0])):
        solo, _ = serve.Server().run([(0.0,) + spec[1:]])
        a, b = key(solo[0]), mixed[i]
        assert torch.equal(a[0], b[0]), i
        assert np.array_equal(np.asarray(a[1]), np.asarray(b[1])), i
