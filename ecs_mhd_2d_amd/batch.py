"""Many RPO candidates of uneven grid size and period in batched launches.

Candidates are grouped by N; each group advances as one batch through the
fused kernels (per-candidate T and adaptive controller, finished candidates
turn into no-ops), and the groups run concurrently on their own streams.
The adaptive attempts are launched eagerly, 32 attempts per host poll; the
library replays them as captured hipGraph chunks only when ECSMHD_GRAPH=1 is
set.  Tile sizes come from the library's autotuner.  The unbatched path (one
candidate at a time) is kept for comparison.
"""
from __future__ import annotations

import threading
import time

import numpy as np
import torch

from . import mhd
from .synthetic import spectrum


def make_candidates(count=64, seed=0, sizes=((64, 16), (128, 32), (256, 16)), T_range=(2.0, 8.0),
                    sx_range=(-0.5, 0.5)):
    """The batched-candidates workload: N drawn with the given multiplicities,
    T ~ U[T_range], sx ~ U[sx_range], spectra G(N, seed_i) (synthetic.spectrum)."""
    rng = np.random.default_rng(seed)
    ns = np.concatenate([[n] * m for n, m in sizes])
    rng.shuffle(ns)
    ns = ns[:count]
    out = []
    for i, n in enumerate(ns):
        out.append({"n": int(n), "T": float(rng.uniform(*T_range)), "sx": float(rng.uniform(*sx_range)),
                    "seed": i})
    return out


_PD = {}


def _ctx(n, nu, eta, b0):
    key = (n, nu, eta, tuple(b0))
    if key not in _PD:
        _PD[key] = mhd.make_param_dict(n, nu, eta, b0)
    return mhd.context(_PD[key])


def run_batched(cands, nu=1 / 40, eta=1 / 40, b0=(0.0, 0.1), atol=1e-8, h0=1e-2, max_steps=64, checkpoints=128):
    """Integrate every candidate over its period T with adaptive Lawson RK4(3).

    Returns (results, wall_seconds) with results[i] = (x_T spectrum, info) in
    input order."""
    groups = {}
    for i, c in enumerate(cands):
        groups.setdefault(c["n"], []).append(i)
    inputs = {}
    for n, idx in groups.items():
        inputs[n] = torch.as_tensor(np.stack([spectrum(n, cands[i]["seed"]) for i in idx])).cuda()
        _ctx(n, nu, eta, b0)
    results = [None] * len(cands)
    errors = []
    torch.cuda.synchronize()

    def work(n, idx):
        try:
            s = torch.cuda.Stream()
            with torch.cuda.stream(s):
                ctx = _ctx(n, nu, eta, b0)
                x, infos = ctx.lawson_rk43(inputs[n], [cands[i]["T"] for i in idx], h0, atol, max_steps, checkpoints)
                s.synchronize()
            for j, i in enumerate(idx):
                results[i] = (x[j], infos[j])
        except Exception as e:  # surfaced after join
            errors.append(e)

    t0 = time.perf_counter()
    threads = [threading.Thread(target=work, args=(n, idx)) for n, idx in groups.items()]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    if errors:
        raise errors[0]
    return results, wall


def run_unbatched(cands, nu=1 / 40, eta=1 / 40, b0=(0.0, 0.1), atol=1e-8, h0=1e-2, max_steps=64, checkpoints=128):
    """The same integrations one candidate at a time on the current stream."""
    xs = [torch.as_tensor(spectrum(c["n"], c["seed"])).cuda() for c in cands]
    for c in cands:
        _ctx(c["n"], nu, eta, b0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    results = []
    for c, x in zip(cands, xs):
        xt, infos = _ctx(c["n"], nu, eta, b0).lawson_rk43(x, c["T"], h0, atol, max_steps, checkpoints)
        results.append((xt, infos[0]))
    torch.cuda.synchronize()
    return results, time.perf_counter() - t0


def accepted_steps(results):
    return int(sum(info["accepted"] for _, info in results))
