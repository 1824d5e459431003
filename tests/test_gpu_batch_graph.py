"""Captured hipGraph chunks agree with eager launches; batched candidates of
uneven N agree with the one-at-a-time path (bitwise: every candidate's
arithmetic is independent of the batch it rides in)."""
import os

import numpy as np
import pytest
import torch

from ecs_mhd_2d_amd import _runtime, batch
from ecs_mhd_2d_amd.synthetic import spectrum
from oracle import mhd_oracle as O

pytestmark = pytest.mark.gpu


def ctx_for(n):
    pd = O.make_params(n, 1 / 40, 1 / 40)
    return _runtime.Context(n, 1 / 40, 1 / 40, (0.0, 0.1), pd["forcing"])


def graphs(fn):
    os.environ["ECSMHD_GRAPH"] = "1"
    try:
        return fn()
    finally:
        del os.environ["ECSMHD_GRAPH"]


def test_graph_replays_match_eager_rk4_and_rk43():
    n = 32
    ctx = ctx_for(n)
    f = torch.as_tensor(np.stack([spectrum(n, 1), spectrum(n, 2)])).cuda()
    ref = ctx.lawson_rk4(f, [0.004, 0.005], 40)
    for rep in range(3):   # several replays with eager work and syncs between them
        out = graphs(lambda: ctx.lawson_rk4(f, [0.004, 0.005], 40))
        torch.cuda.synchronize()
        assert torch.equal(out, ref), rep
        ctx.state_vel(f, True)
        torch.cuda.synchronize()
    x0 = torch.as_tensor(spectrum(n, 3)).cuda()
    xr, ir = ctx.lawson_rk43(x0, 0.7, 1e-2, 1e-6, 32, 8)
    for rep in range(3):
        xg, ig = graphs(lambda: ctx.lawson_rk43(x0, 0.7, 1e-2, 1e-6, 32, 8))
        torch.cuda.synchronize()
        assert torch.equal(xg, xr), rep
        assert ig[0]["accepted"] == ir[0]["accepted"] and ig[0]["rejected"] == ir[0]["rejected"]
        assert np.array_equal(ig[0]["hs"], ir[0]["hs"])
    dx0 = torch.as_tensor(spectrum(n, 4)).cuda()
    # eager launches fuse each step's last kernel with the next step's first
    # (C_S5A); the captured chunk keeps them apart, so the two agree to
    # rounding (FMA contraction differs between the kernels), not bitwise
    _, dxr = ctx.lawson_rk43_jvp(x0, dx0, 0.7, 0.2, ir)
    first = None
    for rep in range(3):
        _, dxg = graphs(lambda: ctx.lawson_rk43_jvp(x0, dx0, 0.7, 0.2, ir))
        torch.cuda.synchronize()
        assert float(torch.linalg.vector_norm(dxg - dxr) / torch.linalg.vector_norm(dxr)) < 1e-12, rep
        if first is None:
            first = dxg
        assert torch.equal(dxg, first), rep   # replays of one graph are bitwise identical


def test_budget_exhaustion_with_graph_chunks_matches_oracle_counts():
    """A budget that is not a multiple of the graph chunk: the attempts past
    it only apply the last decision, as the reference's fixed scan does."""
    n = 16
    ctx = ctx_for(n)
    pd = O.make_params(n, 1 / 40, 1 / 40)
    x0 = spectrum(n, 5) * 3
    x, infos = graphs(lambda: ctx.lawson_rk43(torch.as_tensor(x0).cuda(), 2.0, 1e-2, 1e-6, 7, 5))
    rx, rinfo = O.lawson_rk43(x0, pd, 2.0, 1e-2, 1e-6, 7, 5, record=True)
    gi = infos[0]
    assert not gi["completed"] and not rinfo["completed"]
    assert gi["accepted"] + gi["rejected"] == 35 == rinfo["accepted"] + rinfo["rejected"]
    assert gi["accepted"] == rinfo["accepted"]
    assert abs(gi["s"] - rinfo["s"]) < 1e-9
    torch.cuda.synchronize()
    assert np.linalg.norm(x.cpu().numpy() - rx) / np.linalg.norm(rx) < 1e-9


def test_batched_uneven_candidates_equal_unbatched():
    cands = batch.make_candidates(count=6, seed=3, sizes=((32, 2), (64, 2), (128, 2)), T_range=(0.3, 0.6))
    rb, _ = batch.run_batched(cands, atol=1e-7, max_steps=64, checkpoints=16)
    rg, _ = graphs(lambda: batch.run_batched(cands, atol=1e-7, max_steps=64, checkpoints=16))
    ru, _ = batch.run_unbatched(cands, atol=1e-7, max_steps=64, checkpoints=16)
    for (xb, ib), (xu, iu), (xg, _) in zip(rb, ru, rg):
        assert ib["completed"] and iu["completed"]
        assert ib["accepted"] == iu["accepted"] and ib["rejected"] == iu["rejected"]
        assert torch.equal(xb, xu) and torch.equal(xg, xu)
