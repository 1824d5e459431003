"""Floquet block power iteration (floquet.py:42-99) on the GPU against the
oracle's restatement: R and the orthonormal block after a few iterations."""
import numpy as np
import pytest
import torch

from ecs_mhd_2d_amd import floquet, mhd
from ecs_mhd_2d_amd.synthetic import fields as synth_fields
from oracle import mhd_oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode", ["TDRK4", "Lawson_RK4"])
def test_block_power_iteration_matches_oracle(mode):
    print(f"floquet {mode}", flush=True)
    n, B, maxit = 16, 4, 3
    extra = {"num_checkpoints": 2, "ministeps": 5, "shift_reflect_ny": 3, "rot": True}
    pd = mhd.make_param_dict(n, 1 / 40, 1 / 40)
    pd.update(extra)
    opd = O.make_params(n, 1 / 40, 1 / 40)
    opd.update(extra)
    f = synth_fields(n, 3)
    rng = np.random.default_rng(0)
    t0 = rng.standard_normal((B, 2, n, n))
    out = floquet.block_power_iteration({"fields": torch.as_tensor(f).cuda(), "T": 0.5, "sx": 0.2}, pd,
                                        block_size=B, maxit=maxit, mode=mode, tang=torch.as_tensor(t0).cuda())
    R, Q = O.floquet_block_power({"fields": f, "T": 0.5, "sx": 0.2}, opd, t0, maxit, mode)
    assert np.abs(out["R"].cpu().numpy() - R).max() < 1e-9 * np.abs(R).max()
    assert np.abs(out["Q"].cpu().numpy() - Q).max() < 1e-9
