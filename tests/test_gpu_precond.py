"""Diagonal estimators (preconditioners.py:225-359) on the GPU: exact on
operators diagonal in their basis, and the batched Jacobian probe equal to
one-at-a-time JVPs."""
import numpy as np
import pytest
import torch

from ecs_mhd_2d_amd import loss_functions as LF, mhd, preconditioners as P, utils
from ecs_mhd_2d_amd.synthetic import fields as synth_fields

pytestmark = pytest.mark.gpu


def test_spatial_estimator_is_exact_on_a_diagonal_operator():
    n = 16
    d = torch.as_tensor(np.random.default_rng(0).standard_normal((2, n, n))).cuda()
    M, diag = P.diagonal_preconditioner_spatial({"fields": d}, None, lambda z: d * z, k=2, batch=3)
    assert torch.allclose(diag, d, rtol=0, atol=1e-15)
    v = torch.arange(2 * n * n + 2, dtype=torch.float64, device="cuda") + 1.0
    Mv = M(v)
    assert Mv[0] == v[0] and Mv[-1] == v[-1] and torch.allclose(Mv[1:-1], v[1:-1] / d.reshape(-1))


def test_fourier_estimator_is_exact_on_the_linear_propagator():
    """Zero state, no forcing, b0 = 0: J z = z - exp(L T) mask z, diagonal in
    Fourier space with entries 1 - exp(-nu k^2 T) on the 2/3 modes (1 off them)."""
    n = 16
    pd = mhd.make_param_dict(n, 0.3, 0.3, b0=(0.0, 0.0), forcing_amp=0.0)
    pd.update({"num_checkpoints": 1, "ministeps": 4, "shift_reflect_ny": 0, "rot": False})
    inp = {"fields": torch.zeros(2, n, n, dtype=torch.float64, device="cuda"), "T": 0.7, "sx": 0.0}
    jb = P.batched_fields_jvp(inp, pd, "Lawson_RK4")
    _, _, dhat = P.diagonal_preconditioner_fourier(inp, jb, k=2, batch=3)
    expect = 1 - torch.exp(0.7 * mhd.dissipation(pd)) * pd["mask"]
    assert torch.allclose(dhat.real, expect.expand_as(dhat.real), atol=1e-12)
    assert dhat.imag.abs().max() < 1e-12


def test_batched_probe_equals_single_jvps():
    n = 16
    pd = mhd.make_param_dict(n, 1 / 40, 1 / 40)
    pd.update({"num_checkpoints": 2, "ministeps": 4, "shift_reflect_ny": 3, "rot": True})
    inp = {"fields": torch.as_tensor(synth_fields(n, 2)).cuda(), "T": 0.4, "sx": 0.1}
    z = torch.as_tensor(np.random.default_rng(1).standard_normal((3, 2, n, n))).cuda()
    jb = P.batched_fields_jvp(inp, pd, "TDRK4")(z)
    lin = LF.Linearization(inp, pd, "TDRK4")
    for b in range(3):
        one = lin.jvp({"fields": z[b], "T": 0.0, "sx": 0.0})["fields"]
        assert torch.equal(jb[b], one)
