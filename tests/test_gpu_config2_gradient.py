"""Adjoint-descent gradient at full size: N=256, T=4.3, 1088 Lawson-RK4 steps
(17 segments of 64, checkpoint/recompute), checked against central finite
differences (eps = 1e-6) on 16 seeded unit directions over [fields, T, sx],
and against the tangent by the adjoint dot test."""
import numpy as np
import pytest
import torch

from ecs_mhd_2d_amd import loss_functions as LF, mhd, utils
from ecs_mhd_2d_amd.synthetic import fields as synth_fields

pytestmark = pytest.mark.gpu


def setup():
    n = 256
    pd = mhd.make_param_dict(n, 1 / 40, 1 / 40)
    pd.update({"num_checkpoints": 17, "ministeps": 64, "shift_reflect_ny": 0, "rot": False})
    inp = {"fields": torch.as_tensor(synth_fields(n, 2)).cuda(), "T": 4.3, "sx": 0.1}
    return n, pd, inp


def test_gradient_matches_central_differences_on_16_directions():
    n, pd, inp = setup()
    (loss, _), grad = LF.loss_RPO_value_and_grad(inp, pd)
    g = utils.flatten(grad)
    x = utils.flatten(inp)
    eps = 1e-6
    worst = 0.0
    for i in range(16):
        rng = np.random.default_rng(1000 + i)
        d = torch.as_tensor(rng.standard_normal(x.numel())).cuda()
        d = d / torch.linalg.vector_norm(d)
        lp, _ = LF.loss_RPO(utils.unflatten(x + eps * d, n), pd, "Lawson_RK4")
        lm, _ = LF.loss_RPO(utils.unflatten(x - eps * d, n), pd, "Lawson_RK4")
        fd = float(lp - lm) / (2 * eps)
        an = float(torch.dot(g, d))
        err = abs(fd - an) / max(abs(an), 1e-3 * float(torch.linalg.vector_norm(g)))
        worst = max(worst, err)
    print("loss", float(loss), "|grad|", float(torch.linalg.vector_norm(g)), "worst rel FD err", worst, flush=True)
    assert worst < 1e-5


def test_objective_vjp_dot_test_at_full_size():
    n, pd, inp = setup()
    lin = LF.Linearization(inp, pd, "Lawson_RK4")
    rng = np.random.default_rng(7)
    v = torch.as_tensor(rng.standard_normal(2 * n * n + 2)).cuda()
    u = torch.as_tensor(rng.standard_normal(2 * n * n + 2)).cuda()
    Jv = utils.flatten(lin.jvp(utils.unflatten(v, n)))
    JTu = utils.flatten(LF.objective_RPO_vjp(inp, utils.unflatten(u, n), pd))
    lhs, rhs = float(torch.dot(u, Jv)), float(torch.dot(JTu, v))
    print("dot test", lhs, rhs, flush=True)
    assert abs(lhs - rhs) < 1e-9 * max(abs(lhs), 1.0)
