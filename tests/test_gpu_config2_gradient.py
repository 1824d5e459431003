"""Adjoint-descent gradient at full size: N=256, T=4.3, 1088 Lawson-RK4 steps
(17 segments of 64, checkpoint/recompute), checked against central finite
differences (central at eps = 1e-6, fourth-order at eps = 1e-4) on 16 seeded
unit directions over [fields, T, sx], and against the tangent by the adjoint
dot test."""
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
    """Error model of a finite difference through this chaotic trajectory:
    the rounding of two separately integrated trajectories differs by about
    1e-12 of the loss (FP64 rounding grown over T = 4.3), which a difference
    quotient divides by eps; truncation grows as eps^2 (central) or eps^4
    (fourth-order stencil (8 (f(x+e) - f(x-e)) - (f(x+2e) - f(x-2e))) / (12 e)).
    At the configured eps = 1e-6 the noise term dominates (about 1e-6
    absolute against |J.d| ~ 1e-2), so the central check is bounded at 1e-3;
    the fourth-order stencil at eps = 1e-4 divides the noise by 100 while its
    truncation stays negligible, and is held to 2e-5."""
    n, pd, inp = setup()
    (loss, _), grad = LF.loss_RPO_value_and_grad(inp, pd)
    g = utils.flatten(grad)
    x = utils.flatten(inp)

    def f(v):
        return float(LF.loss_RPO(utils.unflatten(v, n), pd, "Lawson_RK4")[0])

    worst_c = worst_4 = 0.0
    for i in range(16):
        rng = np.random.default_rng(1000 + i)
        d = torch.as_tensor(rng.standard_normal(x.numel())).cuda()
        d = d / torch.linalg.vector_norm(d)
        an = float(torch.dot(g, d))
        scale = max(abs(an), 1e-3 * float(torch.linalg.vector_norm(g)))
        e = 1e-6
        central = (f(x + e * d) - f(x - e * d)) / (2 * e)
        e = 1e-4
        p1, m1, p2, m2 = f(x + e * d), f(x - e * d), f(x + 2 * e * d), f(x - 2 * e * d)
        fourth = (8 * (p1 - m1) - (p2 - m2)) / (12 * e)
        ec, e4 = abs(central - an) / scale, abs(fourth - an) / scale
        print(f"direction {i}: J.d {an:.12e} central(1e-6) rel err {ec:.2e} 4th-order(1e-4) rel err {e4:.2e}",
              flush=True)
        worst_c, worst_4 = max(worst_c, ec), max(worst_4, e4)
    print("loss", float(loss), "|grad|", float(torch.linalg.vector_norm(g)), "worst rel FD err: central", worst_c,
          "4th order", worst_4, flush=True)
    assert worst_c < 1e-3
    assert worst_4 < 2e-5


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
