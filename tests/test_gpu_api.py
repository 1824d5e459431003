"""The reference-shaped API on the GPU vs the oracle and the reference's own
converged solutions."""
import json
import os

import numpy as np
import pytest
import torch

from ecs_mhd_2d_amd import linalg, loss_functions as LF, mhd, solvers, utils
from ecs_mhd_2d_amd.synthetic import fields as synth_fields
from oracle import mhd_oracle as O

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def host(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy() if isinstance(t, torch.Tensor) else np.asarray(t)


def load_solution(tag):
    d = np.load(os.path.join(GOLD, "solutions", tag + ".npz"), allow_pickle=False)
    n = d["fields"].shape[-1]
    pd = mhd.make_param_dict(n, float(d["nu"]), float(d["eta"]), tuple(d["b0"]))
    pd.update({"num_checkpoints": int(d["num_checkpoints"]), "ministeps": int(d["ministeps"]),
               "shift_reflect_ny": int(d["shift_reflect_ny"]), "rot": bool(d["rot"])})
    inp = {"fields": torch.as_tensor(d["fields"]).cuda(), "T": float(d["T"]), "sx": float(d["sx"])}
    return inp, pd


@pytest.mark.parametrize("tag", ["Re40_EQ1", "Re40_RPO4", "Re40_RPO5", "Re40_RPO6", "Re40_TW3", "Re50_RPO2"])
def test_reference_solutions_converge_on_gpu(tag):
    """The reference's converged orbits are periodic under the fused GPU path:
    residual equal to the oracle's (committed) value within 1e-10 relative."""
    print(f"solution {tag}", flush=True)
    inp, pd = load_solution(tag)
    out = LF.objective_RPO(inp, pd, "Lawson_RK4")
    r = float(torch.linalg.vector_norm(out["fields"]) / torch.linalg.vector_norm(inp["fields"]))
    with open(os.path.join(GOLD, "residuals.json")) as fh:
        expect = json.load(fh)[tag]
    print(tag, r, expect, flush=True)
    assert abs(r - expect) < 1e-10


def small_problem(n=16, mode="Lawson_RK4"):
    pd = mhd.make_param_dict(n, 1 / 40, 1 / 40)
    pd.update({"num_checkpoints": 2, "ministeps": 5, "shift_reflect_ny": 3, "rot": True,
               "adaptive_dict": {"atol": 1e-6, "checkpoints": 8, "max_steps_per_checkpoint": 32}})
    opd = O.make_params(n, 1 / 40, 1 / 40)
    opd.update({k: pd[k] for k in ("num_checkpoints", "ministeps", "shift_reflect_ny", "rot", "adaptive_dict")})
    f = synth_fields(n, 7)
    inp = {"fields": torch.as_tensor(f).cuda(), "T": 0.6, "sx": 0.25}
    oinp = {"fields": f, "T": 0.6, "sx": 0.25}
    return pd, opd, inp, oinp


@pytest.mark.parametrize("mode", ["Lawson_RK4", "Lawson_RK43"])
def test_objective_and_jvp(mode):
    pd, opd, inp, oinp = small_problem()
    out = LF.objective_RPO(inp, pd, mode)
    ref, _ = O.objective_RPO(oinp, opd, mode)
    assert rel(host(out["fields"]), ref["fields"]) < 1e-10
    rng = np.random.default_rng(3)
    d = {"fields": rng.standard_normal((2, 16, 16)), "T": 0.3, "sx": -0.4}
    td = {"fields": torch.as_tensor(d["fields"]).cuda(), "T": 0.3, "sx": -0.4}
    j = LF.objective_RPO_jvp(inp, td, pd, mode)
    rj = O.objective_RPO_jvp(oinp, d, opd, mode)
    assert rel(host(utils.flatten(j)), O.flatten(rj)) < 1e-10


def test_objective_vjp_and_loss_gradient():
    pd, opd, inp, oinp = small_problem()
    rng = np.random.default_rng(4)
    cot = {"fields": rng.standard_normal((2, 16, 16)), "T": 0.7, "sx": 0.2}
    tc = {"fields": torch.as_tensor(cot["fields"]).cuda(), "T": 0.7, "sx": 0.2}
    g = LF.objective_RPO_vjp(inp, tc, pd)
    rg = O.objective_RPO_vjp(oinp, cot, opd)
    assert rel(host(utils.flatten(g)), O.flatten(rg)) < 1e-10
    (loss, _), grad = LF.loss_RPO_value_and_grad(inp, pd)
    rl, rgrad = O.loss_RPO_value_and_grad(oinp, opd)
    assert abs(float(loss) - rl) < 1e-12 * rl
    assert rel(host(utils.flatten(grad)), O.flatten(rgrad)) < 1e-10


def test_gmres_matches_oracle_on_rpo_jacobian():
    """GMRES on the RPO Jacobian: H, the Krylov residual history and the final
    rel_res agree with the oracle's GMRES fed the oracle's matvec."""
    pd, opd, inp, oinp = small_problem()
    n, inner = 16, 12
    f = LF.objective_RPO(inp, pd, "Lawson_RK4")
    b = utils.flatten(f) + 1e-3 * torch.ones(2 * n * n + 2, dtype=torch.float64, device="cuda")
    lin = LF.Linearization(inp, pd, "Lawson_RK4")
    A = lambda v: utils.flatten(lin.jvp(utils.unflatten(v, n)))
    x, rr, hist = linalg.gmres(A, b, inner, return_history=True)
    bh = host(b)
    OA = lambda v: O.flatten(O.objective_RPO_jvp(oinp, O.unflatten(v, n), opd, "Lawson_RK4"))
    ox, orr, oH, _ = O.gmres(OA, bh, inner)
    np.testing.assert_allclose(hist["H"], oH, rtol=1e-8, atol=1e-10)
    assert abs(rr - orr) < 1e-8 * max(orr, 1e-3)
    assert rel(host(x), ox) < 1e-8
    h = hist["krylov_residuals"]
    assert all(h[i + 1] <= h[i] * (1 + 1e-12) for i in range(len(h) - 1))


def test_newton_iteration_reduces_residual_on_perturbed_solution():
    inp, pd = load_solution("Re40_RPO6")
    gen = torch.Generator(device="cuda").manual_seed(0)
    pert = 1e-4 * torch.randn(inp["fields"].shape, generator=gen, device="cuda", dtype=torch.float64)
    x = {"fields": solvers.dealias(inp["fields"] + pert, pd), "T": inp["T"], "sx": inp["sx"]}
    r0 = float(torch.linalg.vector_norm(LF.objective_RPO(x, pd, "Lawson_RK4")["fields"]))
    x1, log = solvers.newton_gmres_update(x, pd, "Lawson_RK4", inner=16)
    r1 = float(torch.linalg.vector_norm(LF.objective_RPO(x1, pd, "Lawson_RK4")["fields"]))
    print("newton", r0, r1, log["gmres_rel_res"], log["damp"], flush=True)
    assert r1 < 0.5 * r0
