"""Hookstep Newton (linalg.py:102-337, newton_hookstep_v2.py:95-121) on the
GPU against the oracle's restatement on the same RPO problem, and the
reorthogonalised Arnoldi sweep against numpy."""
import numpy as np
import pytest
import torch

from ecs_mhd_2d_amd import linalg, loss_functions as LF, solvers, utils
from oracle import mhd_oracle as O
from tests.test_gpu_api import host, load_solution, rel, small_problem

pytestmark = pytest.mark.gpu


def test_reorthogonalised_arnoldi_matches_numpy():
    rng = np.random.default_rng(2)
    n, m = 3000, 20
    M = rng.standard_normal((n, n)) / np.sqrt(n) + np.eye(n)
    Md = torch.as_tensor(M).cuda()
    b = rng.standard_normal(n)
    Q, H = linalg._arnoldi_reorth(lambda v: torch.mv(Md, v), torch.as_tensor(b).cuda(), m)
    oQ, oH = O._arnoldi_reorth(lambda v: M @ v, b, m)
    np.testing.assert_allclose(H, oH, rtol=1e-10, atol=1e-12)
    Qh = host(Q)
    assert np.abs(Qh @ Qh.T - np.eye(m + 1)).max() < 1e-13


def test_hookstep_v2_matches_oracle_on_rpo_problem():
    pd, opd, inp, oinp = small_problem()
    n, m = 16, 10
    f = LF.objective_RPO(inp, pd, "Lawson_RK4")
    b = utils.flatten(f)
    lin = LF.Linearization(inp, pd, "Lawson_RK4")
    A = lambda v: utils.flatten(lin.jvp(utils.unflatten(v, n)))
    obj = lambda d: LF.objective_RPO(d, pd, "Lawson_RK4")
    for s0 in (100.0, 0.05):      # unconstrained, and trust region active
        new, s, rr = linalg.newton_gmres_hookstep_v2(A, b, m, s0, obj, inp)
        OA = lambda v: O.flatten(O.objective_RPO_jvp(oinp, O.unflatten(v, n), opd, "Lawson_RK4"))
        of = lambda v: O.flatten(O.objective_RPO(O.unflatten(v, n), opd, "Lawson_RK4")[0])
        ox, os_, orr, odx = O.hookstep_v2(OA, host(b), m, s0, of, O.flatten(oinp))
        print(f"s0 {s0}: s {s} / {os_}, rel_res {rr} / {orr}", flush=True)
        assert s == os_
        assert rel(host(utils.flatten(new)), ox) < 1e-9
        assert abs(rr - orr) < 1e-7 * max(orr, 1e-6)


def test_hookstep_v1_matches_oracle_on_normal_equations():
    pd, opd, inp, oinp = small_problem()
    n, m = 16, 8
    f = LF.objective_RPO(inp, pd, "Lawson_RK4")
    b = utils.flatten(f)
    lin = LF.Linearization(inp, pd, "Lawson_RK4")
    J = lambda v: utils.flatten(lin.jvp(utils.unflatten(v, n)))
    JT = lambda v: utils.flatten(LF.objective_RPO_vjp(inp, utils.unflatten(v, n), pd, "Lawson_RK4"))
    AtA = lambda v: JT(J(v))
    Atb = JT(b)
    obj = lambda d: LF.objective_RPO(d, pd, "Lawson_RK4")
    new, s, rr1, rr2 = linalg.newton_gmres_hookstep(AtA, Atb, m, 100.0, obj, b, J, -b, inp)
    OJ = lambda v: O.flatten(O.objective_RPO_jvp(oinp, O.unflatten(v, n), opd, "Lawson_RK4"))
    OJT = lambda v: O.flatten(O.objective_RPO_vjp(oinp, O.unflatten(v, n), opd))
    of = lambda v: O.flatten(O.objective_RPO(O.unflatten(v, n), opd, "Lawson_RK4")[0])
    bh = host(b)
    ox, os_, orr1, orr2, _ = O.hookstep_v1(lambda v: OJT(OJ(v)), OJT(bh), m, 100.0, of, bh, OJ, -bh, O.flatten(oinp))
    print(f"v1: s {s} / {os_}, rr1 {rr1} / {orr1}, rr2 {rr2} / {orr2}", flush=True)
    assert s == os_
    assert rel(host(utils.flatten(new)), ox) < 1e-8
    assert abs(rr1 - orr1) < 1e-6 * max(orr1, 1e-6)


def test_newton_hookstep_iteration_reduces_residual():
    inp, pd = load_solution("Re40_RPO6")
    gen = torch.Generator(device="cuda").manual_seed(1)
    pert = 1e-4 * torch.randn(inp["fields"].shape, generator=gen, device="cuda", dtype=torch.float64)
    x = {"fields": solvers.dealias(inp["fields"] + pert, pd), "T": inp["T"], "sx": inp["sx"]}
    r0 = float(torch.linalg.vector_norm(LF.objective_RPO(x, pd, "Lawson_RK4")["fields"]))
    x1, s, log = solvers.newton_hookstep_update(x, pd, "Lawson_RK4", 100.0, inner=16)
    r1 = float(torch.linalg.vector_norm(LF.objective_RPO(x1, pd, "Lawson_RK4")["fields"]))
    print("hookstep", r0, r1, log["rel_res"], s, flush=True)
    assert r1 < 0.5 * r0 and s == 200.0


def test_hookstep_windows_give_the_one_by_one_result():
    """Both hookstep forms with their trust-region tries integrated 3, 4 or 8
    at a time (objective_RPO_batch) return the bits, s and residuals of the
    one-by-one loop; from s0 = 100 the v2 form fails four times (s = 100 ..
    12.5) and takes the fifth try (s = 6.25, returned doubled)."""
    pd, opd, inp, oinp = small_problem()
    n = 16
    f = LF.objective_RPO(inp, pd, "Lawson_RK4")
    b = utils.flatten(f)
    lin = LF.Linearization(inp, pd, "Lawson_RK4")
    J = lambda v: utils.flatten(lin.jvp(utils.unflatten(v, n)))
    obj = lambda d: LF.objective_RPO(d, pd, "Lawson_RK4")
    ob = lambda ds: LF.objective_RPO_batch(ds, pd, "Lawson_RK4")
    res = {w: linalg.newton_gmres_hookstep_v2(J, b, 10, 100.0, obj, inp, objective_batch=ob, window=w)
           for w in (1, 3, 4, 8)}
    assert res[1][1] == 12.5
    for w in (3, 4, 8):
        assert res[w][1] == res[1][1] and res[w][2] == res[1][2]
        assert torch.equal(utils.flatten(res[w][0]), utils.flatten(res[1][0]))
    JT = lambda v: utils.flatten(lin.vjp(utils.unflatten(v, n)))
    AtA = lambda v: JT(J(v))
    Atb = JT(b)
    r1 = linalg.newton_gmres_hookstep(AtA, Atb, 8, 100.0, obj, b, J, -b, inp, objective_batch=ob, window=1)
    for w in (3, 8):
        rw = linalg.newton_gmres_hookstep(AtA, Atb, 8, 100.0, obj, b, J, -b, inp, objective_batch=ob, window=w)
        assert rw[1] == r1[1] and rw[2] == r1[2] and rw[3] == r1[3]
        assert torch.equal(utils.flatten(rw[0]), utils.flatten(r1[0]))
    print("hookstep windows: v2 s", res[1][1], "v1 s", r1[1], flush=True)
