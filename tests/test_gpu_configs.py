"""The benchmark workloads at their own sizes, on the GPU.

* Newton-GMRES configuration: N=128, RK4(3) atol 1e-8, a guess evolved 8192
  Lawson-RK4 steps of h=0.005 from G(128, 0).  The primal and the tangent
  replay are compared with the oracle on a short period (the oracle takes
  seconds there); the full iteration (T=4.3, sx=0.1, Krylov dimension 64)
  is checked through properties that do not need the oracle: the adaptive
  integration completes, the Arnoldi relation A q_k = sum_j H[j,k] q_j
  holds, the residual history does not increase and ends at the true
  residual, and the line search lowers |f|.
* Batched candidates: N in {64, 128, 256} with three at N=256, T ~ U[2, 8],
  atol 1e-8; batched results equal one-at-a-time results bit for bit, and
  RK4(3) primal, tangent and reverse pass at N=256 match the oracle.
* The reference's converged solutions: the final state of each trajectory
  equals the oracle's (committed in tests/golden/final_states.npz).
"""
import json
import os

import numpy as np
import pytest
import torch

from ecs_mhd_2d_amd import _runtime, batch, linalg, loss_functions as LF, mhd, solvers, utils
from ecs_mhd_2d_amd.synthetic import spectrum
from oracle import mhd_oracle as O

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300))


def host(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


_GUESS = {}


def newton_guess():
    """G(128, 0) evolved 8192 Lawson-RK4 steps of h = 0.005 (real fields)."""
    if "f" not in _GUESS:
        n = 128
        pd = mhd.make_param_dict(n, 1 / 40, 1 / 40)
        f = mhd.context(pd).lawson_rk4(torch.as_tensor(spectrum(n, 0)).cuda(), 0.005, 8192)
        _GUESS["f"] = torch.fft.irfft2(f, s=(n, n)).contiguous()
    return _GUESS["f"]


def newton_pd(n=128):
    pd = mhd.make_param_dict(n, 1 / 40, 1 / 40)
    pd.update({"shift_reflect_ny": 0, "rot": False,
               "adaptive_dict": {"atol": 1e-8, "checkpoints": 64, "max_steps_per_checkpoint": 64}})
    return pd


def test_newton_configuration_rk43_primal_and_tangent_match_oracle():
    """N=128, atol 1e-8 on the Newton guess, T = 0.25: decisions identical,
    step sizes to 1e-7, final state to 1e-10, tangent replay (dT = 0.3) to
    1e-10."""
    print("N=128 rk43 primal/jvp vs oracle, T=0.25", flush=True)
    n = 128
    pd = O.make_params(n, 1 / 40, 1 / 40)
    ctx = _runtime.Context(n, 1 / 40, 1 / 40, (0.0, 0.1), pd["forcing"])
    x0 = np.fft.rfft2(host(newton_guess()))
    dx0 = spectrum(n, 6)
    T, dT = 0.25, 0.3
    x, infos = ctx.lawson_rk43(torch.as_tensor(x0).cuda(), T, 1e-2, 1e-8, 64, 64)
    rx, rinfo = O.lawson_rk43(x0, pd, T, 1e-2, 1e-8, 64, 64, record=True)
    gi = infos[0]
    sched = [(gi["hs"][i], bool(gi["acc"][i])) for i in range(gi["attempts"])]
    print("attempts", gi["attempts"], "rejected", gi["rejected"], flush=True)
    assert gi["completed"] and [a for _, a in sched] == [a for _, a in rinfo["schedule"]]
    np.testing.assert_allclose([h for h, _ in sched], [h for h, _ in rinfo["schedule"]], rtol=1e-7)
    assert rel(host(x), rx) < 1e-10
    xj, dxj = ctx.lawson_rk43_jvp(torch.as_tensor(x0).cuda(), torch.as_tensor(dx0).cuda(), T, dT, infos)
    _, rdx = O.lawson_rk43_jvp(x0, dx0, pd, T, dT, sched)
    assert rel(host(dxj), rdx) < 1e-10


def test_newton_configuration_full_iteration_properties():
    """T = 4.3, sx = 0.1, Krylov 64 on the Newton guess."""
    print("N=128 full Newton iteration, Krylov 64", flush=True)
    n = 128
    pd = newton_pd(n)
    inp = {"fields": newton_guess(), "T": 4.3, "sx": 0.1}
    lin = LF.Linearization(inp, pd, "Lawson_RK43")
    assert lin.info["completed"] and lin.info["attempts"] > 1000
    b = utils.flatten(lin.objective())
    assert bool(torch.isfinite(b).all())
    A = lambda v: utils.flatten(lin.jvp(utils.unflatten(v, n)))
    x, rr, hist = linalg.gmres(A, b, 64, return_history=True)
    H, Q = hist["H"], hist["Q"]
    for k in (0, 31, 63):
        w = A(Q[k])
        hq = torch.mv(Q[: k + 2].T.contiguous(), torch.as_tensor(H[: k + 2, k], device="cuda"))
        assert float(torch.linalg.vector_norm(w - hq) / torch.linalg.vector_norm(w)) < 1e-10, k
    res = hist["krylov_residuals"]
    print("residual history", ["%.3e" % r for r in res], flush=True)
    # non-increasing in exact arithmetic; each value is a least-squares solve that
    # rounds at ~eps cond(H_k), far below the 1e-10 slack
    assert all(res[k + 1] <= res[k] + 1e-10 for k in range(len(res) - 1))
    print("rel_res", rr, "estimate", res[-1], "qtb", hist["krylov_residuals_qtb"][-1], "orth",
          hist["orthogonality_loss"], flush=True)
    # while it is above the floor that rounding sets (~1e-8 here), the Arnoldi
    # estimate is the true residual |A x_k - b| / |b| of the iterate it describes,
    # x_k = Q_k y_k with y_k = argmin |H_k y - |b| e_1|
    k = 20
    nb = float(torch.linalg.vector_norm(b))
    e1 = np.zeros(k + 1)
    e1[0] = nb
    yk = np.linalg.lstsq(H[: k + 1, :k], e1, rcond=None)[0]
    xk = torch.mv(Q[:k].T.contiguous(), torch.as_tensor(yk, device="cuda"))
    true_k = float(torch.linalg.vector_norm(A(xk) - b)) / nb
    assert abs(true_k - res[k - 1]) < 1e-6 * res[k - 1], (true_k, res[k - 1])
    # the 64-vector solve (the reference's Q^T b right-hand side, linalg.py:92-97) reaches that floor
    assert rr < 1e-6 and res[-1] <= rr
    new, log = solvers.newton_gmres_update(inp, pd, "Lawson_RK43", inner=64)
    assert log["completed"] and log["gmres_rel_res"] == rr      # same bits as the solve above
    f_new = utils.flatten(LF.objective_RPO(new, pd, "Lawson_RK43"))
    assert log["damp"] >= 2.0 ** -19
    # |f| at x - damp step (before dealiasing) is below |f(x)| by the line search; the dealiased
    # iterate is within the dealiasing's change of that
    assert float(torch.linalg.vector_norm(f_new)) < log["f_norm"] * 1.01


def test_newton_configuration_transpose_dot_test():
    """The transpose of the headline matvec at its full size (VJP through the
    whole RK4(3) trajectory of T = 4.3, about 1400 attempts, in checkpointed
    segments whose forward passes run on a second stream) against the matvec:
    <u, J v> = <J^T u, v>; and the one-stream reverse pass gives the same bits."""
    print("N=128 transpose dot test, T=4.3", flush=True)
    n = 128
    pd = newton_pd(n)
    inp = {"fields": newton_guess(), "T": 4.3, "sx": 0.1}
    lin = LF.Linearization(inp, pd, "Lawson_RK43")
    assert lin.info["completed"] and lin.info["attempts"] > 1000
    rng = np.random.default_rng(12)
    v = torch.as_tensor(rng.standard_normal(2 * n * n + 2)).cuda()
    u = torch.as_tensor(rng.standard_normal(2 * n * n + 2)).cuda()
    Jv = utils.flatten(lin.jvp(utils.unflatten(v, n)))
    JTu = utils.flatten(lin.vjp(utils.unflatten(u, n)))
    lhs, rhs = float(torch.dot(u, Jv)), float(torch.dot(JTu, v))
    print("dot test", lhs, rhs, flush=True)
    assert abs(lhs - rhs) < 1e-9 * max(abs(lhs), 1.0)
    ctx = mhd.context(pd)
    ctx.set_rev_overlap(False)
    try:
        one = utils.flatten(lin.vjp(utils.unflatten(u, n)))
    finally:
        ctx.set_rev_overlap(True)
    assert torch.equal(one, JTu)


def test_batched_candidates_equal_one_at_a_time():
    """Eight candidates (N = 64 x2, 128 x3, 256 x3), T ~ U[2, 8], atol 1e-8:
    the batched, grouped, concurrent run gives each candidate the bits and
    the schedule of its solo run."""
    print("batched candidates, 8 incl. 3 at N=256", flush=True)
    cands = batch.make_candidates(count=8, seed=3, sizes=((64, 2), (128, 3), (256, 3)))
    assert sum(c["n"] == 256 for c in cands) == 3
    rb, _ = batch.run_batched(cands)
    ru, _ = batch.run_unbatched(cands)
    for c, (xb, ib), (xu, iu) in zip(cands, rb, ru):
        assert ib["completed"] and iu["completed"]
        assert torch.equal(xb, xu), c
        assert ib["accepted"] == iu["accepted"] and ib["rejected"] == iu["rejected"]
        assert np.array_equal(ib["hs"], iu["hs"])


def test_rk43_n256_primal_tangent_and_reverse_match_oracle():
    """N=256, atol 1e-8, T = 0.05 from G(256, 5): schedule and state equal to
    the oracle's, the tangent replay to 1e-10, and the reverse pass through
    the dot test against the oracle's tangent:
    <c, J d> = <J^T c, d> + Tbar dT to 1e-10."""
    print("N=256 rk43 primal/jvp/vjp vs oracle", flush=True)
    n = 256
    pd = O.make_params(n, 1 / 40, 1 / 40)
    ctx = _runtime.Context(n, 1 / 40, 1 / 40, (0.0, 0.1), pd["forcing"])
    x0, d, c = spectrum(n, 5), spectrum(n, 6), spectrum(n, 7)
    T, dT = 0.05, 0.3
    x, infos = ctx.lawson_rk43(torch.as_tensor(x0).cuda(), T, 1e-2, 1e-8, 64, 8)
    rx, rinfo = O.lawson_rk43(x0, pd, T, 1e-2, 1e-8, 64, 8, record=True)
    gi = infos[0]
    sched = [(gi["hs"][i], bool(gi["acc"][i])) for i in range(gi["attempts"])]
    assert gi["completed"] and [a for _, a in sched] == [a for _, a in rinfo["schedule"]]
    assert rel(host(x), rx) < 1e-10
    _, dxj = ctx.lawson_rk43_jvp(torch.as_tensor(x0).cuda(), torch.as_tensor(d).cuda(), T, dT, infos)
    _, rdx = O.lawson_rk43_jvp(x0, d, pd, T, dT, sched)
    assert rel(host(dxj), rdx) < 1e-10
    gb, Tb = ctx.lawson_rk43_vjp(torch.as_tensor(x0).cuda(), torch.as_tensor(c).cuda(), T, infos, ckpt=16)
    lhs = O.rdot(c, rdx)
    rhs = O.rdot(host(gb), d) + Tb[0] * dT
    assert abs(lhs - rhs) < 1e-10 * abs(lhs), (lhs, rhs)


@pytest.mark.parametrize("tag", ["Re40_EQ1", "Re40_RPO4", "Re40_RPO5", "Re40_RPO6", "Re40_TW3", "Re50_RPO2"])
def test_reference_solutions_final_states_match_oracle(tag):
    """The final state of each converged solution's Lawson-RK4 trajectory
    (dt = T / steps, before the symmetry) equals the oracle's compact final
    spectrum.  The trajectories are 640-3264 steps long and these orbits are
    unstable: rounding differences of ~1e-15 per step grow by the orbit's
    Floquet multipliers, so the bound is 1e-10 (the states agree to ~1e-12)."""
    print(f"final state {tag}", flush=True)
    from tests.test_gpu_api import load_solution
    inp, pd = load_solution(tag)
    fT, _ = LF.integrate(torch.fft.rfft2(inp["fields"]).contiguous(), inp["T"], pd, "Lawson_RK4")
    n = fT.shape[-2]
    kmax = (n - 1) // 3
    got = host(fT)
    comp = np.concatenate([got[:, : kmax + 1, : kmax + 1], got[:, n - kmax:, : kmax + 1]], axis=1)
    ref = np.load(os.path.join(GOLD, "final_states.npz"), allow_pickle=False)[tag]
    r = rel(comp, ref)
    print(tag, r, flush=True)
    assert r < 1e-10
    outside = got.copy()
    outside[:, : kmax + 1, : kmax + 1] = 0
    outside[:, n - kmax:, : kmax + 1] = 0
    assert np.abs(outside).max() == 0.0
