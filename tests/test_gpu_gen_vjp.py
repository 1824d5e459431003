"""Reverse pass of classical RK4, Lawson RK6 and two-derivative RK4 on the
GPU: against the brute-force transpose of the oracle's tangent (N=16), and
the dot test <c, J d> = <J^T c, d> + hbar dh against the GPU tangent (N=64,
several checkpoint segments, batch of two with their own steps)."""
import numpy as np
import pytest
import torch

from ecs_mhd_2d_amd import _runtime
from ecs_mhd_2d_amd.synthetic import spectrum
from oracle import mhd_oracle as O

pytestmark = pytest.mark.gpu


def setup(n):
    pd = O.make_params(n, 1 / 40, 1 / 40)
    return pd, _runtime.Context(n, 1 / 40, 1 / 40, (0.0, 0.1), pd["forcing"])


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


ORACLE_JVP = {
    "rk4": lambda f, d, T, dT, s, pd: O.rk4_jvp(f, d, T, dT, s, pd),
    "lawson_rk6": lambda f, d, T, dT, s, pd: O.lawson_rk6_jvp(f, d, T, dT, s, pd, True),
    "tdrk4": lambda f, d, T, dT, s, pd: O.tdrk4_jvp(f, d, T, dT, s, pd),
}


@pytest.mark.parametrize("scheme", ["rk4", "lawson_rk6", "tdrk4"])
def test_reverse_pass_is_the_exact_transpose(scheme):
    print(f"{scheme}_vjp dense n=16", flush=True)
    n, steps, h = 16, 3, 0.01
    pd, ctx = setup(n)
    f, c = spectrum(n, 11), spectrum(n, 12)
    gb, hb = getattr(ctx, scheme + "_vjp")(torch.as_tensor(f).cuda(), torch.as_tensor(c).cuda(), h, steps, ckpt=2)
    jv = lambda d, dt: ORACLE_JVP[scheme](f, d, h * steps, dt, steps, pd)[1]
    rf, rt = O.dense_vjp(jv, f.shape, c)
    assert rel(gb.cpu().numpy(), rf) < 1e-10
    # hbar = d/dh = steps * d/dT
    assert abs(hb[0] - rt * steps) < 1e-9 * max(1.0, abs(rt * steps))


@pytest.mark.parametrize("scheme", ["rk4", "lawson_rk6", "tdrk4"])
def test_dot_test_against_gpu_tangent(scheme):
    print(f"{scheme}_vjp dot n=64", flush=True)
    n, steps = 64, 5
    pd, ctx = setup(n)
    f = torch.as_tensor(np.stack([spectrum(n, 21), spectrum(n, 22)])).cuda()
    d = torch.as_tensor(np.stack([spectrum(n, 23), spectrum(n, 24)])).cuda()
    c = torch.as_tensor(np.stack([spectrum(n, 25), spectrum(n, 26)])).cuda()
    dts, ddts = [0.004, 0.005], [0.3, -0.2]
    _, jd = getattr(ctx, scheme + "_jvp")(f, d, dts, ddts, steps)
    gb, hb = getattr(ctx, scheme + "_vjp")(f, c, dts, steps, ckpt=2)
    for b in range(2):
        lhs = O.rdot(c[b].cpu().numpy(), jd[b].cpu().numpy())
        rhs = O.rdot(gb[b].cpu().numpy(), d[b].cpu().numpy()) + hb[b] * ddts[b]
        assert abs(lhs - rhs) < 1e-10 * abs(lhs), (b, lhs, rhs)


@pytest.mark.parametrize("mode", ["RK4", "TDRK4", "Lawson_RK6"])
def test_objective_vjp_and_gradient_in_other_modes(mode):
    """objective_RPO_vjp and loss_RPO_value_and_grad in the other modes: dot
    test against the objective JVP, and the gradient against the oracle's
    central differences of the loss."""
    from ecs_mhd_2d_amd import loss_functions as LF, mhd, utils
    print(f"objective vjp {mode}", flush=True)
    n = 16
    extra = {"num_checkpoints": 2, "ministeps": 3, "steps": 6, "shift_reflect_ny": 3, "rot": True}
    pd = mhd.make_param_dict(n, 1 / 40, 1 / 40)
    pd.update(extra)
    opd = O.make_params(n, 1 / 40, 1 / 40)
    opd.update(extra)
    rng = np.random.default_rng(3)
    f = np.fft.irfft2(spectrum(n, 5))
    inp = {"fields": torch.as_tensor(f).cuda(), "T": 0.4, "sx": 0.2}
    u = torch.as_tensor(rng.standard_normal(2 * n * n + 2)).cuda()
    v = torch.as_tensor(rng.standard_normal(2 * n * n + 2)).cuda()
    Jv = utils.flatten(LF.objective_RPO_jvp(inp, utils.unflatten(v, n), pd, mode))
    JTu = utils.flatten(LF.objective_RPO_vjp(inp, utils.unflatten(u, n), pd, mode))
    lhs, rhs = float(torch.dot(u, Jv)), float(torch.dot(JTu, v))
    assert abs(lhs - rhs) < 1e-10 * abs(lhs)
    (loss, _), grad = LF.loss_RPO_value_and_grad(inp, pd, mode)
    oinp = {"fields": f, "T": 0.4, "sx": 0.2}
    d = {"fields": rng.standard_normal((2, n, n)), "T": 0.3, "sx": -0.2}
    eps = 1e-6
    lp = O.loss_RPO({k: oinp[k] + eps * d[k] for k in oinp}, opd, mode)[0]
    lm = O.loss_RPO({k: oinp[k] - eps * d[k] for k in oinp}, opd, mode)[0]
    g = utils.flatten(grad).cpu().numpy()
    an = float(np.dot(g, O.flatten(d)))
    assert abs(float(loss) - O.loss_RPO(oinp, opd, mode)[0]) < 1e-12 * float(loss)
    assert abs((lp - lm) / (2 * eps) - an) < 1e-6 * max(1.0, abs(an))


def test_newton_with_transpose_in_tdrk4_mode():
    """The reference's default Newton setting (newton.py:70,84,87: mode TDRK4,
    use_transpose=True, inner 16) reduces the residual of a perturbed orbit.
    GMRES on the normal equations A^T A squares the condition number, so 16
    Krylov vectors buy less than plain GMRES does (r1 = 0.72 r0 here, against
    >2x for the plain form in test_gpu_api): this is a property of the
    setting, not a tolerance.  The iteration itself is pinned against the
    oracle with the exact transpose (test_gpu_iterations.py,
    test_newton_with_transpose_in_tdrk4_mode_matches_oracle), and the reverse
    pass by the dot tests above."""
    from ecs_mhd_2d_amd import loss_functions as LF, solvers
    from tests.test_gpu_api import load_solution
    inp, pd = load_solution("Re40_RPO6")
    gen = torch.Generator(device="cuda").manual_seed(2)
    pert = 1e-4 * torch.randn(inp["fields"].shape, generator=gen, device="cuda", dtype=torch.float64)
    x = {"fields": solvers.dealias(inp["fields"] + pert, pd), "T": inp["T"], "sx": inp["sx"]}
    r0 = float(torch.linalg.vector_norm(LF.objective_RPO(x, pd, "TDRK4")["fields"]))
    x1, log = solvers.newton_gmres_update(x, pd, "TDRK4", inner=16, use_transpose=True)
    r1 = float(torch.linalg.vector_norm(LF.objective_RPO(x1, pd, "TDRK4")["fields"]))
    print("newton TDRK4 transpose", r0, r1, log["gmres_rel_res"], log["damp"], flush=True)
    assert r1 < 0.8 * r0 and log["gmres_rel_res"] < 0.05


def test_rk43_reverse_pass_is_the_exact_transpose():
    """Reverse pass of the adaptive stepper along its recorded schedule
    (rejections first: the carried k1 quirk enters the first accepted step):
    the brute-force transpose of the oracle's tangent along the same schedule."""
    print("lawson_rk43_vjp dense n=16", flush=True)
    n = 16
    pd, ctx = setup(n)
    x0, c = spectrum(n, 41), spectrum(n, 13)
    T = 0.15
    _, infos = ctx.lawson_rk43(torch.as_tensor(x0).cuda(), T, 0.7, 1e-6, 32, 8)
    info = infos[0]
    sched = [(info["hs"][i], bool(info["acc"][i])) for i in range(info["attempts"])]
    assert info["completed"] and info["rejected"] >= 2
    gb, Tb = ctx.lawson_rk43_vjp(torch.as_tensor(x0).cuda(), torch.as_tensor(c).cuda(), T, infos, ckpt=8)
    rf, rt = O.dense_vjp(lambda d, dt: O.lawson_rk43_jvp(x0, d, pd, T, dt, sched)[1], x0.shape, c)
    assert rel(gb.cpu().numpy(), rf) < 1e-10
    assert abs(Tb[0] - rt) < 1e-9 * max(1.0, abs(rt))


def test_rk43_reverse_pass_dot_test_batched():
    """Batch of two candidates with different schedules (the shorter one takes
    zero-length steps at the end): dot test against the GPU tangent."""
    print("lawson_rk43_vjp dot n=32", flush=True)
    n = 32
    pd, ctx = setup(n)
    x0 = torch.as_tensor(np.stack([spectrum(n, 51), 1.5 * spectrum(n, 52)])).cuda()
    d = torch.as_tensor(np.stack([spectrum(n, 53), spectrum(n, 54)])).cuda()
    c = torch.as_tensor(np.stack([spectrum(n, 55), spectrum(n, 56)])).cuda()
    T, dT = [0.5, 0.7], [0.3, -0.4]
    _, infos = ctx.lawson_rk43(x0, T, 0.5, 1e-7, 64, 8)
    assert infos[0]["accepted"] != infos[1]["accepted"]
    _, jd = ctx.lawson_rk43_jvp(x0, d, T, dT, infos)
    gb, Tb = ctx.lawson_rk43_vjp(x0, c, T, infos, ckpt=16)
    for b in range(2):
        lhs = O.rdot(c[b].cpu().numpy(), jd[b].cpu().numpy())
        rhs = O.rdot(gb[b].cpu().numpy(), d[b].cpu().numpy()) + Tb[b] * dT[b]
        assert abs(lhs - rhs) < 1e-10 * abs(lhs), (b, lhs, rhs)


def test_objective_vjp_dot_test_in_rk43_mode():
    from ecs_mhd_2d_amd import loss_functions as LF, mhd, utils
    n = 16
    pd = mhd.make_param_dict(n, 1 / 40, 1 / 40)
    pd.update({"shift_reflect_ny": 2, "rot": True,
               "adaptive_dict": {"atol": 1e-7, "checkpoints": 8, "max_steps_per_checkpoint": 32}})
    rng = np.random.default_rng(4)
    inp = {"fields": torch.as_tensor(np.fft.irfft2(spectrum(n, 6))).cuda(), "T": 0.3, "sx": 0.1}
    u = torch.as_tensor(rng.standard_normal(2 * n * n + 2)).cuda()
    v = torch.as_tensor(rng.standard_normal(2 * n * n + 2)).cuda()
    Jv = utils.flatten(LF.objective_RPO_jvp(inp, utils.unflatten(v, n), pd, "Lawson_RK43"))
    JTu = utils.flatten(LF.objective_RPO_vjp(inp, utils.unflatten(u, n), pd, "Lawson_RK43"))
    lhs, rhs = float(torch.dot(u, Jv)), float(torch.dot(JTu, v))
    assert abs(lhs - rhs) < 1e-10 * abs(lhs)


@pytest.mark.parametrize("mode", ["Lawson_RK43", "TDRK4", "RK4", "Lawson_RK6", "Lawson_RK4"])
def test_linearization_vjp_reuses_the_primal(mode):
    """Linearization(x).vjp (the vjp function of one primal, applied to several
    cotangents as newton.py:168-169 does; it keeps the reverse pass's
    checkpoints after its first call) computes the same bits as separate
    objective_RPO_vjp calls, each of which solves the primal again."""
    print(f"mode {mode}", flush=True)
    from ecs_mhd_2d_amd import loss_functions as LF, mhd, utils
    n = 16
    pd = mhd.make_param_dict(n, 1 / 40, 1 / 40)
    pd.update({"shift_reflect_ny": 2, "rot": True, "num_checkpoints": 2, "ministeps": 5, "steps": 10,
               "adaptive_dict": {"atol": 1e-7, "checkpoints": 8, "max_steps_per_checkpoint": 32}})
    rng = np.random.default_rng(9)
    inp = {"fields": torch.as_tensor(np.fft.irfft2(spectrum(n, 7))).cuda(), "T": 0.3, "sx": 0.1}
    lin = LF.Linearization(inp, pd, mode)
    for _ in range(2):
        u = utils.unflatten(torch.as_tensor(rng.standard_normal(2 * n * n + 2)).cuda(), n)
        a = utils.flatten(lin.vjp(u))
        b = utils.flatten(LF.objective_RPO_vjp(inp, u, pd, mode))
        assert torch.equal(a, b)


@pytest.mark.parametrize("n,B,steps", [(128, 1, 256), (128, 3, 200), (256, 2, 130)])
def test_lawson_rk4_reverse_overlap_gives_the_same_bits(n, B, steps):
    """The Lawson RK4 reverse pass with the next segment recomputed on a second
    stream while the current one is swept backward (two step tapes) returns
    the bits of the one-stream pass: 4, 4 (ragged last segment) and 3
    segments, batches 1-3; and the dot test against the tangent, so the
    overlapped pass is checked on its own too."""
    print(f"lawson rk4 reverse overlap n={n} B={B} steps={steps}", flush=True)
    pd = O.make_params(n, 1 / 40, 1 / 40)
    ctx = _runtime.Context(n, 1 / 40, 1 / 40, (0.0, 0.1), pd["forcing"])
    f0 = torch.as_tensor(np.stack([spectrum(n, s) for s in range(B)])).cuda()
    g = torch.as_tensor(np.stack([spectrum(n, 30 + s) for s in range(B)])).cuda()
    df = torch.as_tensor(np.stack([spectrum(n, 60 + s) for s in range(B)])).cuda()
    h = 0.5 / steps
    res = {}
    for ov in (False, True):
        ctx.set_rev_overlap(ov)
        res[ov] = ctx.lawson_rk4_vjp(f0, g, h, steps)
    assert torch.equal(res[True][0], res[False][0])
    assert list(res[True][1]) == list(res[False][1])
    # dot test per candidate: <g, J df> = <J^T g, df> (real inner product on stored entries)
    _, jdf = ctx.lawson_rk4_jvp(f0, df, h, 0.0, steps)
    for b in range(B):
        lhs = O.rdot(g[b].cpu().numpy(), jdf[b].cpu().numpy())
        rhs = O.rdot(res[True][0][b].cpu().numpy(), df[b].cpu().numpy())
        assert abs(lhs - rhs) < 1e-10 * abs(lhs), (b, lhs, rhs)


@pytest.mark.parametrize("scheme,n,B,steps,ckpt", [("rk4", 128, 1, 70, 16), ("lawson_rk6", 128, 2, 40, 16),
                                                   ("tdrk4", 128, 2, 40, 16), ("rk4", 256, 1, 33, 8),
                                                   ("tdrk4", 256, 1, 21, 8)])
def test_generic_reverse_step_tape_and_overlap(scheme, n, B, steps, ckpt, monkeypatch):
    """The generic reverse pass with a step tape per segment (one forward pass
    fills it; no step is recomputed before its transpose) gives the same bits
    with the next segment's forward pass on a second stream as on one, and
    matches the pass that recomputes each step (ECSMHD_SEG_TAPE=0) to 1e-12:
    several segments with a ragged last one, batches 1-2."""
    print(f"{scheme} reverse step tape n={n} B={B} steps={steps} ckpt={ckpt}", flush=True)
    pd, ctx = setup(n)
    f = torch.as_tensor(np.stack([spectrum(n, 70 + s) for s in range(B)])).cuda()
    c = torch.as_tensor(np.stack([spectrum(n, 80 + s) for s in range(B)])).cuda()
    dts = [0.004 + 0.001 * s for s in range(B)]
    vjp = getattr(ctx, scheme + "_vjp")
    res = {}
    for ov in (False, True):
        ctx.set_rev_overlap(ov)
        res[ov] = vjp(f, c, dts, steps, ckpt=ckpt)
    assert torch.equal(res[True][0], res[False][0])
    assert list(res[True][1]) == list(res[False][1])
    monkeypatch.setenv("ECSMHD_SEG_TAPE", "0")
    old = vjp(f, c, dts, steps, ckpt=ckpt)
    monkeypatch.delenv("ECSMHD_SEG_TAPE")
    print("bitwise equal to the recomputing pass:", torch.equal(old[0], res[True][0]), flush=True)
    for b in range(B):
        assert rel(res[True][0][b].cpu().numpy(), old[0][b].cpu().numpy()) < 1e-12
        assert abs(res[True][1][b] - old[1][b]) < 1e-12 * max(1.0, abs(old[1][b]))


@pytest.mark.parametrize("n,B,T", [(64, 2, 0.6), (128, 1, 0.5), (256, 1, 0.15)])
def test_rk43_reverse_step_tape_and_overlap(n, B, T, monkeypatch):
    """The same for the RK4(3) reverse pass along a recorded schedule with
    rejections (the tape holds the raw first-stage value the carried-k1 factor
    scales): overlap on and off give the same bits, and the recomputing pass
    agrees to 1e-12; and the dot test against the RK4(3) tangent."""
    print(f"rk43 reverse step tape n={n} B={B} T={T}", flush=True)
    pd, ctx = setup(n)
    x0 = torch.as_tensor(np.stack([(1 + 0.3 * s) * spectrum(n, 90 + s) for s in range(B)])).cuda()
    c = torch.as_tensor(np.stack([spectrum(n, 100 + s) for s in range(B)])).cuda()
    d = torch.as_tensor(np.stack([spectrum(n, 110 + s) for s in range(B)])).cuda()
    Ts, dTs = [T * (1 + 0.2 * s) for s in range(B)], [0.3 - 0.5 * s for s in range(B)]
    _, infos = ctx.lawson_rk43(x0, Ts, 0.5, 1e-8, 64, 64)
    assert all(i["completed"] for i in infos)
    assert sum(i["rejected"] for i in infos) >= 1
    # three or more segments
    ckpt = max(2, (max(i["accepted"] for i in infos) + 2) // 3)
    print("accepted", [i["accepted"] for i in infos], "ckpt", ckpt, flush=True)
    res = {}
    for ov in (False, True):
        ctx.set_rev_overlap(ov)
        res[ov] = ctx.lawson_rk43_vjp(x0, c, Ts, infos, ckpt=ckpt)
    assert torch.equal(res[True][0], res[False][0])
    assert list(res[True][1]) == list(res[False][1])
    monkeypatch.setenv("ECSMHD_SEG_TAPE", "0")
    old = ctx.lawson_rk43_vjp(x0, c, Ts, infos, ckpt=ckpt)
    monkeypatch.delenv("ECSMHD_SEG_TAPE")
    print("bitwise equal to the recomputing pass:", torch.equal(old[0], res[True][0]), flush=True)
    for b in range(B):
        assert rel(res[True][0][b].cpu().numpy(), old[0][b].cpu().numpy()) < 1e-12
        assert abs(res[True][1][b] - old[1][b]) < 1e-12 * max(1.0, abs(old[1][b]))
    _, jd = ctx.lawson_rk43_jvp(x0, d, Ts, dTs, infos)
    for b in range(B):
        lhs = O.rdot(c[b].cpu().numpy(), jd[b].cpu().numpy())
        rhs = O.rdot(res[True][0][b].cpu().numpy(), d[b].cpu().numpy()) + res[True][1][b] * dTs[b]
        assert abs(lhs - rhs) < 1e-10 * abs(lhs), (b, lhs, rhs)


@pytest.mark.parametrize("scheme,n,B,steps", [("rk4", 64, 2, 70), ("lawson_rk6", 32, 1, 40), ("tdrk4", 128, 1, 70)])
def test_two_phase_reverse_pass_gives_the_one_call_bits(scheme, n, B, steps):
    """ecs_gen_fwd_ckpt leaves the forward integrator's final state (bitwise)
    and checkpoints from which ecs_gen_vjp_ckpt, applied to two cotangents,
    returns the bits of the one-call reverse pass for each."""
    print(f"two-phase {scheme} n={n} B={B} steps={steps}", flush=True)
    pd, ctx = setup(n)
    f = torch.as_tensor(np.stack([spectrum(n, 120 + s) for s in range(B)])).cuda()
    dts = [0.004 + 0.001 * s for s in range(B)]
    fT, ck = ctx.gen_fwd_ckpt(scheme, f, dts, steps, 16)
    fwd = getattr(ctx, scheme)(f, dts, steps)
    assert torch.equal(fT, fwd)
    vjp = getattr(ctx, scheme + "_vjp")
    for seed in (130, 140):
        c = torch.as_tensor(np.stack([spectrum(n, seed + s) for s in range(B)])).cuda()
        a = ctx.gen_vjp_ckpt(scheme, ck, c, dts, steps, 16)
        b = vjp(f, c, dts, steps, ckpt=16)
        assert torch.equal(a[0], b[0]) and list(a[1]) == list(b[1])


def test_rk43_two_phase_reverse_pass_gives_the_one_call_bits():
    print("two-phase rk43 n=64 B=2", flush=True)
    n = 64
    pd, ctx = setup(n)
    x0 = torch.as_tensor(np.stack([spectrum(n, 150), 1.4 * spectrum(n, 151)])).cuda()
    Ts = [0.6, 0.8]
    _, infos = ctx.lawson_rk43(x0, Ts, 0.5, 1e-8, 64, 64)
    assert infos[0]["accepted"] != infos[1]["accepted"]
    ck = ctx.lawson_rk43_fwd_ckpt(x0, Ts, infos, 16)
    for seed in (160, 170):
        c = torch.as_tensor(np.stack([spectrum(n, seed), spectrum(n, seed + 1)])).cuda()
        a = ctx.lawson_rk43_vjp_ckpt(ck, c, Ts, infos, 16)
        b = ctx.lawson_rk43_vjp(x0, c, Ts, infos, ckpt=16)
        assert torch.equal(a[0], b[0]) and list(a[1]) == list(b[1])


def test_rk43_loss_gradient_against_the_frozen_schedule_jvp():
    """loss_RPO_value_and_grad in mode Lawson_RK43 (its backward reuses the
    forward's schedule and final state): <grad, d> equals the derivative of
    the loss along d with the step sizes frozen (h is stop-gradient,
    timestepping.py:308), 2 mean(o * J d) for the objective o and its JVP."""
    from ecs_mhd_2d_amd import loss_functions as LF, mhd, utils
    print("rk43 loss gradient", flush=True)
    n = 32
    pd = mhd.make_param_dict(n, 1 / 40, 1 / 40)
    pd.update({"shift_reflect_ny": 2, "rot": True,
               "adaptive_dict": {"atol": 1e-8, "checkpoints": 16, "max_steps_per_checkpoint": 64}})
    rng = np.random.default_rng(21)
    inp = {"fields": torch.as_tensor(np.fft.irfft2(spectrum(n, 8))).cuda(), "T": 0.7, "sx": 0.15}
    (loss, _), grad = LF.loss_RPO_value_and_grad(inp, pd, "Lawson_RK43")
    lin = LF.Linearization(inp, pd, "Lawson_RK43")
    o = lin.objective()["fields"]
    assert abs(float(loss) - float(torch.mean(o * o))) < 1e-13 * float(loss)
    for _ in range(2):
        d = utils.unflatten(torch.as_tensor(rng.standard_normal(2 * n * n + 2)).cuda(), n)
        lhs = float(torch.dot(utils.flatten(grad), utils.flatten(d)))
        rhs = float(2.0 * torch.mean(o * lin.jvp(d)["fields"]))
        assert abs(lhs - rhs) < 1e-10 * max(abs(rhs), 1e-300), (lhs, rhs)


def test_value_and_grad_leaves_the_callers_fields_alone():
    from ecs_mhd_2d_amd import loss_functions as LF, mhd
    n = 16
    pd = mhd.make_param_dict(n, 1 / 40, 1 / 40)
    pd.update({"num_checkpoints": 2, "ministeps": 4, "shift_reflect_ny": 0, "rot": False})
    f = torch.as_tensor(np.fft.irfft2(spectrum(n, 9))).cuda()
    LF.loss_RPO_value_and_grad({"fields": f, "T": 0.2, "sx": 0.0}, pd)
    assert not f.requires_grad and f.grad is None


@pytest.mark.parametrize("n", [32, 64])
def test_reverse_passes_at_batch_16_give_the_solo_bits(n):
    """From batch 16 the adjoint column kernels (the fused combination + B_G,
    B_OUT + combinations + next B_G) take multi-column tiles; each candidate
    of a batch of 16 gets the bits of its solo reverse pass (RK4, TDRK4 and
    RK4(3) along per-candidate schedules)."""
    print(f"reverse passes batch 16 vs solo n={n}", flush=True)
    pd, ctx = setup(n)
    B = 16
    f = torch.as_tensor(np.stack([(1 + 0.05 * s) * spectrum(n, 300 + s) for s in range(B)])).cuda()
    c = torch.as_tensor(np.stack([spectrum(n, 400 + s) for s in range(B)])).cuda()
    dts = [0.003 + 0.0002 * s for s in range(B)]
    for scheme in ("rk4", "tdrk4"):
        vjp = getattr(ctx, scheme + "_vjp")
        gb, hb = vjp(f, c, dts, 20, ckpt=8)
        for b in (0, 7, 15):
            g1, h1 = vjp(f[b:b + 1], c[b:b + 1], [dts[b]], 20, ckpt=8)
            assert torch.equal(gb[b:b + 1], g1) and hb[b] == h1[0], (scheme, b)
    Ts = [0.2 + 0.02 * s for s in range(B)]
    _, infos = ctx.lawson_rk43(f, Ts, 0.5, 1e-7, 64, 16)
    gb, Tb = ctx.lawson_rk43_vjp(f, c, Ts, infos, ckpt=16)
    for b in (0, 9, 15):
        _, inf1 = ctx.lawson_rk43(f[b:b + 1], [Ts[b]], 0.5, 1e-7, 64, 16)
        g1, T1 = ctx.lawson_rk43_vjp(f[b:b + 1], c[b:b + 1], [Ts[b]], inf1, ckpt=16)
        assert torch.equal(gb[b:b + 1], g1) and Tb[b] == T1[0], ("rk43", b)
