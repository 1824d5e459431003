"""Parity of the HIP kernels (through the C-ABI) against the CPU oracle.

Tolerances: relative L2 in FP64.  The RHS is one pass of transforms, so it
agrees to ~1e-14; trajectories are held to 1e-11 after 64 steps (the bar
for states), tangents to 1e-10.
"""
import numpy as np
import pytest
import torch

from ecs_mhd_2d_amd import _runtime
from ecs_mhd_2d_amd.synthetic import spectrum
from oracle import mhd_oracle as O

pytestmark = pytest.mark.gpu

NU = ETA = 1 / 40
B0 = (0.0, 0.1)


def setup(n):
    pd = O.make_params(n, NU, ETA, B0)
    ctx = _runtime.Context(n, NU, ETA, B0, pd["forcing"])
    return pd, ctx


def dev(a):
    return torch.as_tensor(np.ascontiguousarray(a)).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("n", [16, 32, 64, 128, 256])
@pytest.mark.parametrize("diss", [False, True])
def test_state_vel(n, diss):
    print(f"state_vel n={n} diss={diss}", flush=True)
    pd, ctx = setup(n)
    f = np.stack([spectrum(n, 1), spectrum(n, 2) * 0.5])
    # include a non-dealiased component: state_vel must ignore it except via dissipation
    rng = np.random.default_rng(3)
    f = f + 1e-3 * np.fft.rfft2(rng.standard_normal((2, 2, n, n)))
    out = host(ctx.state_vel(dev(f), diss))
    ref = np.stack([O.state_vel(f[b], pd, diss) for b in range(2)])
    # FP64 rounding of two 2D transforms grows with log N (pocketfft and ours)
    assert rel(out, ref) < (1e-12 if n >= 256 else 1e-13)


@pytest.mark.parametrize("n", [32, 128])
def test_state_vel_jvp(n):
    print(f"state_vel_jvp n={n}", flush=True)
    pd, ctx = setup(n)
    f = spectrum(n, 4)
    df = spectrum(n, 5)
    out, dout = ctx.state_vel_jvp(dev(f), dev(df), True)
    assert rel(host(out), O.state_vel(f, pd, True)) < 1e-13
    assert rel(host(dout), O.state_vel_jvp(f, df, pd, True)) < 1e-13


@pytest.mark.parametrize("n,steps", [(16, 64), (32, 64), (128, 8)])
def test_lawson_rk4(n, steps):
    print(f"lawson_rk4 n={n} steps={steps}", flush=True)
    pd, ctx = setup(n)
    f = np.stack([spectrum(n, 11), spectrum(n, 12)])
    dts = [0.005, 0.004]
    out = host(ctx.lawson_rk4(dev(f), dts, steps))
    for b in range(2):
        ref = O.lawson_rk4(f[b], dts[b] * steps, steps, pd, True)
        assert rel(out[b], ref) < 1e-11


def test_lawson_rk4_unmasked_keeps_outside_modes():
    n = 32
    pd, ctx = setup(n)
    rng = np.random.default_rng(9)
    f = np.fft.rfft2(rng.standard_normal((2, n, n)))
    out = host(ctx.lawson_rk4(dev(f), 0.003, 5, masked=False))
    ref = O.lawson_rk4(f, 0.015, 5, pd, False)
    assert rel(out, ref) < 1e-11


@pytest.mark.parametrize("n", [32, 64])
def test_lawson_rk4_jvp(n):
    print(f"lawson_rk4_jvp n={n}", flush=True)
    pd, ctx = setup(n)
    f = spectrum(n, 21)
    df = spectrum(n, 22)
    steps, dt, ddt = 16, 0.005, 0.3 / 16
    x, dx = ctx.lawson_rk4_jvp(dev(f), dev(df), dt, ddt, steps)
    rx, rdx = O.lawson_rk4_jvp(f, df, dt * steps, ddt * steps, steps, pd, True)
    assert rel(host(x), rx) < 1e-11
    assert rel(host(dx), rdx) < 1e-10


@pytest.mark.parametrize("n", [32, 64])
def test_lawson_rk43(n):
    print(f"lawson_rk43 n={n}", flush=True)
    pd, ctx = setup(n)
    x0 = spectrum(n, 31)
    rng = np.random.default_rng(5)
    x0 = x0 + 1e-3 * np.fft.rfft2(rng.standard_normal((2, n, n)))  # outside-mask modes are carried
    T, atol = 1.0, 1e-6
    x, infos = ctx.lawson_rk43(dev(x0), T, 1e-2, atol, 64, 8)
    rx, rinfo = O.lawson_rk43(x0, pd, T, 1e-2, atol, 64, 8, record=True)
    gi = infos[0]
    sched = [(gi["hs"][i], bool(gi["acc"][i])) for i in range(gi["attempts"])]
    print("gpu", gi["accepted"], gi["rejected"], "oracle", rinfo["accepted"], rinfo["rejected"], flush=True)
    assert gi["completed"] and rinfo["completed"]
    assert gi["fevals"] == 1 + 4 * (gi["accepted"] + gi["rejected"])
    assert [a for _, a in sched] == [a for _, a in rinfo["schedule"]]
    # err = max|irfft2(h (k5 - k4)/6)| is a difference of nearly equal stages, so
    # its relative rounding is ~eps |k| / |k5 - k4|; h follows it as err^(-1/4)
    np.testing.assert_allclose([h for h, _ in sched], [h for h, _ in rinfo["schedule"]], rtol=1e-7)
    assert rel(host(x), rx) < 1e-10


def test_lawson_rk43_jvp():
    n = 32
    pd, ctx = setup(n)
    x0 = spectrum(n, 41)
    dx0 = spectrum(n, 42)
    T, dT, atol = 0.9, 0.4, 1e-6
    x, infos = ctx.lawson_rk43(dev(x0), T, 1e-2, atol, 32, 8)
    sched = [(infos[0]["hs"][i], bool(infos[0]["acc"][i])) for i in range(infos[0]["attempts"])]
    xj, dxj = ctx.lawson_rk43_jvp(dev(x0), dev(dx0), T, dT, infos)
    rx, rdx = O.lawson_rk43_jvp(x0, dx0, pd, T, dT, sched)
    assert rel(host(xj), host(x)) < 1e-12
    assert rel(host(xj), rx) < 1e-10
    assert rel(host(dxj), rdx) < 1e-10


def test_mgs_and_dot():
    from ecs_mhd_2d_amd import _lib
    lib = _lib.load()
    rng = np.random.default_rng(0)
    n, inner = 5000, 6
    A = rng.standard_normal((n, n // 50))
    Q = torch.zeros((inner + 1, n), dtype=torch.float64, device="cuda")
    b = rng.standard_normal(n)
    Q[0] = torch.as_tensor(b / np.linalg.norm(b))
    H = torch.zeros((inner, inner + 2), dtype=torch.float64, device="cuda")
    Qh = np.zeros((n, inner + 1))
    Qh[:, 0] = b / np.linalg.norm(b)
    Hh = np.zeros((inner + 1, inner))
    for k in range(inner):
        wv = A @ (A.T @ Qh[:, k])
        w = torch.as_tensor(wv).cuda()
        _lib.check(lib.ecs_mgs(Q.data_ptr(), n, n, k, w.data_ptr(), H[k].data_ptr(),
                               _runtime._stream()), "ecs_mgs")
        for j in range(k + 1):
            Hh[j, k] = wv @ Qh[:, j]
            wv = wv - Hh[j, k] * Qh[:, j]
        Hh[k + 1, k] = np.linalg.norm(wv)
        Qh[:, k + 1] = wv / Hh[k + 1, k]
    Hd = host(H)
    for k in range(inner):
        np.testing.assert_allclose(Hd[k, :k + 2], Hh[:k + 2, k], rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(host(Q).T, Qh, atol=1e-12)
    out = torch.zeros(1, dtype=torch.float64, device="cuda")
    a, c = torch.as_tensor(A[:, 0]).cuda(), torch.as_tensor(A[:, 1]).cuda()
    _lib.check(lib.ecs_dot(a.data_ptr(), c.data_ptr(), n, out.data_ptr(), _runtime._stream()), "ecs_dot")
    assert abs(host(out)[0] - A[:, 0] @ A[:, 1]) < 1e-10


@pytest.mark.parametrize("n", [32, 128])
def test_state_vel_vjp(n):
    print(f"state_vel_vjp n={n}", flush=True)
    pd, ctx = setup(n)
    f = spectrum(n, 51)
    g = spectrum(n, 52)
    for diss in (False, True):
        out = host(ctx.state_vel_vjp(dev(f), dev(g), diss))
        ref = O.state_vel_vjp(f, g, pd, diss)
        assert rel(out, ref) < 1e-12


@pytest.mark.parametrize("n,steps,ckpt", [(32, 10, 4), (64, 16, 16)])
def test_lawson_rk4_vjp(n, steps, ckpt):
    """Checkpointed reverse pass vs the oracle's stored-tape adjoint, and the
    dot test <fbar, J df> = <J^T fbar, df> against the GPU tangent."""
    print(f"lawson_rk4_vjp n={n} steps={steps} ckpt={ckpt}", flush=True)
    pd, ctx = setup(n)
    f = np.stack([spectrum(n, 61), spectrum(n, 62)])
    fbar = np.stack([spectrum(n, 63), spectrum(n, 64)])
    dts = [0.005, 0.004]
    gb, hb = ctx.lawson_rk4_vjp(dev(f), dev(fbar), dts, steps, True, ckpt)
    gb = host(gb)
    for b in range(2):
        rg, rh = O.lawson_rk4_vjp(f[b], dts[b] * steps, steps, pd, fbar[b], True)
        assert rel(gb[b], rg) < 1e-10
        assert abs(hb[b] - rh * steps) < 1e-9 * max(1.0, abs(rh * steps))
    df = np.stack([spectrum(n, 65), spectrum(n, 66)])
    ddt = [0.3, -0.2]
    _, dx = ctx.lawson_rk4_jvp(dev(f), dev(df), dts, ddt, steps)
    dx = host(dx)
    for b in range(2):
        lhs = O.rdot(fbar[b], dx[b])
        rhs = O.rdot(gb[b], df[b]) + hb[b] * ddt[b]
        assert abs(lhs - rhs) < 1e-10 * abs(lhs)


@pytest.mark.parametrize("n", [32, 128])
def test_batched_uneven_rk43_jvp_equals_solo_and_oracle(n):
    """Tangent replay of candidates with different periods (so different
    numbers of accepted steps, some with rejections) in one batch: every
    candidate bit for bit equal to its own one-candidate replay, and to the
    oracle's replay of its schedule."""
    print(f"n={n}", flush=True)
    pd, ctx = setup(n)
    Ts, dTs, atol = [0.35, 0.9, 0.6], [0.4, -0.3, 0.1], 1e-7
    x0 = np.stack([spectrum(n, 60 + b) for b in range(3)])
    dx0 = np.stack([spectrum(n, 70 + b) for b in range(3)])
    x, infos = ctx.lawson_rk43(dev(x0), np.array(Ts), 1e-2, atol, 64, 8)
    assert len({i["accepted"] for i in infos}) == 3
    xj, dxj = ctx.lawson_rk43_jvp(dev(x0), dev(dx0), np.array(Ts), np.array(dTs), infos)
    assert torch.equal(xj, x) or rel(host(xj), host(x)) < 1e-12
    for b in range(3):
        xs, inf_s = ctx.lawson_rk43(dev(x0[b:b + 1]), Ts[b], 1e-2, atol, 64, 8)
        assert inf_s[0]["accepted"] == infos[b]["accepted"] and inf_s[0]["rejected"] == infos[b]["rejected"]
        _, dxs = ctx.lawson_rk43_jvp(dev(x0[b:b + 1]), dev(dx0[b:b + 1]), Ts[b], dTs[b], inf_s)
        assert torch.equal(dxs[0], dxj[b]), b
        if n <= 32:
            sched = [(infos[b]["hs"][i], bool(infos[b]["acc"][i])) for i in range(infos[b]["attempts"])]
            _, rdx = O.lawson_rk43_jvp(x0[b], dx0[b], pd, Ts[b], dTs[b], sched)
            assert rel(host(dxj[b]), rdx) < 1e-10, b


@pytest.mark.parametrize("n", [32, 128])
def test_reverse_pass_batch_16_equals_solo(n):
    """The adjoint kernels take one column / row per block below batch 16 and
    wider tiles from 16 on; tiles do not change the arithmetic, so every
    candidate of a batch of 16 gets the bits of its one-candidate run."""
    print(f"n={n}", flush=True)
    pd, ctx = setup(n)
    B, steps = 16, 6
    f = np.stack([spectrum(n, 70 + b) for b in range(B)])
    fbar = np.stack([spectrum(n, 90 + b) for b in range(B)])
    dts = [0.003 + 0.0002 * b for b in range(B)]
    gb, hb = ctx.lawson_rk4_vjp(dev(f), dev(fbar), dts, steps, True, 4)
    gb = host(gb)
    for b in (0, 7, 15):
        g1, h1 = ctx.lawson_rk4_vjp(dev(f[b:b + 1]), dev(fbar[b:b + 1]), [dts[b]], steps, True, 4)
        assert np.array_equal(gb[b], host(g1)[0]), b
        assert hb[b] == h1[0], b
