"""Classical RK4, Lawson RK6 and two-derivative RK4 on the GPU (through the
C-ABI) against the oracle: trajectories to 1e-11 and tangents to 1e-10
relative L2, per-candidate steps in one batch, the unmasked schemes carrying
the modes outside the 2/3 region, and the objective JVP in each mode."""
import numpy as np
import pytest
import torch

from ecs_mhd_2d_amd import _runtime, loss_functions as LF, mhd
from ecs_mhd_2d_amd.synthetic import spectrum
from oracle import mhd_oracle as O

pytestmark = pytest.mark.gpu

NU = ETA = 1 / 40


def setup(n):
    pd = O.make_params(n, NU, ETA)
    return pd, _runtime.Context(n, NU, ETA, (0.0, 0.1), pd["forcing"])


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def oracle_run(scheme, f, dt, steps, pd):
    T = dt * steps
    return {"rk4": lambda: O.rk4(f, T, steps, pd), "lawson_rk6": lambda: O.lawson_rk6(f, T, steps, pd, True),
            "tdrk4": lambda: O.tdrk4(f, T, steps, pd)}[scheme]()


def oracle_jvp(scheme, f, df, dt, ddt, steps, pd):
    T, dT = dt * steps, ddt * steps
    return {"rk4": lambda: O.rk4_jvp(f, df, T, dT, steps, pd),
            "lawson_rk6": lambda: O.lawson_rk6_jvp(f, df, T, dT, steps, pd, True),
            "tdrk4": lambda: O.tdrk4_jvp(f, df, T, dT, steps, pd)}[scheme]()


@pytest.mark.parametrize("scheme", ["rk4", "lawson_rk6", "tdrk4"])
@pytest.mark.parametrize("n", [32, 128])
def test_primal_batch_matches_oracle(scheme, n):
    print(f"{scheme} n={n}", flush=True)
    pd, ctx = setup(n)
    f = np.stack([spectrum(n, 1), spectrum(n, 2)])
    dts, steps = [0.004, 0.005], 12
    out = getattr(ctx, scheme)(torch.as_tensor(f).cuda(), dts, steps).cpu().numpy()
    for b in range(2):
        assert rel(out[b], oracle_run(scheme, f[b], dts[b], steps, pd)) < 1e-11, b


@pytest.mark.parametrize("scheme", ["rk4", "lawson_rk6", "tdrk4"])
@pytest.mark.parametrize("n", [32, 128])
def test_tangent_matches_oracle(scheme, n):
    print(f"{scheme}_jvp n={n}", flush=True)
    pd, ctx = setup(n)
    f = np.stack([spectrum(n, 3), spectrum(n, 4)])
    df = np.stack([spectrum(n, 5), spectrum(n, 6)])
    dts, ddts, steps = [0.005, 0.004], [0.02, -0.01], 8
    x, dx = getattr(ctx, scheme + "_jvp")(torch.as_tensor(f).cuda(), torch.as_tensor(df).cuda(), dts, ddts, steps)
    x, dx = x.cpu().numpy(), dx.cpu().numpy()
    for b in range(2):
        rx, rdx = oracle_jvp(scheme, f[b], df[b], dts[b], ddts[b], steps, pd)
        assert rel(x[b], rx) < 1e-11, b
        assert rel(dx[b], rdx) < 1e-10, b


def test_n256_short_runs():
    n = 256
    pd, ctx = setup(n)
    f = spectrum(n, 7)
    for scheme in ("rk4", "lawson_rk6", "tdrk4"):
        print(f"{scheme} n=256", flush=True)
        out = getattr(ctx, scheme)(torch.as_tensor(f).cuda(), 0.004, 3).cpu().numpy()
        assert rel(out, oracle_run(scheme, f, 0.004, 3, pd)) < 1e-11, scheme


def test_unmasked_schemes_carry_outside_modes():
    """RK4 and TDRK4 have no mask: modes outside the 2/3 region (and the mean)
    are carried unchanged; Lawson RK6 with the mask zeroes them."""
    n = 32
    pd, ctx = setup(n)
    rng = np.random.default_rng(9)
    f = spectrum(n, 8) + 1e-2 * np.fft.rfft2(rng.standard_normal((2, n, n)))
    outside = ~pd["mask"]
    for scheme in ("rk4", "tdrk4"):
        out = getattr(ctx, scheme)(torch.as_tensor(f).cuda(), 0.005, 6).cpu().numpy()
        assert np.array_equal(out[:, outside], f[:, outside])
        assert rel(out, oracle_run(scheme, f, 0.005, 6, pd)) < 1e-11
    out = ctx.lawson_rk6(torch.as_tensor(f).cuda(), 0.005, 6).cpu().numpy()
    assert np.all(out[:, outside] == 0)
    assert rel(out, oracle_run("lawson_rk6", f, 0.005, 6, pd)) < 1e-11


@pytest.mark.parametrize("mode", ["RK4", "Lawson_RK6", "TDRK4"])
def test_objective_and_jvp_in_other_modes(mode):
    print(f"objective {mode}", flush=True)
    n = 32
    pdo = O.make_params(n, NU, ETA)
    extra = {"num_checkpoints": 2, "ministeps": 5, "steps": 10, "shift_reflect_ny": 3, "rot": True}
    pdo.update(extra)
    pd = mhd.make_param_dict(n, NU, ETA)
    pd.update(extra)
    rng = np.random.default_rng(11)
    inp = {"fields": np.fft.irfft2(spectrum(n, 9)), "T": 0.35, "sx": 0.15}
    tan = {"fields": rng.standard_normal((2, n, n)), "T": 0.4, "sx": -0.3}
    gin = {"fields": torch.as_tensor(inp["fields"]).cuda(), "T": inp["T"], "sx": inp["sx"]}
    gtan = {"fields": torch.as_tensor(tan["fields"]).cuda(), "T": tan["T"], "sx": tan["sx"]}
    obj = LF.objective_RPO(gin, pd, mode)
    ref, _ = O.objective_RPO(inp, pdo, mode)
    assert rel(obj["fields"].cpu().numpy(), ref["fields"]) < 1e-11
    jv = LF.objective_RPO_jvp(gin, gtan, pd, mode)
    rj = O.objective_RPO_jvp(inp, tan, pdo, mode)
    assert rel(jv["fields"].cpu().numpy(), rj["fields"]) < 1e-10
    assert abs(float(jv["T"]) - rj["T"]) < 1e-10 * max(1.0, abs(rj["T"]))
    assert abs(float(jv["sx"]) - rj["sx"]) < 1e-10 * max(1.0, abs(rj["sx"]))
