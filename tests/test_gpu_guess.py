"""Guess pipeline on the GPU against the oracle: the recurrence-diagram
kernel, the symmetry search, a short turbulent series with Lawson RK6, and
an RPO guess from it evaluated in mode "Lawson_RK6"."""
import numpy as np
import pytest
import torch

from ecs_mhd_2d_amd import loss_functions as LF, mhd, symmetry, turbulence
from ecs_mhd_2d_amd.synthetic import spectrum
from oracle import mhd_oracle as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def test_recurrence_kernel_matches_oracle():
    n, S = 16, 37      # S not a multiple of the 16 x 16 tile
    fs = np.stack([spectrum(n, s) * (1 + 0.01 * s) for s in range(S)])
    d = turbulence.recurrence_diagram(torch.as_tensor(fs).cuda()).cpu().numpy()
    ref = O.recurrence_diagram(fs)
    assert np.array_equal(d, d.T) and np.all(np.diag(d) == 0)
    assert np.abs(d - ref).max() < 1e-12 * ref.max()


def test_find_optimal_symmetry_matches_oracle():
    n, nx = 32, 16
    pd = mhd.make_param_dict(n, 1 / 40, 1 / 40)
    opd = O.make_params(n, 1 / 40, 1 / 40)
    f2 = spectrum(n, 4)
    f1p = O.rot180(O.shift_reflect(O.shift_x(f2, 2 * np.pi / nx * 7, opd), 5, opd))
    f1r = spectrum(n, 5)
    for f1 in (f1p, f1r):
        got = symmetry.find_optimal_symmetry(torch.as_tensor(f1).cuda(), torch.as_tensor(f2).cuda(), pd, nx)
        ref = O.find_optimal_symmetry(f1, f2, opd, nx)
        assert got[1:] == ref[1:] and abs(got[0] - ref[0]) < 1e-15
    assert got is not None


def test_short_turbulent_series_and_guess_match_oracle():
    n, dt, ministeps, steps, transient = 32, 1 / 256, 4, 6, 16
    turb, pd = turbulence.generate_turbulence(n=n, dt=dt, ministeps=ministeps, steps=steps,
                                              transient_steps=transient, b0=(0.0, 0.3), seed=5)
    opd = O.make_params(n, 1 / 40, 1 / 40, b0=(0.0, 0.3))
    rng = np.random.default_rng(5)
    f = opd["mask"] * np.fft.rfft2(10.0 * rng.standard_normal((2, n, n)))
    fs = O.turbulence_series(f, opd, dt, ministeps, steps, transient)
    assert rel(turb["fs"].cpu().numpy(), fs) < 1e-11
    assert rel(turb["dist"].cpu().numpy(), O.recurrence_diagram(fs)) < 1e-11
    inp, pd = turbulence.create_state_from_turb(turb, (2, 5), pd, nx=16)
    oinp, osteps, ony, orot = O.create_state_from_turb(fs, (2, 5), opd, dt, ministeps, nx=16)
    assert (pd["steps"], pd["shift_reflect_ny"], pd["rot"]) == (osteps, ony, orot)
    assert abs(inp["sx"] - oinp["sx"]) < 1e-15 and inp["T"] == oinp["T"]
    assert "dt" not in pd and "ministeps" not in pd
    opd.update({"steps": osteps, "shift_reflect_ny": ony, "rot": orot})
    obj = LF.objective_RPO(inp, pd, "Lawson_RK6")
    ref, _ = O.objective_RPO(oinp, opd, "Lawson_RK6")
    assert rel(obj["fields"].cpu().numpy(), ref["fields"]) < 1e-10
