"""Results under the kernel variants that ECSMHD_VARIANTS forces, run as its
own process (variants are fixed once per process and grid size):

    ECSMHD_VARIANTS="64:9,9,3,3" python tests/variant_check.py out.npz

For every grid size named in ECSMHD_VARIANTS: the RHS and its tangent at
batch 3, a 32-step Lawson RK4 trajectory with its tangent, the primal
trajectory at batch 3, 16 steps of classical RK4 and of Lawson RK6, and an adaptive
Lawson RK4(3) solve with the tangent replay of its schedule (batch 2, uneven
periods, so one candidate finishes early).  The fixed-step trajectory is also
checked against the oracle (1e-11, tangent 1e-10), and the primal of the
tangent replay against the solve it replays (1e-12).  Saves every array to the
npz named on the command line, so two runs can be compared bit for bit, and
exits non-zero on failure.
"""
import os
import sys

import numpy as np
import torch

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
from ecs_mhd_2d_amd import _runtime  # noqa: E402
from ecs_mhd_2d_amd.synthetic import spectrum  # noqa: E402
from oracle import mhd_oracle as O  # noqa: E402


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def main():
    out_path = sys.argv[1]
    grids = [int(tok.split(":")[0]) for tok in os.environ["ECSMHD_VARIANTS"].split(";") if tok]
    ok = True
    res = {}
    for n in grids:
        pd = O.make_params(n, 1 / 40, 1 / 40)
        ctx = _runtime.Context(n, 1 / 40, 1 / 40, (0.0, 0.1), pd["forcing"])
        f = torch.as_tensor(np.stack([spectrum(n, s) for s in (1, 2, 3)])).cuda()
        df = torch.as_tensor(np.stack([spectrum(n, s) for s in (4, 5, 6)])).cuda()
        res[f"rhs{n}"] = ctx.state_vel(f, True).cpu().numpy()
        o, do = ctx.state_vel_jvp(f, df, True)
        res[f"rhs_jvp{n}"], res[f"rhs_jvp_t{n}"] = o.cpu().numpy(), do.cpu().numpy()
        steps, dt, ddt = 32, 0.005, 0.3 / 32
        x0, dx0 = spectrum(n, 7), spectrum(n, 8)
        xg, dxg = ctx.lawson_rk4_jvp(torch.as_tensor(x0).cuda(), torch.as_tensor(dx0).cuda(), dt, ddt, steps)
        xr, dxr = O.lawson_rk4_jvp(x0, dx0, dt * steps, ddt * steps, steps, pd, True)
        res[f"rk4{n}"], res[f"rk4_t{n}"] = xg.cpu().numpy(), dxg.cpu().numpy()
        e_x, e_dx = rel(res[f"rk4{n}"], xr), rel(res[f"rk4_t{n}"], dxr)
        # primal fixed-step Lawson RK4, classical RK4 and Lawson RK6 at batch 3
        # (the column kernels' primal stage and generic-stage modes)
        xp = ctx.lawson_rk4(torch.as_tensor(np.stack([x0, spectrum(n, 13), spectrum(n, 14)])).cuda(), dt, steps)
        res[f"rk4p{n}"] = xp.cpu().numpy()
        e_p = rel(res[f"rk4p{n}"][0], xr)
        res[f"crk4{n}"] = ctx.rk4(f, dt, 16).cpu().numpy()
        res[f"lrk6{n}"] = ctx.lawson_rk6(f, dt, 16).cpu().numpy()
        xa = torch.as_tensor(np.stack([spectrum(n, 9), spectrum(n, 10)])).cuda()
        dxa = torch.as_tensor(np.stack([spectrum(n, 11), spectrum(n, 12)])).cuda()
        T = np.array([0.3, 0.5])
        xT, infos = ctx.lawson_rk43(xa, T, 0.01, 1e-6, 64, 4)
        xj, dxj = ctx.lawson_rk43_jvp(xa, dxa, T, np.array([0.2, -0.1]), infos)
        res[f"rk43{n}"], res[f"rk43_jvp{n}"], res[f"rk43_jvp_t{n}"] = (xT.cpu().numpy(), xj.cpu().numpy(),
                                                                     dxj.cpu().numpy())
        e_rep = rel(res[f"rk43_jvp{n}"], res[f"rk43{n}"])
        good = e_x < 1e-11 and e_p < 1e-11 and e_dx < 1e-10 and e_rep < 1e-12 and all(i["completed"] for i in infos)
        ok = ok and good
        print(f"n={n} rk4 {e_x:.2e} primal batch {e_p:.2e} tangent {e_dx:.2e} replayed primal {e_rep:.1e} "
              f"steps {[i['accepted'] for i in infos]} {'ok' if good else 'FAIL'}", flush=True)
    np.savez(out_path, **res)
    return 0 if ok else 1


if __name__ == "__main__":
    sys.exit(main())
