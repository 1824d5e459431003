"""Parity of one forced FFT radix plan against the oracle, run as its own
process (the plan is fixed once per process and grid size):

    ECSMHD_PLAN=narrow python tests/plan_check.py

For every grid size: the RHS at batch 3 (1e-13, 1e-12 at N=256), a
64-step Lawson RK4 trajectory (1e-11) and its tangent (1e-10), the RHS VJP at
batch 3 (1e-12) and an 8-step checkpointed Lawson RK4 VJP (1e-10; the adjoint
kernels take the plan too), plus the plan the library reports.  Prints one line per case and exits non-zero on failure.
"""
import ctypes
import os
import sys

import numpy as np
import torch

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
from ecs_mhd_2d_amd import _lib, _runtime  # noqa: E402
from ecs_mhd_2d_amd.synthetic import spectrum  # noqa: E402
from oracle import mhd_oracle as O  # noqa: E402


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def main():
    want = {"wide": 16, "narrow": 8}[os.environ.get("ECSMHD_PLAN", "wide")]
    ok = True
    for n in (16, 32, 64, 128, 256):
        pd = O.make_params(n, 1 / 40, 1 / 40)
        ctx = _runtime.Context(n, 1 / 40, 1 / 40, (0.0, 0.1), pd["forcing"])
        f = np.stack([spectrum(n, s) for s in (1, 2, 3)])
        out = ctx.state_vel(torch.as_tensor(f).cuda(), True).cpu().numpy()
        e_rhs = rel(out, np.stack([O.state_vel(f[b], pd, True) for b in range(3)]))
        steps = 64 if n <= 128 else 16
        x0, dx0 = spectrum(n, 4), spectrum(n, 5)
        dt, ddt = 0.005, 0.3 / steps
        xg, dxg = ctx.lawson_rk4_jvp(torch.as_tensor(x0).cuda(), torch.as_tensor(dx0).cuda(), dt, ddt, steps)
        xr, dxr = O.lawson_rk4_jvp(x0, dx0, dt * steps, ddt * steps, steps, pd, True)
        e_x, e_dx = rel(xg.cpu().numpy(), xr), rel(dxg.cpu().numpy(), dxr)
        # the reverse pass transforms with the same plan: RHS VJP and the
        # checkpointed Lawson RK4 VJP (the reverse of 8 of the steps above)
        g = np.stack([spectrum(n, s) for s in (6, 7, 8)])
        vj = ctx.state_vel_vjp(torch.as_tensor(f).cuda(), torch.as_tensor(g).cuda(), True).cpu().numpy()
        e_vjp = rel(vj, np.stack([O.state_vel_vjp(f[b], g[b], pd, True) for b in range(3)]))
        gb, hb = ctx.lawson_rk4_vjp(torch.as_tensor(x0).cuda(), torch.as_tensor(dx0).cuda(), dt, 8, True, 4)
        rg, rh = O.lawson_rk4_vjp(x0, dt * 8, 8, pd, dx0, True)
        e_rvjp = max(rel(gb.cpu().numpy(), rg), abs(hb[0] - rh * 8) / max(1.0, abs(rh * 8)))
        plans = (ctypes.c_int * 4)()
        _lib.check(_lib.load().ecs_tuned_plans(n, plans), "ecs_tuned_plans")
        good = (e_rhs < (1e-12 if n >= 256 else 1e-13) and e_x < 1e-11 and e_dx < 1e-10
                and e_vjp < 1e-12 and e_rvjp < 1e-10 and list(plans) == [want] * 4)
        ok = ok and good
        print(f"n={n} plans={list(plans)} rhs {e_rhs:.2e} rk4 {e_x:.2e} tangent {e_dx:.2e} "
              f"rhs vjp {e_vjp:.2e} rk4 vjp {e_rvjp:.2e} {'ok' if good else 'FAIL'}", flush=True)
    return 0 if ok else 1


if __name__ == "__main__":
    sys.exit(main())
