"""Write the golden fixtures under tests/golden/ (run in the build container).

1. Converged solutions shipped with the reference (data, read with
   numpy.load(allow_pickle=False), no code executed): a subset of
   /root/reference/solutions/Re40/*.npz (N=128, nu=eta=0.025) and
   /root/reference/solutions/Re50/RPO2.npz (N=256, nu=eta=0.02).  Only the
   numeric entries are kept (fields, T, sx, nu, eta, b0, steps, ministeps,
   num_checkpoints, shift_reflect_ny, rot); the forcing of every file is
   -4 cos(4 y) (its forcing_str).
2. residuals.json: the relative RPO residual |irfft2(f0 - g Phi_T f0)| / |f0|
   of each solution under the oracle's Lawson RK4 with dt = T / steps
   (ministeps x num_checkpoints as stored).
3. dictio/Re40_EQ1.npz: one solution file with every entry of the reference's
   dictionaryIO layout (input_keys, param_keys, forcing_str, ...), re-saved
   entry by entry, for the load_dicts format test.
4. final_states.npz: for each committed solution, the oracle's final state
   of the Lawson-RK4 trajectory of mismatch_RPO (before the symmetry), as the
   compact spectrum of the 2/3 region, rows kx = 0..kmax then -kmax..-1,
   columns ky = 0..kmax (the masked stepper leaves every other mode zero).
   Made from the committed fixtures alone: ``python make_golden.py final``.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import mhd_oracle as O  # noqa: E402

REF = "/root/reference/solutions"
PICK = ["Re40/EQ1", "Re40/RPO4", "Re40/RPO5", "Re40/RPO6", "Re40/TW3", "Re50/RPO2"]
KEYS = ["fields", "T", "sx", "nu", "eta", "b0", "steps", "ministeps", "num_checkpoints", "shift_reflect_ny", "rot"]


def main():
    os.makedirs(os.path.join(HERE, "solutions"), exist_ok=True)
    res = {}
    for name in PICK:
        d = np.load(os.path.join(REF, name + ".npz"), allow_pickle=False)
        assert str(d["forcing_str"]) == "lambda x,y : -4*jnp.cos(4*y)"
        out = {k: np.asarray(d[k]) for k in KEYS}
        tag = name.replace("/", "_")
        np.savez(os.path.join(HERE, "solutions", tag + ".npz"), **out)
        n = out["fields"].shape[-1]
        pd = O.make_params(n, float(out["nu"]), float(out["eta"]), tuple(out["b0"]))
        pd.update({"num_checkpoints": int(out["num_checkpoints"]), "ministeps": int(out["ministeps"]),
                   "shift_reflect_ny": int(out["shift_reflect_ny"]), "rot": bool(out["rot"])})
        inp = {"fields": out["fields"], "T": float(out["T"]), "sx": float(out["sx"])}
        obj, _ = O.objective_RPO(inp, pd, "Lawson_RK4")
        r = float(np.linalg.norm(obj["fields"]) / np.linalg.norm(out["fields"]))
        res[tag] = r
        print(tag, n, r, flush=True)
    with open(os.path.join(HERE, "residuals.json"), "w") as fh:
        json.dump(res, fh, indent=1, sort_keys=True)
    os.makedirs(os.path.join(HERE, "dictio"), exist_ok=True)
    d = np.load(os.path.join(REF, "Re40/EQ1.npz"), allow_pickle=False)
    np.savez(os.path.join(HERE, "dictio", "Re40_EQ1.npz"), **{k: np.asarray(d[k]) for k in d.files})


def final_states():
    out = {}
    for name in PICK:
        tag = name.replace("/", "_")
        d = np.load(os.path.join(HERE, "solutions", tag + ".npz"), allow_pickle=False)
        n = d["fields"].shape[-1]
        pd = O.make_params(n, float(d["nu"]), float(d["eta"]), tuple(d["b0"]))
        nc, ms = int(d["num_checkpoints"]), int(d["ministeps"])
        f = np.fft.rfft2(d["fields"])
        for _ in range(nc):
            f = O.lawson_rk4(f, float(d["T"]) / nc, ms, pd, True)
        kmax = (n - 1) // 3
        out[tag] = np.concatenate([f[:, : kmax + 1, : kmax + 1], f[:, n - kmax:, : kmax + 1]], axis=1)
        print(tag, n, nc * ms, np.linalg.norm(out[tag]), flush=True)
    np.savez(os.path.join(HERE, "final_states.npz"), **out)


if __name__ == "__main__":
    if len(sys.argv) > 1 and sys.argv[1] == "final":
        final_states()
    else:
        main()
        final_states()
