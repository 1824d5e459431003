"""The on-disk format (dictionaryIO.py:11-78) and the oracle's guess pipeline
(find_optimal_symmetry, recurrence diagram), on the CPU."""
import os

import numpy as np
import torch

from ecs_mhd_2d_amd import dictionaryIO, mhd
from ecs_mhd_2d_amd.synthetic import spectrum
from oracle import mhd_oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_load_reference_layout_and_rebuild_grid():
    """A solution file in the reference's layout: keys split by input_keys /
    param_keys, numbers as tensors, rot a bool, forcing rebuilt from
    forcing_str without eval."""
    inp, pd = dictionaryIO.load_dicts(os.path.join(GOLD, "dictio", "Re40_EQ1.npz"), device="cpu")
    assert set(inp) == {"T", "fields", "sx"}
    assert set(pd) == {"forcing_str", "nu", "eta", "b0", "steps", "shift_reflect_ny", "rot", "ministeps",
                       "num_checkpoints"}
    assert isinstance(pd["rot"], bool) and isinstance(pd["forcing_str"], str)
    assert inp["fields"].shape == (2, 128, 128) and inp["fields"].dtype == torch.float64
    pd = dictionaryIO.recompute_grid_information(inp, pd)
    ref = mhd.make_param_dict(128, 0.025, 0.025, device="cpu")
    assert torch.equal(pd["forcing"], ref["forcing"])
    assert torch.equal(pd["mask"], ref["mask"]) and torch.equal(pd["to_u"], ref["to_u"])


def test_save_load_round_trip(tmp_path):
    pd = mhd.make_param_dict(16, 0.02, 0.03, b0=(0.0, 0.2), device="cpu")
    pd.update({"forcing_str": "lambda x,y : -4*jnp.cos(4*y)", "steps": 96, "rot": True, "shift_reflect_ny": 3})
    inp = {"fields": torch.randn(2, 16, 16, dtype=torch.float64), "T": 2.5, "sx": -0.1}
    f = str(tmp_path / "state.npz")
    dictionaryIO.save_dicts(f, inp, dictionaryIO.remove_grid_information(pd))
    inp2, pd2 = dictionaryIO.load_dicts(f, device="cpu")
    assert torch.equal(inp2["fields"], inp["fields"]) and float(inp2["T"]) == 2.5 and float(inp2["sx"]) == -0.1
    assert "mask" not in pd2 and pd2["rot"] is True and int(pd2["steps"]) == 96
    pd2 = dictionaryIO.recompute_grid_information(inp2, pd2)
    assert torch.equal(pd2["forcing"], pd["forcing"])


def test_forcing_str_is_parsed_not_evaluated():
    x = torch.zeros(4, 4, dtype=torch.float64)
    y = torch.linspace(0, 1, 16, dtype=torch.float64).reshape(4, 4)
    f = dictionaryIO.forcing_from_str("lambda x,y : 2.5*jnp.sin(3*y)", x, y)
    assert torch.equal(f, 2.5 * torch.sin(3 * y))
    for bad in ("lambda x,y : __import__('os').system('true')", "lambda x,y : -4*jnp.cos(4*y) + x"):
        try:
            dictionaryIO.forcing_from_str(bad, x, y)
        except ValueError:
            continue
        raise AssertionError(bad)


def test_oracle_find_optimal_symmetry_recovers_a_planted_symmetry():
    n, nx = 16, 16
    pd = O.make_params(n, 1 / 40, 1 / 40)
    f2 = spectrum(n, 3)
    for g in ((2 * np.pi / nx * 5, 3, True), (2 * np.pi / nx * 11, 6, False), (0.0, 1, True)):
        f1 = O.shift_x(f2, g[0], pd)
        f1 = O.shift_reflect(f1, g[1], pd)
        if g[2]:
            f1 = O.rot180(f1)
        got = O.find_optimal_symmetry(f1, f2, pd, nx)
        assert got[1:] == g[1:] and abs(got[0] - g[0]) < 1e-15


def test_oracle_recurrence_diagram_properties():
    n = 16
    base = [spectrum(n, s) for s in range(5)]
    fs = np.stack(base + [2 * base[0]])
    d = O.recurrence_diagram(fs)
    assert np.array_equal(d, d.T) and np.all(np.diag(d) == 0)
    # |2 a| - |a| = |a|: the distance to a doubled snapshot is the snapshot's norm
    assert abs(d[0, 5] - np.linalg.norm(np.abs(base[0]))) < 1e-12 * d[0, 5]
    for i in range(6):
        for j in range(6):
            for k in range(6):
                assert d[i, k] <= d[i, j] + d[j, k] + 1e-12
