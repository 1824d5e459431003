"""Saving and loading of states with their parameters
(jax_scripts/lib/dictionaryIO.py:11-78), in the reference's .npz layout:
every entry of input_dict and param_dict as an array, plus 'input_keys' and
'param_keys' naming which dict each entry came from.

Files are read with numpy.load(allow_pickle=False): the layout holds only
numeric and unicode-string arrays, so nothing needs unpickling (the
reference opens them with allow_pickle=True).  The forcing is rebuilt from
'forcing_str' by matching the expression form the reference writes
("lambda x,y : A*jnp.cos(K*y)", cos or sin, of x or y), never by eval.
"""
from __future__ import annotations

import re

import numpy as np
import torch

from . import mhd

GRID_KEYS = ["x", "y", "kx", "ky", "mask", "to_u", "to_v", "inv_lap", "forcing"]


def _np(v):
    if isinstance(v, torch.Tensor):
        return v.detach().cpu().numpy()
    return np.array(v)


def save_dicts(filename, input_dict, param_dict):
    """np.savez of input_dict | param_dict plus the two key lists (dictionaryIO.py:11-28)."""
    combined = dict(input_dict)
    combined.update(param_dict)
    combined["input_keys"] = list(input_dict.keys())
    combined["param_keys"] = list(param_dict.keys())
    np.savez(filename, **{k: _np(v) for k, v in combined.items()})


def _convert(val, device):
    if isinstance(val, np.ndarray) and np.issubdtype(val.dtype, np.number):
        return torch.as_tensor(val, device=device)
    return val.item() if val.shape == () else val


def load_dicts(filename, device=None):
    """(input_dict, param_dict) from a file of save_dicts (dictionaryIO.py:31-48):
    numeric arrays become torch tensors (on ``device``, default the current
    CUDA device if there is one), strings and booleans plain values."""
    if device is None:
        device = "cuda" if torch.cuda.is_available() else "cpu"
    with np.load(filename, allow_pickle=False) as d:
        input_keys = [str(k) for k in d["input_keys"]]
        param_keys = [str(k) for k in d["param_keys"]]
        input_dict = {k: _convert(d[k], device) for k in input_keys}
        param_dict = {k: _convert(d[k], device) for k in param_keys}
    return input_dict, param_dict


def remove_grid_information(param_dict):
    """param_dict without the grid tables (dictionaryIO.py:52-62)."""
    pd = dict(param_dict)
    for key in GRID_KEYS:
        pd.pop(key, None)
    return pd


_FORCING = re.compile(r"^\s*lambda\s+x\s*,\s*y\s*:\s*([-+]?[0-9.eE+-]+)\s*\*\s*(?:jnp|np)\.(cos|sin)\s*\(\s*"
                      r"([-+]?[0-9.eE+-]+)\s*\*\s*([xy])\s*\)\s*$")


def forcing_from_str(forcing_str, x, y):
    """The field A*cos(K*y) (or sin, or of x) that forcing_str describes; any
    other expression raises ValueError (no eval)."""
    m = _FORCING.match(str(forcing_str))
    if not m:
        raise ValueError(f"unsupported forcing_str {forcing_str!r}: expected 'lambda x,y : A*jnp.cos(K*y)'")
    amp, fn, k, var = float(m.group(1)), m.group(2), float(m.group(3)), m.group(4)
    arg = k * (y if var == "y" else x)
    return amp * (torch.cos(arg) if fn == "cos" else torch.sin(arg))


def recompute_grid_information(input_dict, param_dict):
    """Inverse of remove_grid_information (dictionaryIO.py:66-78): the
    construct_domain tables of the fields' grid and the forcing."""
    fields = input_dict["fields"]
    n = fields.shape[-1]
    device = fields.device if isinstance(fields, torch.Tensor) else "cpu"
    param_dict.update(mhd.construct_domain(n, torch.float64, device))
    param_dict["forcing"] = forcing_from_str(param_dict["forcing_str"], param_dict["x"], param_dict["y"])
    return param_dict
