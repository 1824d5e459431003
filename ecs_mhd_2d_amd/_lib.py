"""ctypes binding of libecsmhd.so (the C-ABI declared in include/ecsmhd.h).

The library is built in-tree by ``__graft_entry__.build()`` (hipcc,
--offload-arch=gfx950).  There is no fallback: if the library is missing or a
call fails, an exception is raised.
"""
from __future__ import annotations

import ctypes
import os
import re

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, "libecsmhd.so")
HEADER = os.path.join(os.path.dirname(_HERE), "include", "ecsmhd.h")

c_int, c_double, c_void_p = ctypes.c_int, ctypes.c_double, ctypes.c_void_p
P_double = ctypes.POINTER(ctypes.c_double)
P_int8 = ctypes.POINTER(ctypes.c_int8)
P_int = ctypes.POINTER(ctypes.c_int)


class RK43Info(ctypes.Structure):
    _fields_ = [("completed", c_int), ("accepted", c_int), ("rejected", c_int), ("fevals", c_int),
                ("attempts", c_int), ("pad", c_int), ("s", c_double)]


_SIGS = {
    "ecs_ctx_create": (c_int, [ctypes.POINTER(c_void_p), c_int, c_double, c_double, c_double, c_double,
                                P_double, c_int]),
    "ecs_ctx_destroy": (c_int, [c_void_p]),
    "ecs_last_error": (ctypes.c_char_p, []),
    "ecs_ctx_n": (c_int, [c_void_p]),
    "ecs_ctx_set_rev_overlap": (c_int, [c_void_p, c_int]),
    "ecs_state_vel": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_int, c_void_p]),
    "ecs_state_vel_noalloc": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_int, c_void_p]),
    "ecs_state_vel_jvp": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int, c_void_p]),
    "ecs_state_vel_vjp": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int, c_void_p]),
    "ecs_lawson_rk4": (c_int, [c_void_p, c_void_p, c_int, P_double, c_int, c_int, c_void_p]),
    "ecs_lawson_rk4_jvp": (c_int, [c_void_p, c_void_p, c_void_p, c_int, P_double, P_double, c_int, c_int,
                                   c_void_p]),
    "ecs_lawson_rk4_vjp": (c_int, [c_void_p, c_void_p, c_void_p, P_double, c_int, P_double, c_int, c_int,
                                   c_int, c_void_p]),
    "ecs_ckpt_doubles": (ctypes.c_longlong, [c_void_p, c_int, c_int, c_int]),
    "ecs_lawson_rk4_fwd_ckpt": (c_int, [c_void_p, c_void_p, c_int, P_double, c_int, c_int, c_int, c_void_p,
                                        P_int, c_void_p]),
    "ecs_lawson_rk4_vjp_ckpt": (c_int, [c_void_p, c_void_p, c_void_p, P_double, c_int, P_double, c_int, c_int,
                                        c_int, P_int, c_void_p]),
    "ecs_lawson_rk43": (c_int, [c_void_p, c_void_p, c_int, P_double, c_double, c_double, c_int, c_int,
                                ctypes.POINTER(RK43Info), P_double, P_int8, c_void_p]),
    "ecs_lawson_rk43_jvp": (c_int, [c_void_p, c_void_p, c_void_p, c_int, P_double, P_double, P_double, P_int8,
                                    P_int, c_int, c_void_p]),
    "ecs_time_kernel": (c_int, [c_void_p, c_int, c_int, c_int, P_double, c_void_p]),
    "ecs_replay_probe": (c_int, [c_void_p, c_int, c_int, P_double, P_int]),
    "ecs_rk4": (c_int, [c_void_p, c_void_p, c_int, P_double, c_int, c_void_p]),
    "ecs_rk4_jvp": (c_int, [c_void_p, c_void_p, c_void_p, c_int, P_double, P_double, c_int, c_void_p]),
    "ecs_lawson_rk6": (c_int, [c_void_p, c_void_p, c_int, P_double, c_int, c_int, c_void_p]),
    "ecs_lawson_rk6_jvp": (c_int, [c_void_p, c_void_p, c_void_p, c_int, P_double, P_double, c_int, c_int,
                                   c_void_p]),
    "ecs_tdrk4": (c_int, [c_void_p, c_void_p, c_int, P_double, c_int, c_void_p]),
    "ecs_tdrk4_jvp": (c_int, [c_void_p, c_void_p, c_void_p, c_int, P_double, P_double, c_int, c_void_p]),
    "ecs_lawson_rk43_vjp": (c_int, [c_void_p, c_void_p, c_void_p, P_double, c_int, P_double, P_double, P_int8,
                                    P_int, c_int, c_int, c_void_p]),
    "ecs_rk4_vjp": (c_int, [c_void_p, c_void_p, c_void_p, P_double, c_int, P_double, c_int, c_int, c_void_p]),
    "ecs_lawson_rk6_vjp": (c_int, [c_void_p, c_void_p, c_void_p, P_double, c_int, P_double, c_int, c_int, c_int,
                                   c_void_p]),
    "ecs_tdrk4_vjp": (c_int, [c_void_p, c_void_p, c_void_p, P_double, c_int, P_double, c_int, c_int, c_void_p]),
    "ecs_gen_fwd_ckpt": (c_int, [c_void_p, c_void_p, c_int, P_double, c_int, c_int, c_int, c_int, c_void_p,
                                 c_void_p]),
    "ecs_gen_vjp_ckpt": (c_int, [c_void_p, c_void_p, c_void_p, P_double, c_int, P_double, c_int, c_int, c_int,
                                 c_int, c_void_p]),
    "ecs_lawson_rk43_fwd_ckpt": (c_int, [c_void_p, c_void_p, c_int, P_double, P_double, P_int8, P_int, c_int,
                                         c_int, c_void_p, c_void_p]),
    "ecs_lawson_rk43_vjp_ckpt": (c_int, [c_void_p, c_void_p, c_void_p, P_double, c_int, P_double, P_double,
                                         P_int8, P_int, c_int, c_int, c_void_p]),
    "ecs_tuned_tiles": (c_int, [c_int, c_int, P_int]),
    "ecs_tuned_plans": (c_int, [c_int, P_int]),
    "ecs_tuned_splits": (c_int, [c_int, c_int, P_int]),
    "ecs_tuned_variants": (c_int, [c_int, c_int, P_int]),
    "ecs_tuned_replay_variants": (c_int, [c_int, c_int, P_int]),
    "ecs_mgs": (c_int, [c_void_p, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p]),
    "ecs_mgs_reorth": (c_int, [c_void_p, c_int, c_int, c_int, c_void_p, c_void_p, c_int, c_void_p]),
    "ecs_dot": (c_int, [c_void_p, c_void_p, c_int, c_void_p, c_void_p]),
    "ecs_recurrence": (c_int, [c_void_p, c_int, ctypes.c_longlong, c_void_p, c_void_p]),
}

_lib = None


def header_symbols() -> list[str]:
    """Names of every function declared in include/ecsmhd.h."""
    with open(HEADER) as fh:
        txt = fh.read()
    return sorted(set(re.findall(r"\b(ecs_[a-z0-9_]+)\s*\(", txt)))


def load():
    """Load the in-tree library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBPATH):
        raise RuntimeError(f"{LIBPATH} is missing: build it with __graft_entry__.build() "
                           "(the HIP path has no fallback)")
    lib = ctypes.CDLL(LIBPATH)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc: int, what: str):
    if rc != 0:
        msg = load().ecs_last_error().decode(errors="replace")
        raise RuntimeError(f"{what} failed (code {rc}): {msg}")


def dptr(values) -> ctypes.Array:
    arr = (ctypes.c_double * len(values))(*[float(v) for v in values])
    return arr
