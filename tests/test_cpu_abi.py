"""The C-ABI library loads on a machine without a GPU and exports every
function include/ecsmhd.h declares; argument checks fail cleanly."""
import ctypes

from ecs_mhd_2d_amd import _lib


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    names = _lib.header_symbols()
    assert "ecs_lawson_rk43_jvp" in names and "ecs_state_vel" in names
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) <= set(_lib._SIGS), "binding table misses a declared symbol"


def test_argument_checks_return_errors_without_touching_a_gpu():
    lib = _lib.load()
    h = ctypes.c_void_p()
    assert lib.ecs_ctx_create(ctypes.byref(h), 100, 0.1, 0.1, 0.0, 0.1, None, 0) == 1
    assert b"unsupported" in lib.ecs_last_error()
    assert lib.ecs_state_vel(None, None, None, 1, 0, None) == 1
    assert lib.ecs_lawson_rk43(None, None, 1, None, 0.01, 1e-6, 4, 4, None, None, None, None) == 1
    assert lib.ecs_mgs(None, 1, 1, 0, None, None, None) == 1
    assert lib.ecs_ckpt_doubles(None, 1, 1, 1) == -1
    # the two-phase reverse passes refuse a missing context, buffer or a scheme outside 0..2
    assert lib.ecs_gen_fwd_ckpt(None, None, 1, None, 1, 0, 0, 1, None, None) == 1
    assert lib.ecs_gen_vjp_ckpt(None, None, None, None, 1, None, 1, 3, 0, 1, None) == 1
    assert lib.ecs_lawson_rk43_fwd_ckpt(None, None, 1, None, None, None, None, 1, 1, None, None) == 1
    assert lib.ecs_lawson_rk43_vjp_ckpt(None, None, None, None, 1, None, None, None, None, 1, 1, None) == 1
    assert b"ecs_lawson_rk43_vjp_ckpt" in lib.ecs_last_error()
