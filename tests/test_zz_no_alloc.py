"""ecs_state_vel_noalloc: the allocation-free variant of ecs_state_vel,
called through the C-ABI.  Its results equal ecs_state_vel's bit for bit (the
same kernels in the same FFT plan; tiles do not change the arithmetic) on two
calls in a row, the second smaller, and the oracle's state_vel to the RHS
tolerance; a batch beyond the static scratch is refused on the host; and a
call leaves the device's free memory unchanged."""
import ctypes

import numpy as np
import pytest
import torch

from ecs_mhd_2d_amd import _lib, _runtime
from ecs_mhd_2d_amd.synthetic import spectrum
from oracle import mhd_oracle as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def noalloc(ctx, f, diss):
    lib = _lib.load()
    out = torch.empty_like(f)
    rc = lib.ecs_state_vel_noalloc(ctx.handle, f.data_ptr(), out.data_ptr(), f.shape[0], int(diss),
                                   _runtime._stream())
    return rc, out


def inputs(n, B, seed):
    rng = np.random.default_rng(seed)
    f = np.stack([spectrum(n, seed + b) for b in range(B)])
    # modes outside the 2/3 region: ignored except by the dissipation term
    f = f + 1e-3 * np.fft.rfft2(rng.standard_normal((B, 2, n, n)))
    return f


@pytest.mark.parametrize("n,batches", [(128, (6, 2)), (256, (14, 3)), (32, (5, 1))])
def test_noalloc_state_vel_equals_state_vel(n, batches):
    print(f"state_vel_noalloc n={n} batches={batches}", flush=True)
    pd = O.make_params(n, 1 / 40, 1 / 40)
    ctx = _runtime.Context(n, 1 / 40, 1 / 40, (0.0, 0.1), pd["forcing"])
    for i, B in enumerate(batches):
        f = inputs(n, B, 10 * i + 1)
        ft = torch.as_tensor(f).cuda()
        for diss in (False, True):
            rc, a = noalloc(ctx, ft, diss)
            assert rc == 0, _lib.load().ecs_last_error()
            b = ctx.state_vel(ft, diss)
            torch.cuda.synchronize()
            assert torch.equal(a, b), (B, diss)
            ah = a.cpu().numpy()
            for k in (0, B - 1):
                assert rel(ah[k], O.state_vel(f[k], pd, diss)) < (1e-12 if n >= 256 else 1e-13)


def test_noalloc_refuses_a_batch_beyond_its_scratch():
    n = 256
    pd = O.make_params(n, 1 / 40, 1 / 40)
    ctx = _runtime.Context(n, 1 / 40, 1 / 40, (0.0, 0.1), pd["forcing"])
    f = torch.zeros((15, 2, n, n // 2 + 1), dtype=torch.complex128, device="cuda")
    rc, _ = noalloc(ctx, f, False)
    assert rc == 1
    msg = _lib.load().ecs_last_error().decode()
    assert "scratch" in msg and "at most 14" in msg


def test_noalloc_allocates_no_device_memory():
    """A fresh context: one small call loads the kernels, then a call at the
    largest batch leaves hipMemGetInfo's free bytes where they were (the
    allocating entry point would grow the context's workspace by ~60 MB)."""
    n = 256
    pd = O.make_params(n, 1 / 40, 1 / 40)
    ctx = _runtime.Context(n, 1 / 40, 1 / 40, (0.0, 0.1), pd["forcing"])
    small = torch.as_tensor(inputs(n, 1, 3)).cuda()
    big = torch.as_tensor(inputs(n, 14, 4)).cuda()
    out = torch.empty_like(big)
    assert noalloc(ctx, small, True)[0] == 0
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    lib = _lib.load()
    for _ in range(3):
        assert lib.ecs_state_vel_noalloc(ctx.handle, big.data_ptr(), out.data_ptr(), 14, 1, _runtime._stream()) == 0
    torch.cuda.synchronize()
    assert torch.cuda.mem_get_info()[0] == free0
    ref = ctx.state_vel(big, True)
    torch.cuda.synchronize()
    assert torch.equal(out, ref)
