"""Independent torch-CPU float64 restatement of the reference hot path, used only
to cross-check the oracle's hand-derived derivatives with torch autograd.

Written directly from the reference (mhd_jax.py:105-175 state_vel,
timestepping.py:46-81 lawson_rk4, symmetry.py:4-25, loss_functions.py:99-111),
sharing no code with oracle/mhd_oracle.py.
"""
import math

import torch


def domain(n):
    k = torch.fft.fftfreq(n, d=1 / n, dtype=torch.float64)
    kx = k.reshape(-1, 1)
    ky = k[: n // 2 + 1].reshape(1, -1)
    mask = ((kx.abs() < n / 3) & (ky.abs() < n / 3)).to(torch.float64)
    mask[0, 0] = 0.0
    ksq = kx * kx + ky * ky
    ksq[0, 0] = 1.0
    to_u = ky / ksq * mask
    to_v = -kx / ksq * mask
    y = torch.arange(n, dtype=torch.float64) / n * 2 * math.pi
    return {"kx": kx * mask, "ky": ky * mask, "mask": mask, "to_u": to_u, "to_v": to_v,
            "y": y.reshape(1, -1).expand(n, n)}


def state_vel(f, d, nu, eta, b0, forcing, diss):
    fx = torch.fft.irfft2(1j * d["kx"] * f)
    fy = torch.fft.irfft2(1j * d["ky"] * f)
    fu = torch.fft.irfft2(1j * d["to_u"] * f)
    fv = torch.fft.irfft2(1j * d["to_v"] * f)
    u, v = fu[0], fv[0]
    a, b = fu[1] + b0[0], fv[1] + b0[1]
    adv_w = u * fx[0] + v * fy[0]
    adv_j = a * fx[1] + b * fy[1]
    ksq = (d["kx"] ** 2 + d["ky"] ** 2) * d["mask"]
    w_t = d["mask"] * torch.fft.rfft2(-adv_w + adv_j + forcing)
    j_t = ksq * torch.fft.rfft2(u * b - v * a)
    out = torch.stack([w_t, j_t])
    if diss:
        out = out - torch.stack([nu * ksq, eta * ksq]) * f
    return out


def lawson_rk4(f, t, steps, d, nu, eta, b0, forcing):
    h = t / steps
    ksq = d["kx"] ** 2 + d["ky"] ** 2
    L = -torch.stack([nu * ksq, eta * ksq])
    e = torch.exp(h / 2 * L) * d["mask"]
    for _ in range(steps):
        k1 = h * state_vel(f, d, nu, eta, b0, forcing, False)
        f = e * f
        k1 = e * k1
        k2 = h * state_vel(f + k1 / 2, d, nu, eta, b0, forcing, False)
        k3 = h * state_vel(f + k2 / 2, d, nu, eta, b0, forcing, False)
        f = e * f
        k1, k2, k3 = e * k1, e * k2, e * k3
        k4 = h * state_vel(f + k3, d, nu, eta, b0, forcing, False)
        f = f + k1 / 6 + k2 / 3 + k3 / 3 + k4 / 6
    return f


def symmetry(f, sx, ny, rot, d):
    f = torch.exp(1j * d["kx"] * sx) * f
    s = ny * 2 * math.pi / 8
    f = torch.exp(1j * d["ky"] * s) * f
    if ny % 2 == 1:
        f = -torch.roll(torch.flip(f, dims=[-2]), shifts=1, dims=-2)
    if rot:
        f = torch.tensor([1.0, -1.0], dtype=torch.float64).reshape(2, 1, 1) * torch.conj(f)
    return f


def loss_rpo(fields, T, sx, n, nu, eta, b0, nc, ms, ny, rot):
    d = domain(n)
    forcing = -4 * torch.cos(4 * d["y"])
    f0 = torch.fft.rfft2(fields)
    f = f0
    for _ in range(nc):
        f = lawson_rk4(f, T / nc, ms, d, nu, eta, b0, forcing)
    f = symmetry(f, sx, ny, rot, d)
    return torch.mean(torch.fft.irfft2(f - f0) ** 2)
