"""Independent torch-CPU float64 restatement of the reference hot path, used only
to cross-check the oracle's hand-derived derivatives with torch autograd.

Written directly from the reference (mhd_jax.py:105-175 state_vel,
timestepping.py:46-81 lawson_rk4, :151-335 lawson_rk43, symmetry.py:4-25,
loss_functions.py:99-111), sharing no code with oracle/mhd_oracle.py.
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


def _v(f, d, nu, eta, b0, forcing, diss):
    return state_vel(f, d, nu, eta, b0, forcing, diss)


def rk4_classic(f, t, steps, d, nu, eta, b0, forcing):
    """timestepping.py:17-44 with the dissipative RHS (loss_functions.py:53)."""
    h = t / steps
    for _ in range(steps):
        k1 = h * _v(f, d, nu, eta, b0, forcing, True)
        k2 = h * _v(f + 0.5 * k1, d, nu, eta, b0, forcing, True)
        k3 = h * _v(f + 0.5 * k2, d, nu, eta, b0, forcing, True)
        k4 = h * _v(f + k3, d, nu, eta, b0, forcing, True)
        f = f + (k1 + 2 * k2 + 2 * k3 + k4) / 6
    return f


def lawson_rk6(f, t, steps, d, nu, eta, b0, forcing):
    """timestepping.py:83-125 written stage by stage."""
    h = t / steps
    ksq = d["kx"] ** 2 + d["ky"] ** 2
    e = torch.exp(h / 6 * (-torch.stack([nu * ksq, eta * ksq]))) * d["mask"]
    v = lambda z: h * _v(z, d, nu, eta, b0, forcing, False)
    for _ in range(steps):
        k1 = v(f)
        f, k1 = e * f, e * k1
        k2 = v(f + k1 / 6)
        k3 = v(f + (k1 + k2) / 12)
        f, k1, k2, k3 = e * f, e * k1, e * k2, e * k3
        k4 = v(f + (-4 * k2 + 15 * k3) / 33)
        f, k1, k2, k3, k4 = e * f, e * k1, e * k2, e * k3, e * k4
        k5 = v(f + (-11 * k1 - 29 * k2 + 62 * k3) / 44)
        f, k1, k2, k3, k4, k5 = e * f, e * k1, e * k2, e * k3, e * k4, e * k5
        k6 = v(f + (9 * k1 + 8 * k2 - 12 * k3 + 3 * k4 + 14 * k5) / 33)
        f, k1, k2, k3, k4, k5, k6 = e * f, e * k1, e * k2, e * k3, e * k4, e * k5, e * k6
        k7 = v(f + (-17 * k1 - 20 * k2 + 48 * k3 + 48 * k4 - 52 * k5 + 33 * k6) / 48)
        f, k1, k2, k3, k4, k5, k6, k7 = e * f, e * k1, e * k2, e * k3, e * k4, e * k5, e * k6, e * k7
        k8 = v(f + (20 * k1 + 12 * k2 - 31 * k3 - k4 + 34 * k5 - 11 * k6 + 16 * k7) / 39)
        f = f + (13 * (k1 + k8) / 200 + 4 * (k3 + k7) / 25 + 11 * (k4 + k6) / 40)
    return f


def tdrk4(f, t, steps, d, nu, eta, b0, forcing):
    """timestepping.py:127-149: the second stage derivative is a forward-mode
    derivative of the RHS, taken here with torch.func.jvp as the reference
    takes it with jax.jvp."""
    h = t / steps
    vf = lambda z: _v(z, d, nu, eta, b0, forcing, True)
    jac = lambda z, k: torch.func.jvp(vf, (z,), (k,))[1]
    for _ in range(steps):
        k1 = h * vf(f)
        a1 = h * jac(f, k1)
        g = f + k1 / 2 + a1 / 8
        k2 = h * vf(g)
        a2 = h * jac(g, k2)
        f = f + k1 + a1 / 6 + a2 / 3
    return f


def lawson_rk43(x, t, d, nu, eta, b0, forcing, h=1e-2, atol=1e-4, max_steps_per_checkpoint=1024,
                checkpoints=1):
    """timestepping.py:151-335 written from the reference's loop body.

    Dynamics in pseudo-time s in [0, 1]: dx/ds = t (L x + v(x)) (:242-243).
    Each attempt multiplies the carried k1 by e = exp(t L h / 2) twice
    (:267, :276); a rejected attempt keeps x0 but carries that e^2 k1 into
    the next attempt (:297).
    Returns (x, [(h, accepted), ...], s).
    """
    ksq = d["kx"] ** 2 + d["ky"] ** 2
    tL = t * (-torch.stack([nu * ksq, eta * ksq]))
    vel = lambda z: t * state_vel(z, d, nu, eta, b0, forcing, False)
    s = 0.0
    k1 = vel(x)
    attempts = []
    for _ in range(int(max_steps_per_checkpoint) * int(checkpoints)):
        if abs(s - 1.0) < 1e-10:
            break
        e = torch.exp(tL * h / 2)
        xa = e * x
        ka = e * k1
        k2 = vel(xa + h * ka / 2)
        k3 = vel(xa + h * k2 / 2)
        xa, ka, k2, k3 = e * xa, e * ka, e * k2, e * k3
        k4 = vel(xa + h * k3)
        xf = xa + h * (ka / 6 + k2 / 3 + k3 / 3 + k4 / 6)
        k5 = vel(xf)
        err = float(torch.fft.irfft2(h * (k5 - k4) / 6).abs().max())
        ok = err < atol
        attempts.append((h, ok))
        if ok:
            x, k1, s = xf, k5, s + h
        else:
            k1 = ka
        h = h / 2 if math.isnan(err) else h * 0.9 * (atol / err) ** 0.25
        over = s + h - 1.0
        if math.isnan(over) or over > 0:
            h = h - over
    return x, attempts, s
