"""CPU oracle for the 2D MHD periodic-orbit hot path -- TEST INFRASTRUCTURE ONLY.

This module is the checker, never the thing measured or shipped.  Only files
under ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it.  The product path (``ecs_mhd_2d_amd``) never calls
into it and fails loudly when its HIP library is missing.

It is a plain FP64 numpy restatement of the reference's algorithm
(mgolden30/ECS_MHD_2D, ``jax_scripts/lib``), one function per reference
function, each citing the file:line it follows.  FFTs are numpy's pocketfft
with numpy's default ("backward") normalisation, which is the convention of
``jnp.fft`` (inverse carries 1/N^2).

Derivatives (JVP/VJP) are hand-derived here, because the reference gets them
from JAX autodiff.  Conventions:

* a spectral cotangent ``c`` of a complex array ``z`` holds d(loss)/d(Re z) in
  its real part and d(loss)/d(Im z) in its imaginary part, i.e. it is the
  gradient for the real inner product <a, b> = sum Re(conj(a) * b);
* the real-space API cotangents (``fields``, ``T``, ``sx``) are ordinary
  gradients.

How each function is pinned (tests/test_oracle.py): closed forms (single
Fourier mode, laminar Kolmogorov-Alfven equilibrium), invariants (ideal
energy / cross-helicity conservation, symmetry group laws and equivariance),
exact linear decay, empirical 4th order, the reference's converged solutions
(``solutions/*.npz``, restated as committed fixtures), central finite
differences and adjoint dot tests for every derivative, and an independent
torch-CPU float64 restatement differentiated by torch autograd.
The one part nothing pins bit-exactly is the RK4(3) accept/reject sequence
at the ``err < atol`` boundary (see ``lawson_rk43``), which is pinned only
through its invariants (fevals = 1 + 4 * attempts, completion, order).
"""
from __future__ import annotations

import numpy as np

# 2D real transforms: numpy's pocketfft (the default), or scipy.fft's with
# several worker threads for the CPU baseline timing (set_fft_workers).
_FFT = {"workers": 1}


def _rfft2(x, s=None):
    if _FFT["workers"] > 1:
        import scipy.fft
        return scipy.fft.rfft2(x, s=s, workers=_FFT["workers"])
    return np.fft.rfft2(x, s=s)


def _irfft2(x, s=None):
    if _FFT["workers"] > 1:
        import scipy.fft
        return scipy.fft.irfft2(x, s=s, workers=_FFT["workers"])
    return np.fft.irfft2(x, s=s)


def set_fft_workers(k: int) -> None:
    """Route the oracle's 2D transforms through scipy.fft with k threads (k > 1)
    or numpy.fft (k = 1).  Same transforms; only the CPU timing changes."""
    _FFT["workers"] = max(1, int(k))


# mhd_jax.py:47 hard-codes pi as this literal; symmetry.py:11 uses jnp.pi.
# Both round to the same double.
PI_LITERAL = 3.14159265358979323846


# --------------------------------------------------------------------------
# spectral domain  (mhd_jax.py:20-94, :97-102)
# --------------------------------------------------------------------------
def construct_domain(n: int) -> dict:
    """Grid, wavenumbers, 2/3 mask and uncurl tables (mhd_jax.py:20-94).

    x varies along axis -2, y along axis -1 (meshgrid 'ij', :52).  kx, ky,
    to_u, to_v are pre-multiplied by the mask (:87-90), so derivatives are
    zero outside the mask and the (0,0) mode is excluded (:65-70).
    """
    grid = np.arange(n, dtype=np.float64) / n * 2 * PI_LITERAL
    x, y = np.meshgrid(grid, grid, indexing="ij")
    k = np.fft.fftfreq(n, d=1 / n)
    kx = k.reshape(-1, 1)
    ky = k[: n // 2 + 1].reshape(1, -1)
    mask = (np.abs(kx) < n / 3) & (np.abs(ky) < n / 3)
    mask[0, 0] = False
    k_sq = kx * kx + ky * ky
    k_sq[0, 0] = 1
    inv_lap = -1 / k_sq
    inv_lap[0, 0] = 0
    inv_lap = inv_lap * mask
    to_u = ky / k_sq
    to_v = -kx / k_sq
    kx = kx * mask
    ky = ky * mask
    to_u = to_u * mask
    to_v = to_v * mask
    return {"x": x, "y": y, "kx": kx, "ky": ky, "mask": mask,
            "to_u": to_u, "to_v": to_v, "inv_lap": inv_lap}


def make_params(n: int, nu: float, eta: float, b0=(0.0, 0.1),
                forcing_amp: float = -4.0, forcing_k: float = 4.0) -> dict:
    """param_dict as the drivers build it: domain + nu, eta, b0, forcing.

    The forcing is -4 cos(4 y) in every shipped solution
    (generate_turbulence.py:49-50, ``forcing_str`` of solutions/*.npz).
    """
    pd = construct_domain(n)
    pd.update({"nu": float(nu), "eta": float(eta),
               "b0": np.array(b0, dtype=np.float64),
               "forcing": forcing_amp * np.cos(forcing_k * pd["y"])})
    return pd


def dissipation(pd: dict) -> np.ndarray:
    """L = -(kx^2 + ky^2) * [nu; eta] with masked kx, ky (mhd_jax.py:97-102)."""
    k_sq = pd["kx"] ** 2 + pd["ky"] ** 2
    coeffs = np.array([pd["nu"], pd["eta"]]).reshape(2, 1, 1)
    return -k_sq * coeffs


# --------------------------------------------------------------------------
# right-hand side  (mhd_jax.py:105-175) and its tangent / adjoint
# --------------------------------------------------------------------------
def _real_fields(f: np.ndarray, pd: dict):
    """The 8 real-space fields of state_vel (mhd_jax.py:136-149).

    Returns fx, fy, fu, fv, each [..., 2, N, N]; fu/fv of the current j get
    the mean field b0 added (:145-149).
    """
    fx = _irfft2(1j * pd["kx"] * f)
    fy = _irfft2(1j * pd["ky"] * f)
    fu = _irfft2(1j * pd["to_u"] * f)
    fv = _irfft2(1j * pd["to_v"] * f)
    return fx, fy, fu, fv


def state_vel(f: np.ndarray, pd: dict, include_dissipation: bool) -> np.ndarray:
    """MHD state velocity of (omega, j) in Fourier space (mhd_jax.py:105-175)."""
    kx, ky, mask = pd["kx"], pd["ky"], pd["mask"]
    fx, fy, fu, fv = _real_fields(f, pd)
    b0 = pd["b0"]
    fu = fu + np.array([0.0, b0[0]]).reshape(2, 1, 1)
    fv = fv + np.array([0.0, b0[1]]).reshape(2, 1, 1)
    adv = fu * fx + fv * fy
    k_sq = (kx * kx + ky * ky) * mask
    dwdt = -adv[..., 0, :, :] + adv[..., 1, :, :] + pd["forcing"]
    dwdt = mask * _rfft2(dwdt)
    djdt = fu[..., 0, :, :] * fv[..., 1, :, :] - fv[..., 0, :, :] * fu[..., 1, :, :]
    djdt = k_sq * _rfft2(djdt)
    out = np.stack([dwdt, djdt], axis=-3)
    if include_dissipation:
        out = out + dissipation(pd) * f
    return out


def state_vel_jvp(f, df, pd, include_dissipation=False):
    """Tangent of state_vel: d/de state_vel(f + e df) at e = 0.

    Linearisation of mhd_jax.py:136-175; the same products as the reference's
    hand-derived tangent (deprecated/derive_adjoint_looping/functions/
    tangent_vector_velocity.m:1-43), written for the (omega, j) spectral form.
    """
    kx, ky, mask = pd["kx"], pd["ky"], pd["mask"]
    fx, fy, fu, fv = _real_fields(f, pd)
    b0 = pd["b0"]
    fu = fu + np.array([0.0, b0[0]]).reshape(2, 1, 1)
    fv = fv + np.array([0.0, b0[1]]).reshape(2, 1, 1)
    tx, ty, tu, tv = _real_fields(df, pd)
    dadv = tu * fx + fu * tx + tv * fy + fv * ty
    k_sq = (kx * kx + ky * ky) * mask
    dw = mask * _rfft2(-dadv[..., 0, :, :] + dadv[..., 1, :, :])
    dj = k_sq * _rfft2(tu[..., 0, :, :] * fv[..., 1, :, :] + fu[..., 0, :, :] * tv[..., 1, :, :]
                             - tv[..., 0, :, :] * fu[..., 1, :, :] - fv[..., 0, :, :] * tu[..., 1, :, :])
    out = np.stack([dw, dj], axis=-3)
    if include_dissipation:
        out = out + dissipation(pd) * df
    return out


def _col_weight(n: int, c: int) -> np.ndarray:
    """Per-column weight over the rfft half plane: 1 on ky=0 and ky=N/2, c elsewhere."""
    w = np.full(n // 2 + 1, float(c))
    w[0] = 1.0
    w[-1] = 1.0
    return w.reshape(1, -1)


def irfft2_T(g: np.ndarray) -> np.ndarray:
    """Real transpose of numpy's irfft2 (complex cotangent convention above).

    irfft2(c)[n] = (1/N^2) sum_k w'_k Re(c_k e^{i theta_kn}) with w' = 1 on
    the ky=0/N/2 columns and 2 elsewhere, so its transpose is
    (w'/N^2) * rfft2(g).
    """
    n = g.shape[-1]
    return _col_weight(n, 2) / (n * n) * _rfft2(g)


def rfft2_T(c: np.ndarray) -> np.ndarray:
    """Real transpose of numpy's rfft2: N^2 * irfft2(w * c), w = 1 on ky=0/N/2, 1/2 elsewhere."""
    n = c.shape[-2]
    w = np.full((1, n // 2 + 1), 0.5)
    w[0, 0] = 1.0
    w[0, -1] = 1.0
    return n * n * _irfft2(w * c, s=(n, n))


def state_vel_vjp(f, g, pd, include_dissipation=False):
    """Adjoint of state_vel_jvp: returns J(f)^T g (real inner product on spectra).

    Transpose of the linearisation of mhd_jax.py:136-175.  The reference's
    own hand-derived cotangent RHS (deprecated/derive_adjoint_looping/
    functions/cotangent_vector_velocity.m:1-59) is the same operator in a
    real-space formulation.
    """
    kx, ky, mask = pd["kx"], pd["ky"], pd["mask"]
    fx, fy, fu, fv = _real_fields(f, pd)
    b0 = pd["b0"]
    fu = fu + np.array([0.0, b0[0]]).reshape(2, 1, 1)
    fv = fv + np.array([0.0, b0[1]]).reshape(2, 1, 1)
    k_sq = (kx * kx + ky * ky) * mask
    # cotangents of the two real-space products
    pw = rfft2_T(mask * g[..., 0, :, :])
    pj = rfft2_T(k_sq * g[..., 1, :, :])
    u, v, a, b = fu[..., 0, :, :], fv[..., 0, :, :], fu[..., 1, :, :], fv[..., 1, :, :]
    wx, wy, jx, jy = fx[..., 0, :, :], fy[..., 0, :, :], fx[..., 1, :, :], fy[..., 1, :, :]
    # p_w = -(u wx + v wy) + (a jx + b jy);  p_j = u b - v a
    bar_u = -pw * wx + pj * b
    bar_v = -pw * wy - pj * a
    bar_a = pw * jx - pj * v
    bar_b = pw * jy + pj * u
    bar_wx = -pw * u
    bar_wy = -pw * v
    bar_jx = pw * a
    bar_jy = pw * b
    bfx = np.stack([bar_wx, bar_jx], axis=-3)
    bfy = np.stack([bar_wy, bar_jy], axis=-3)
    bfu = np.stack([bar_u, bar_a], axis=-3)
    bfv = np.stack([bar_v, bar_b], axis=-3)
    # d = irfft2(i m f)  ->  fbar += conj(i m) irfft2_T(dbar) = -i m irfft2_T(dbar)
    out = (-1j * pd["kx"] * irfft2_T(bfx) - 1j * pd["ky"] * irfft2_T(bfy)
           - 1j * pd["to_u"] * irfft2_T(bfu) - 1j * pd["to_v"] * irfft2_T(bfv))
    if include_dissipation:
        out = out + dissipation(pd) * g
    return out


def rdot(a, b) -> float:
    """Real inner product sum Re(conj(a) b) over all elements."""
    return float(np.sum(np.real(np.conj(a) * b)))


# --------------------------------------------------------------------------
# Lawson RK4, fixed step  (timestepping.py:46-81), with tangent and adjoint
# --------------------------------------------------------------------------
def _lawson_e(pd, dt, masked: bool):
    L = dissipation(pd)
    e = np.exp(dt / 2 * L)
    if masked:
        e = e * pd["mask"]
    return L, e


def lawson_rk4(f, t, steps, pd, mask=True):
    """Lawson RK4 (timestepping.py:46-81); v_fn excludes dissipation
    (loss_functions.py:61) and e = exp(h/2 L) * mask (:70-72)."""
    h = t / steps
    _, e = _lawson_e(pd, h, mask)
    v = lambda z: state_vel(z, pd, False)
    for _ in range(int(steps)):
        k1 = h * v(f)
        f = e * f
        k1 = e * k1
        k2 = h * v(f + k1 / 2)
        k3 = h * v(f + k2 / 2)
        f = e * f
        k1 = e * k1
        k2 = e * k2
        k3 = e * k3
        k4 = h * v(f + k3)
        f = f + k1 / 6 + k2 / 3 + k3 / 3 + k4 / 6
    return f


def lawson_rk4_jvp(f, df, t, dt_t, steps, pd, mask=True):
    """Forward-mode derivative of lawson_rk4 w.r.t. (f, t).

    h = t/steps depends on t, so dh = dt_t/steps enters through every stage
    scale and through e = exp(h/2 L) (de = dh/2 L e).  Returns (f_T, df_T).
    """
    h = t / steps
    dh = dt_t / steps
    L, e = _lawson_e(pd, h, mask)
    de = dh / 2 * L * e
    v = lambda z: state_vel(z, pd, False)
    jv = lambda z, dz: state_vel_jvp(z, dz, pd, False)
    for _ in range(int(steps)):
        v1 = v(f)
        k1, dk1 = h * v1, dh * v1 + h * jv(f, df)
        f, df = e * f, de * f + e * df
        k1, dk1 = e * k1, de * k1 + e * dk1
        y2, dy2 = f + k1 / 2, df + dk1 / 2
        v2 = v(y2)
        k2, dk2 = h * v2, dh * v2 + h * jv(y2, dy2)
        y3, dy3 = f + k2 / 2, df + dk2 / 2
        v3 = v(y3)
        k3, dk3 = h * v3, dh * v3 + h * jv(y3, dy3)
        f, df = e * f, de * f + e * df
        k1, dk1 = e * k1, de * k1 + e * dk1
        k2, dk2 = e * k2, de * k2 + e * dk2
        k3, dk3 = e * k3, de * k3 + e * dk3
        y4, dy4 = f + k3, df + dk3
        v4 = v(y4)
        k4, dk4 = h * v4, dh * v4 + h * jv(y4, dy4)
        f = f + k1 / 6 + k2 / 3 + k3 / 3 + k4 / 6
        df = df + dk1 / 6 + dk2 / 3 + dk3 / 3 + dk4 / 6
    return f, df


def lawson_rk4_vjp(f0, t, steps, pd, fbar, mask=True):
    """Reverse-mode derivative of lawson_rk4: returns (f0bar, tbar) for cotangent fbar.

    Stores every step's input (the reference rematerialises per segment with
    jax.checkpoint, loss_functions.py:63; the result is the same).  One step
    is written as  y2 = e(f + k1/2), y3 = e f + k2/2, y4 = e^2 f + e k3,
    f' = e^2 f + e^2 k1/6 + e k2/3 + e k3/3 + k4/6 with k_s = h v(y_s).
    """
    h = t / steps
    L, e = _lawson_e(pd, h, mask)
    v = lambda z: state_vel(z, pd, False)
    vt = lambda z, g: state_vel_vjp(z, g, pd, False)
    tape = []
    f = f0
    for _ in range(int(steps)):
        tape.append(f)
        f = lawson_rk4(f, h, 1, pd, mask)
    hbar = 0.0
    ebar = np.zeros(e.shape)
    gf = fbar
    for f in reversed(tape):
        v1 = v(f); k1 = h * v1
        y2 = e * (f + k1 / 2); v2 = v(y2); k2 = h * v2
        y3 = e * f + k2 / 2; v3 = v(y3); k3 = h * v3
        y4 = e * e * f + e * k3; v4 = v(y4); k4 = h * v4
        # f' = e^2 f + e^2 k1/6 + e k2/3 + e k3/3 + k4/6
        k4b = gf / 6
        fb = e * e * gf
        k1b = e * e * gf / 6
        k2b = e * gf / 3
        k3b = e * gf / 3
        ebar += np.real(np.conj(gf) * (2 * e * f + 2 * e * k1 / 6 + k2 / 3 + k3 / 3))
        # k4 = h v(y4)
        hbar += rdot(v4, k4b)
        y4b = vt(y4, h * k4b)
        fb += e * e * y4b
        k3b += e * y4b
        ebar += np.real(np.conj(y4b) * (2 * e * f + k3))
        # k3 = h v(y3), y3 = e f + k2/2
        hbar += rdot(v3, k3b)
        y3b = vt(y3, h * k3b)
        fb += e * y3b
        k2b += y3b / 2
        ebar += np.real(np.conj(y3b) * f)
        # k2 = h v(y2), y2 = e f + e k1/2
        hbar += rdot(v2, k2b)
        y2b = vt(y2, h * k2b)
        fb += e * y2b
        k1b += e * y2b / 2
        ebar += np.real(np.conj(y2b) * (f + k1 / 2))
        # k1 = h v(f)
        hbar += rdot(v1, k1b)
        fb += vt(f, h * k1b)
        gf = fb
    # e = exp(h/2 L) (* mask): de/dh = L/2 e
    hbar += float(np.sum(ebar * (L / 2) * e))
    return gf, hbar / steps


# --------------------------------------------------------------------------
# adaptive Lawson RK4(3)  (timestepping.py:151-335)
# --------------------------------------------------------------------------
def lawson_rk43(x, pd, t, h=1e-2, atol=1e-4, max_steps_per_checkpoint=1024,
                checkpoints=1, record=False, _plant_k1_restart=False):
    """Adaptive Lawson RK4 with embedded 3rd-order error (timestepping.py:151-335).

    Reproduces: pseudo-time s in [0,1] with mod_L = t L and v' = t v
    (:242-243); e = exp(mod_L h/2) unmasked (:263); the FSAL k5 (:282);
    err = max|irfft2(h (k5 - k4)/6)| (:288-291); hs[acc+rej] = h (:294);
    on rejection the carried k1 is the doubly multiplied e^2 k1 (:267,276,297);
    h <- h 0.9 (atol/err)^(1/4), or h/2 on NaN (:300-302); clip to s = 1
    (:304-307); done iff |s-1| < 1e-10 (:313); attempt budget
    checkpoints * max_steps_per_checkpoint (:315-321).

    The accept test ``err < atol`` is bit-sensitive: a last-ulp difference in
    err at the boundary flips the decision, so only invariants of the
    decision sequence are pinned.  With ``record`` the per-attempt
    (h, accepted) schedule is returned in info['schedule'].

    ``_plant_k1_restart`` is a planted mistake for the tests only: a rejected
    attempt then restarts from k1 = v'(x0) instead of carrying e^2 k1, so a
    test can show that it tells the two apart.
    """
    L = dissipation(pd)
    mod_L = t * L
    v = lambda z: t * state_vel(z, pd, False)
    budget = int(max_steps_per_checkpoint) * int(checkpoints)
    hs = np.zeros(budget)
    s, k1, fevals, acc, rej = 0.0, v(x), 1, 0, 0
    schedule = []
    for _ in range(budget):
        if abs(s - 1.0) < 1e-10:
            continue
        x0 = x
        e = np.exp(mod_L * h / 2)
        xx = e * x0
        k1 = e * k1
        k2 = v(xx + h * k1 / 2)
        k3 = v(xx + h * k2 / 2)
        xx = e * xx
        k1 = e * k1
        k2 = e * k2
        k3 = e * k3
        k4 = v(xx + h * k3)
        xf = xx + h * (k1 / 6 + k2 / 3 + k3 / 3 + k4 / 6)
        k5 = v(xf)
        fevals += 4
        err = np.max(np.abs(_irfft2(h * (k5 - k4) / 6)))
        hs[acc + rej] = h
        accepted = bool(err < atol)
        schedule.append((h, accepted))
        if accepted:
            x, k1, s, acc = xf, k5, s + h, acc + 1
        else:
            x, rej = x0, rej + 1
            if _plant_k1_restart:
                k1 = v(x0)
        with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
            h = h / 2 if np.isnan(err) else h * 0.9 * (atol / err) ** (1 / 4)
            over = s + h - 1
            h = h - (over if (np.isnan(over) or over > 0) else 0.0)
    info = {"completed": abs(s - 1.0) < 1e-10, "s": s, "accepted": acc,
            "rejected": rej, "fevals": fevals, "hs": hs}
    if record:
        info["schedule"] = schedule
    return x, info


def lawson_rk43_jvp(x, dx, pd, t, dt_t, schedule):
    """Forward-mode derivative of lawson_rk43 along a recorded schedule.

    h is stop-gradient (timestepping.py:308) and the accept/reject decisions
    follow the primal, so the tangent replays ``schedule`` [(h, accepted)].
    t enters through mod_L = t L and v' = t v (:242-243).  A rejected attempt
    still changes the carried k1 to e^2 k1 (:267,276,297), and so its tangent.
    """
    L = dissipation(pd)
    v = lambda z: state_vel(z, pd, False)
    jv = lambda z, dz: state_vel_jvp(z, dz, pd, False)
    v0 = v(x)
    k1, dk1 = t * v0, dt_t * v0 + t * jv(x, dx)
    for h, accepted in schedule:
        e = np.exp(t * L * h / 2)
        de = L * h / 2 * dt_t * e
        xx, dxx = e * x, de * x + e * dx
        k1n, dk1n = e * k1, de * k1 + e * dk1
        y2, dy2 = xx + h * k1n / 2, dxx + h * dk1n / 2
        v2 = v(y2); k2, dk2 = t * v2, dt_t * v2 + t * jv(y2, dy2)
        y3, dy3 = xx + h * k2 / 2, dxx + h * dk2 / 2
        v3 = v(y3); k3, dk3 = t * v3, dt_t * v3 + t * jv(y3, dy3)
        xx, dxx = e * xx, de * xx + e * dxx
        k1n, dk1n = e * k1n, de * k1n + e * dk1n
        k2, dk2 = e * k2, de * k2 + e * dk2
        k3, dk3 = e * k3, de * k3 + e * dk3
        y4, dy4 = xx + h * k3, dxx + h * dk3
        v4 = v(y4); k4, dk4 = t * v4, dt_t * v4 + t * jv(y4, dy4)
        xf = xx + h * (k1n / 6 + k2 / 3 + k3 / 3 + k4 / 6)
        dxf = dxx + h * (dk1n / 6 + dk2 / 3 + dk3 / 3 + dk4 / 6)
        if accepted:
            v5 = v(xf)
            x, dx = xf, dxf
            k1, dk1 = t * v5, dt_t * v5 + t * jv(xf, dxf)
        else:
            k1, dk1 = k1n, dk1n
    return x, dx


# --------------------------------------------------------------------------
# the other fixed-step integrators of mismatch_RPO  (timestepping.py:17-149)
# --------------------------------------------------------------------------
def state_vel_hess(a, b, pd):
    """Second derivative of state_vel: d/de state_vel_jvp(f + e a, b) at e = 0.

    The nonlinearity of mhd_jax.py:136-164 is a bilinear form plus the b0
    terms, which are linear, so this is state_vel_jvp at the point a with
    b0 = 0 and no dissipation, and it does not depend on f.
    """
    pd0 = dict(pd)
    pd0["b0"] = [0.0, 0.0]
    return state_vel_jvp(a, b, pd0, False)


def rk4(f, t, steps, pd):
    """Classical RK4 (timestepping.py:17-44) on the RHS with dissipation
    (mode "RK4", loss_functions.py:49-56)."""
    h = t / steps
    v = lambda z: state_vel(z, pd, True)
    for _ in range(int(steps)):
        k1 = h * v(f)
        k2 = h * v(f + k1 / 2)
        k3 = h * v(f + k2 / 2)
        k4 = h * v(f + k3)
        f = f + k1 / 6 + k2 / 3 + k3 / 3 + k4 / 6
    return f


def rk4_jvp(f, df, t, dt_t, steps, pd):
    """Forward-mode derivative of rk4 w.r.t. (f, t); dh = dt_t / steps."""
    h, dh = t / steps, dt_t / steps
    v = lambda z: state_vel(z, pd, True)
    jv = lambda z, dz: state_vel_jvp(z, dz, pd, True)

    def stage(z, dz):
        vz = v(z)
        return h * vz, dh * vz + h * jv(z, dz)

    for _ in range(int(steps)):
        k1, dk1 = stage(f, df)
        k2, dk2 = stage(f + k1 / 2, df + dk1 / 2)
        k3, dk3 = stage(f + k2 / 2, df + dk2 / 2)
        k4, dk4 = stage(f + k3, df + dk3)
        f = f + k1 / 6 + k2 / 3 + k3 / 3 + k4 / 6
        df = df + dk1 / 6 + dk2 / 3 + dk3 / 3 + dk4 / 6
    return f, df


def dense_vjp(jvp, shape, cbar):
    """Transpose of a real-linear tangent map by brute force on tiny inputs:
    jvp(df, dt) -> dout for spectra of `shape`.  Builds the Jacobian column by
    column (real and imaginary unit perturbations of every stored entry, then
    the scalar t) and returns (J^T cbar restricted to the state, and to t)
    in the real inner product sum Re(conj(a) b) on stored entries."""
    m = int(np.prod(shape))
    cols = []
    zero = np.zeros(shape, dtype=complex)
    for k in range(m):
        for unit in (1.0, 1j):
            d = zero.copy().reshape(-1)
            d[k] = unit
            cols.append(jvp(d.reshape(shape), 0.0))
    tcol = jvp(zero, 1.0)
    fbar = np.zeros(m, dtype=complex)
    cr = cbar.reshape(-1)
    for k in range(m):
        fbar[k] = rdot(cr, cols[2 * k].reshape(-1)) + 1j * rdot(cr, cols[2 * k + 1].reshape(-1))
    return fbar.reshape(shape), rdot(cr, tcol.reshape(-1))


# Lawson RK6 of timestepping.py:83-125, written as data: stage s (0-based)
# evaluates h v(f + sum_j RK6_A[s][j] k_j) after the integrating factor
# e = exp(h/6 L) [* mask] has been applied to f and all k_j whenever
# RK6_E_BEFORE[s] is set; the step returns f + sum_j RK6_B[j] k_j.
RK6_A = [
    [],
    [1 / 6],
    [1 / 12, 1 / 12],
    [0, -4 / 33, 5 / 11],
    [-1 / 4, -29 / 44, 31 / 22, 0],
    [3 / 11, 8 / 33, -4 / 11, 1 / 11, 14 / 33],
    [-17 / 48, -5 / 12, 1, 1, -13 / 12, 11 / 16],
    [20 / 39, 12 / 39, -31 / 39, -1 / 39, 34 / 39, -11 / 39, 16 / 39],
]
RK6_B = [13 / 200, 0, 4 / 25, 11 / 40, 0, 11 / 40, 4 / 25, 13 / 200]
RK6_E_BEFORE = [False, True, False, True, True, True, True, True]


def lawson_rk6(f, t, steps, pd, mask=True):
    """Lawson RK6 (timestepping.py:83-125), e = exp(h/6 L) [* mask], RHS without
    dissipation (mode "Lawson_RK6", loss_functions.py:66-71)."""
    h = t / steps
    L = dissipation(pd)
    e = np.exp(h / 6 * L)
    if mask:
        e = e * pd["mask"]
    v = lambda z: state_vel(z, pd, False)
    for _ in range(int(steps)):
        ks = []
        for s in range(8):
            if RK6_E_BEFORE[s]:
                f = e * f
                ks = [e * k for k in ks]
            y = f
            for a, k in zip(RK6_A[s], ks):
                if a != 0:
                    y = y + a * k
            ks.append(h * v(y))
        f = f + sum(bj * k for bj, k in zip(RK6_B, ks) if bj != 0)
    return f


def lawson_rk6_jvp(f, df, t, dt_t, steps, pd, mask=True):
    """Forward-mode derivative of lawson_rk6 w.r.t. (f, t): dh = dt_t / steps
    enters every k = h v and e (de = dh/6 L e)."""
    h, dh = t / steps, dt_t / steps
    L = dissipation(pd)
    e = np.exp(h / 6 * L)
    if mask:
        e = e * pd["mask"]
    de = dh / 6 * L * e
    v = lambda z: state_vel(z, pd, False)
    jv = lambda z, dz: state_vel_jvp(z, dz, pd, False)
    for _ in range(int(steps)):
        ks, dks = [], []
        for s in range(8):
            if RK6_E_BEFORE[s]:
                f, df = e * f, de * f + e * df
                dks = [de * k + e * dk for k, dk in zip(ks, dks)]
                ks = [e * k for k in ks]
            y, dy = f, df
            for a, k, dk in zip(RK6_A[s], ks, dks):
                if a != 0:
                    y, dy = y + a * k, dy + a * dk
            vy = v(y)
            ks.append(h * vy)
            dks.append(dh * vy + h * jv(y, dy))
        f = f + sum(bj * k for bj, k in zip(RK6_B, ks) if bj != 0)
        df = df + sum(bj * dk for bj, dk in zip(RK6_B, dks) if bj != 0)
    return f, df


def tdrk4(f, t, steps, pd):
    """Two-derivative RK4 (timestepping.py:127-149): a_j = h J(y_j) k_j with J
    the RHS Jacobian (jax.jvp of v_fn, :135); RHS with dissipation (mode
    "TDRK4", loss_functions.py:78-84)."""
    h = t / steps
    v = lambda z: state_vel(z, pd, True)
    jv = lambda z, dz: state_vel_jvp(z, dz, pd, True)
    for _ in range(int(steps)):
        k1 = h * v(f)
        a1 = h * jv(f, k1)
        y = f + k1 / 2 + a1 / 8
        k2 = h * v(y)
        a2 = h * jv(y, k2)
        f = f + k1 + a1 / 6 + a2 / 3
    return f


def tdrk4_jvp(f, df, t, dt_t, steps, pd):
    """Forward-mode derivative of tdrk4 w.r.t. (f, t).  The tangent of
    a = h J(y) k is dh J(y) k + h (H[dy, k] + J(y) dk), H = state_vel_hess
    (the second derivative jax.jvp of the nested jvp produces)."""
    h, dh = t / steps, dt_t / steps
    v = lambda z: state_vel(z, pd, True)
    jv = lambda z, dz: state_vel_jvp(z, dz, pd, True)
    hv = lambda a, b: state_vel_hess(a, b, pd)
    for _ in range(int(steps)):
        v1 = v(f)
        k1, dk1 = h * v1, dh * v1 + h * jv(f, df)
        j1 = jv(f, k1)
        a1, da1 = h * j1, dh * j1 + h * (hv(df, k1) + jv(f, dk1))
        y, dy = f + k1 / 2 + a1 / 8, df + dk1 / 2 + da1 / 8
        v2 = v(y)
        k2, dk2 = h * v2, dh * v2 + h * jv(y, dy)
        j2 = jv(y, k2)
        a2, da2 = h * j2, dh * j2 + h * (hv(dy, k2) + jv(y, dk2))
        f = f + k1 + a1 / 6 + a2 / 3
        df = df + dk1 + da1 / 6 + da2 / 3
    return f, df


# --------------------------------------------------------------------------
# symmetries  (symmetry.py:4-25) with tangent and transpose
# --------------------------------------------------------------------------
def shift_x(f, s, pd):
    """exp(i kx s) f with masked kx (symmetry.py:4-5)."""
    return np.exp(1j * pd["kx"] * s) * f


def shift_reflect(f, n, pd):
    """exp(i ky n pi/4) f; for odd n also kx -> -kx and a sign flip (symmetry.py:7-19)."""
    n = int(n)
    s = (n * 2 * np.pi) / 8
    f = np.exp(1j * pd["ky"] * s) * f
    if n % 2 == 1:
        f = np.flip(f, axis=-2)
        f = np.roll(f, shift=1, axis=-2)
        f = -f
    return f


def shift_reflect_T(g, n, pd):
    """Real transpose of shift_reflect (a unitary map, so its inverse adjoint)."""
    n = int(n)
    s = (n * 2 * np.pi) / 8
    if n % 2 == 1:
        g = -g
        g = np.roll(g, shift=-1, axis=-2)
        g = np.flip(g, axis=-2)
    return np.conj(np.exp(1j * pd["ky"] * s)) * g


def rot180(f):
    """conj(f) * [1, -1] (symmetry.py:21-25)."""
    return np.array([1.0, -1.0]).reshape(2, 1, 1) * np.conj(f)


def rot180_T(g):
    """Real transpose of rot180: sign * conj(g)."""
    return np.array([1.0, -1.0]).reshape(2, 1, 1) * np.conj(g)


def apply_symmetry(f, sx, pd):
    """The symmetry tail of mismatch_RPO (loss_functions.py:91-94)."""
    f = shift_x(f, sx, pd)
    f = shift_reflect(f, pd["shift_reflect_ny"], pd)
    if pd["rot"]:
        f = rot180(f)
    return f


def apply_symmetry_jvp(f, df, sx, dsx, pd):
    """Tangent of apply_symmetry in (f, sx)."""
    a = np.exp(1j * pd["kx"] * sx)
    g, dg = a * f, a * df + 1j * pd["kx"] * dsx * a * f
    g, dg = shift_reflect(g, pd["shift_reflect_ny"], pd), shift_reflect(dg, pd["shift_reflect_ny"], pd)
    if pd["rot"]:
        g, dg = rot180(g), rot180(dg)
    return g, dg


def apply_symmetry_vjp(f, sx, pd, gbar):
    """Transpose of apply_symmetry_jvp: returns (fbar, sxbar)."""
    if pd["rot"]:
        gbar = rot180_T(gbar)
    gbar = shift_reflect_T(gbar, pd["shift_reflect_ny"], pd)
    a = np.exp(1j * pd["kx"] * sx)
    fbar = np.conj(a) * gbar
    sxbar = rdot(gbar, 1j * pd["kx"] * a * f)
    return fbar, sxbar


# --------------------------------------------------------------------------
# guess generation  (generate_turbulence.py:86-167, utils.py:60-86, symmetry.py:28-56)
# --------------------------------------------------------------------------
def turbulence_series(f, pd, dt, ministeps, steps, transient_steps):
    """Transient of Lawson-RK6 steps of dt, then `steps` snapshots, each after
    ministeps steps of h = (dt ministeps) / ministeps (generate_turbulence.py:86-125)."""
    for _ in range(int(transient_steps)):
        f = lawson_rk6(f, dt, 1, pd, True)
    fs = []
    for _ in range(int(steps)):
        f = lawson_rk6(f, dt * ministeps, ministeps, pd, True)
        fs.append(f)
    return np.stack(fs)


def recurrence_diagram(fs):
    """dist[:, i] = norm(|fs| - |fs_i|) over each snapshot (generate_turbulence.py:140-151)."""
    S = fs.shape[0]
    dist = np.zeros((S, S))
    a = np.abs(fs).reshape(S, -1)
    for i in range(S):
        dist[:, i] = np.linalg.norm(a - a[i], axis=1)
    return dist


def find_optimal_symmetry(f1, f2, pd, nx):
    """Brute-force search of symmetry.py:28-56 (first minimum kept)."""
    best_g, best_err = (0.0, 0, False), 1e100
    for i in range(nx):
        s = (2 * np.pi / nx) * i
        for ny in range(8):
            for rot in (True, False):
                f = shift_x(f2, s, pd)
                f = shift_reflect(f, ny, pd)
                if rot:
                    f = rot180(f)
                err = np.mean(np.square(np.abs(f1 - f)))
                if err < best_err:
                    best_err, best_g = err, (s, ny, rot)
    return best_g


def create_state_from_turb(fs, idx, pd, dt, ministeps, nx=64):
    """utils.py:60-86: (input_dict, steps, shift_reflect_ny, rot)."""
    f1, f2 = fs[idx[0] - 1], fs[idx[1] - 1]
    g = find_optimal_symmetry(f1, f2, pd, nx)
    T = dt * ministeps * (idx[1] - idx[0])
    return {"fields": _irfft2(f1), "T": T, "sx": g[0]}, int(ministeps * (idx[1] - idx[0])), g[1], g[2]


def floquet_block_power(inp, pd, tang, maxit, mode="TDRK4"):
    """Block power iteration of floquet.py:42-99 with the oracle's tangents:
    J t = irfft2(mask * g dPhi_T rfft2(t)) at fixed T and sx; numpy (LAPACK
    Householder) QR after each application; R = Q J(Q)^T."""
    f0 = _rfft2(inp["fields"])
    T, sx = float(inp["T"]), float(inp["sx"])

    def jac(t):
        out = []
        for b in range(t.shape[0]):
            df0 = _rfft2(t[b])
            if mode in ("RK4", "TDRK4", "Lawson_RK4"):
                nc, ms = int(pd["num_checkpoints"]), int(pd["ministeps"])
                step = {"RK4": lambda f, d, a, c: rk4_jvp(f, d, a, c, ms, pd),
                        "TDRK4": lambda f, d, a, c: tdrk4_jvp(f, d, a, c, ms, pd),
                        "Lawson_RK4": lambda f, d, a, c: lawson_rk4_jvp(f, d, a, c, ms, pd, True)}[mode]
                f, df = f0, df0
                for _ in range(nc):
                    f, df = step(f, df, T / nc, 0.0)
            else:
                raise ValueError(mode)
            _, dg = apply_symmetry_jvp(f, df, sx, 0.0, pd)
            out.append(_irfft2(dg * pd["mask"]))
        return np.stack(out)

    B = tang.shape[0]
    for _ in range(maxit):
        q, _ = np.linalg.qr(jac(tang).reshape(B, -1).T, mode="reduced")
        tang = q.T.reshape(tang.shape)
    t = tang.reshape(B, -1)
    jt = jac(tang).reshape(B, -1)
    return t @ jt.T, tang


# --------------------------------------------------------------------------
# RPO objective and loss  (loss_functions.py:9-148)
# --------------------------------------------------------------------------
def mismatch_RPO(inp, pd, mode):
    """rfft2 of fields (unmasked, :45), integrate by mode (:48-88), apply symmetries (:91-94)."""
    f0 = _rfft2(inp["fields"])
    T = float(inp["T"])
    if mode == "Lawson_RK4":
        nc, ms = int(pd["num_checkpoints"]), int(pd["ministeps"])
        f = f0
        for _ in range(nc):
            f = lawson_rk4(f, T / nc, ms, pd, True)
        info = None
    elif mode == "Lawson_RK43":
        ad = pd["adaptive_dict"]
        f, info = lawson_rk43(f0, pd, T, h=1e-2, atol=ad["atol"],
                              max_steps_per_checkpoint=ad["max_steps_per_checkpoint"],
                              checkpoints=ad["checkpoints"], record=True)
    elif mode in ("RK4", "TDRK4"):
        nc, ms = int(pd["num_checkpoints"]), int(pd["ministeps"])
        step = rk4 if mode == "RK4" else tdrk4
        f = f0
        for _ in range(nc):
            f = step(f, T / nc, ms, pd)
        info = None
    elif mode == "Lawson_RK6":
        f = lawson_rk6(f0, T, int(pd["steps"]), pd, True)
        info = None
    else:
        raise ValueError(f"You selected mode = {mode}, which does not exist.")
    f = apply_symmetry(f, float(inp["sx"]), pd)
    return f0, f, info


def loss_RPO(inp, pd, mode):
    """mean(irfft2(f - f0)^2) (loss_functions.py:99-111)."""
    f0, f, info = mismatch_RPO(inp, pd, mode)
    diff = _irfft2(f - f0)
    return float(np.mean(np.square(diff))), info


def phase_vectors(f0, pd):
    """Stop-gradient phase-condition vectors vt, vx in real space (loss_functions.py:113-122)."""
    vt = _irfft2(state_vel(f0, pd, True))
    vx = _irfft2(1j * pd["kx"] * f0)
    return vt, vx


def objective_RPO(inp, pd, mode):
    """{fields: irfft2(f0 - f), T: pt, sx: px} (loss_functions.py:132-148).

    pt and px are x - stop_gradient(x), i.e. exactly 0 in value (:127-128).
    """
    f0, f, info = mismatch_RPO(inp, pd, mode)
    return {"fields": _irfft2(f0 - f), "T": 0.0, "sx": 0.0}, info


def objective_RPO_jvp(inp, tan, pd, mode, schedule=None):
    """JVP of objective_RPO at inp along tan (utils.py:108).

    The phase conditions differentiate to mean(vt * dfields), mean(vx * dfields)
    (loss_functions.py:115-128; vt, vx are stop-gradient).
    """
    f0 = _rfft2(inp["fields"])
    df0 = _rfft2(tan["fields"])
    T, dT = float(inp["T"]), float(tan["T"])
    if mode == "Lawson_RK4":
        nc, ms = int(pd["num_checkpoints"]), int(pd["ministeps"])
        f, df = f0, df0
        for _ in range(nc):
            f, df = lawson_rk4_jvp(f, df, T / nc, dT / nc, ms, pd, True)
    elif mode == "Lawson_RK43":
        if schedule is None:
            _, _, info = mismatch_RPO(inp, pd, mode)
            schedule = info["schedule"]
        f, df = lawson_rk43_jvp(f0, df0, pd, T, dT, schedule)
    elif mode in ("RK4", "TDRK4"):
        nc, ms = int(pd["num_checkpoints"]), int(pd["ministeps"])
        step = rk4_jvp if mode == "RK4" else tdrk4_jvp
        f, df = f0, df0
        for _ in range(nc):
            f, df = step(f, df, T / nc, dT / nc, ms, pd)
    elif mode == "Lawson_RK6":
        f, df = lawson_rk6_jvp(f0, df0, T, dT, int(pd["steps"]), pd, True)
    else:
        raise ValueError(mode)
    _, dg = apply_symmetry_jvp(f, df, float(inp["sx"]), float(tan["sx"]), pd)
    vt, vx = phase_vectors(f0, pd)
    return {"fields": _irfft2(df0 - dg),
            "T": float(np.mean(vt * tan["fields"])),
            "sx": float(np.mean(vx * tan["fields"]))}


def loss_RPO_value_and_grad(inp, pd):
    """value_and_grad of loss_RPO in mode Lawson_RK4 (adjoint_descent.py:90-96).

    Returns (loss, {fields, T, sx} gradient).
    """
    n = inp["fields"].shape[-1]
    nc, ms = int(pd["num_checkpoints"]), int(pd["ministeps"])
    T, sx = float(inp["T"]), float(inp["sx"])
    f0 = _rfft2(inp["fields"])
    segs = [f0]
    for _ in range(nc):
        segs.append(lawson_rk4(segs[-1], T / nc, ms, pd, True))
    fT = segs[-1]
    g = apply_symmetry(fT, sx, pd)
    diff = _irfft2(g - f0)
    loss = float(np.mean(np.square(diff)))
    dbar = irfft2_T(diff / (n * n))        # d loss / d (g - f0)
    fTbar, sxbar = apply_symmetry_vjp(fT, sx, pd, dbar)
    f0bar = -dbar
    Tbar = 0.0
    fb = fTbar
    for i in reversed(range(nc)):
        fb, tb = lawson_rk4_vjp(segs[i], T / nc, ms, pd, fb, True)
        Tbar += tb / nc
    f0bar = f0bar + fb
    return loss, {"fields": rfft2_T(f0bar), "T": Tbar, "sx": sxbar}


def objective_RPO_vjp(inp, cot, pd):
    """VJP of objective_RPO in mode Lawson_RK4 (newton.py:168-169): J^T cot."""
    n = inp["fields"].shape[-1]
    nc, ms = int(pd["num_checkpoints"]), int(pd["ministeps"])
    T, sx = float(inp["T"]), float(inp["sx"])
    f0 = _rfft2(inp["fields"])
    segs = [f0]
    for _ in range(nc):
        segs.append(lawson_rk4(segs[-1], T / nc, ms, pd, True))
    fT = segs[-1]
    cbar = irfft2_T(cot["fields"])            # cotangent of (f0 - g)
    fTbar, sxbar = apply_symmetry_vjp(fT, sx, pd, -cbar)
    f0bar = cbar
    Tbar = 0.0
    fb = fTbar
    for i in reversed(range(nc)):
        fb, tb = lawson_rk4_vjp(segs[i], T / nc, ms, pd, fb, True)
        Tbar += tb / nc
    f0bar = f0bar + fb
    vt, vx = phase_vectors(f0, pd)
    n2 = 2 * n * n
    gfields = rfft2_T(f0bar) + (cot["T"] * vt + cot["sx"] * vx) / n2
    return {"fields": gfields, "T": Tbar, "sx": sxbar}


# --------------------------------------------------------------------------
# flattening  (newton.py:106-110: ravel_pytree sorts keys -> [T, fields, sx])
# --------------------------------------------------------------------------
def flatten(d):
    return np.concatenate([np.atleast_1d(np.float64(d["T"])),
                           np.asarray(d["fields"], dtype=np.float64).ravel(),
                           np.atleast_1d(np.float64(d["sx"]))])


def unflatten(x, n):
    return {"T": float(x[0]), "fields": x[1:-1].reshape(2, n, n), "sx": float(x[-1])}


# --------------------------------------------------------------------------
# GMRES, line search, ADAM  (linalg.py:12-98, utils.py:17-58, adam.py:23-61)
# --------------------------------------------------------------------------
def gmres(A, b, inner, outer=1, preconditioner_list=()):
    """Single-pass MGS Arnoldi + least squares (linalg.py:12-98).

    No restarts (``outer`` is ignored), no breakdown guard, one extra matvec
    for the relative residual (:97).  Returns (x, rel_res, H, Q).
    """
    def precond(v):
        for M in preconditioner_list:
            v = M(v)
        return v
    n = b.size
    Q = np.zeros((n, inner + 1))
    H = np.zeros((inner + 1, inner))
    b = precond(b)
    Q[:, 0] = b / np.linalg.norm(b)
    for k in range(inner):
        w = precond(A(Q[:, k]))
        for j in range(k + 1):
            hj = np.dot(w, Q[:, j])
            w = w - hj * Q[:, j]
            H[j, k] = hj
        hk1 = np.linalg.norm(w)
        H[k + 1, k] = hk1
        Q[:, k + 1] = w / hk1
    b2 = Q.T @ b
    y = np.linalg.lstsq(H, b2, rcond=None)[0]
    x = Q[:, :inner] @ y
    rel_res = np.linalg.norm(precond(A(x)) - b) / np.linalg.norm(b)
    return x, float(rel_res), H, Q


def _bisect(g, lower, upper, maxit=128):
    """Bisection of linalg.py:165-177 / :289-301: t = 0 if g(0) > 0, else the
    last midpoint after maxit halvings (g(t) > 0 moves the upper end)."""
    if g(0.0) > 0:
        return 0.0
    t = lower
    for _ in range(maxit):
        t = (lower + upper) / 2
        if g(t) > 0:
            upper = t
        else:
            lower = t
    return t


def _arnoldi_reorth(A, b, m):
    """Arnoldi with one recorded MGS pass and two discarded reorthogonalisation
    passes (linalg.py:137-153, :255-268).  Returns Q (n, m+1), H (m+1, m)."""
    n = b.size
    Q = np.zeros((n, m + 1))
    H = np.zeros((m + 1, m))
    Q[:, 0] = b / np.linalg.norm(b)
    for k in range(m):
        w = A(Q[:, k])
        for j in range(k + 1):
            hj = np.dot(w, Q[:, j])
            w = w - hj * Q[:, j]
            H[j, k] = hj
        for _ in range(2):
            for j in range(k + 1):
                w = w - np.dot(w, Q[:, j]) * Q[:, j]
        H[k + 1, k] = np.linalg.norm(w)
        Q[:, k + 1] = w / H[k + 1, k]
    return Q, H


def hookstep_v2(A, b, m, s, f_vec, x, max_halvings=64):
    """newton_gmres_hookstep_v2 (linalg.py:217-337) on flat vectors: f_vec(x)
    is the flattened objective.  Returns (x - dx, s, rel_res, dx).  The
    trust-region test compares |y| with s*s, and the bisection makes |z| = s,
    as the reference does (:282, :287)."""
    Q, H = _arnoldi_reorth(A, b, m)
    b2 = Q.T @ b
    U, S, Vh = np.linalg.svd(H, full_matrices=False)
    y = Vh.T @ ((U.T @ b2) / S)
    nb = np.linalg.norm(b)
    for _ in range(max_halvings):
        if np.linalg.norm(y) > s * s:
            b4 = Vh @ (H.T @ b2)
            g = lambda t: s - np.linalg.norm(b4 / (S * S + t))
            t = _bisect(g, 0.0, np.sqrt(m) * np.linalg.norm(b4) / s)
            y = Vh.T @ (b4 / (S * S + t))
        dx = Q[:, :m] @ y
        if np.linalg.norm(f_vec(x - dx)) < nb:
            s = s * 2
            break
        s = s / 2
    else:
        raise RuntimeError("hookstep: no step decreased |f|")
    rel_res = np.linalg.norm(A(dx) - b) / nb
    return x - dx, s, float(rel_res), dx


def hookstep_v1(AtA, Atb, m, s, f_vec, f0, J, b, x, max_halvings=64):
    """newton_gmres_hookstep (linalg.py:102-213): Arnoldi on A^T A (square H,
    no last subdiagonal), eigh of H (lower triangle), bisection for the step of
    norm s.  Returns (x - step, s, rel_res_1, rel_res_2, step)."""
    Q, Hf = _arnoldi_reorth(AtA, Atb, m)
    H = Hf[:m, :m]
    D, V = np.linalg.eigh(H)
    b2 = V.T @ (Q[:, :m].T @ Atb)
    for _ in range(max_halvings):
        g = lambda t: s - np.linalg.norm(b2 / (D + t))
        t = 0.0 if g(0.0) > 0 else _bisect(g, 0.0, np.linalg.norm(b) / s)
        step = Q[:, :m] @ (V @ (b2 / (D + t)))
        if np.linalg.norm(f_vec(x - step)) < np.linalg.norm(f0):
            s = s * 2
            break
        s = s / 2
    else:
        raise RuntimeError("hookstep: no step decreased |f|")
    rr1 = np.linalg.norm(AtA(step) - Atb) / np.linalg.norm(Atb)
    rr2 = np.linalg.norm(J(step) + b) / np.linalg.norm(b)
    return x - step, s, float(rr1), float(rr2), step


def line_search(x, step, obj_norm, f_norm, max_iters=20):
    """Halving line search (utils.py:17-58): first damp in 1, 1/2, ... with
    |f(x - damp step)| < |f(x)|; returns (x_new, damp)."""
    damp = 1.0
    for _ in range(max_iters):
        xt = x - damp * step
        if obj_norm(xt) < f_norm:
            return xt, damp
        damp = damp / 2.0
    return x, damp


def adam_update(params, grads, m, v, t, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
    """ADAM over a dict of arrays (adam.py:23-61)."""
    m = {k: beta1 * m[k] + (1 - beta1) * grads[k] for k in params}
    v = {k: beta2 * v[k] + (1 - beta2) * (grads[k] ** 2) for k in params}
    mh = {k: m[k] / (1 - beta1 ** t) for k in params}
    vh = {k: v[k] / (1 - beta2 ** t) for k in params}
    params = {k: params[k] - lr * mh[k] / (np.sqrt(vh[k]) + eps) for k in params}
    return params, m, v


def dealias_fields(fields, pd):
    """irfft2(mask * rfft2(fields)) (newton.py:211-214, adjoint_descent.py:111-114)."""
    return _irfft2(pd["mask"] * _rfft2(fields))


# --------------------------------------------------------------------------
# one Newton-GMRES iteration and one adjoint-descent iteration
# (newton.py:151-215, adjoint_descent.py:93-123)
# --------------------------------------------------------------------------
def newton_iteration(inp, pd, mode, inner, do_line_search=True, default_damp=0.1):
    """newton_gmres_update of newton.py:151-215 without preconditioners.

    b = flatten(objective(x)) (:159-162); A = flatten o jac(x, .) o unflatten
    (:165); (step, rel_res) = gmres(A, b, inner) (:181-182); halving line
    search on |objective(x - damp step)| < |b| (utils.py:17-58) or the fixed
    default_damp; rel_err = |f.fields| / |x_new.fields| before dealiasing
    (:195, relative_error_RPO at :140-142); the fields are dealiased last
    (:211-214).  Returns (x_new, log) with the logged values (:197-208) plus
    H, Q and the GMRES step.
    """
    n = inp["fields"].shape[-1]
    f, info = objective_RPO(inp, pd, mode)
    b = flatten(f)
    sched = info["schedule"] if info is not None else None
    A = lambda v: flatten(objective_RPO_jvp(inp, unflatten(v, n), pd, mode, sched))
    step, rel_res, H, Q = gmres(A, b, inner)
    x = flatten(inp)
    if do_line_search:
        obj_norm = lambda xv: float(np.linalg.norm(flatten(objective_RPO(unflatten(xv, n), pd, mode)[0])))
        x, damp = line_search(x, step, obj_norm, float(np.linalg.norm(b)))
    else:
        damp = default_damp
        x = x - damp * step
    new = unflatten(x, n)
    rel_err = float(np.linalg.norm(f["fields"]) / np.linalg.norm(new["fields"]))
    new = {"fields": dealias_fields(new["fields"], pd), "T": new["T"], "sx": new["sx"]}
    log = {"rel_err": rel_err, "f_norm": float(np.linalg.norm(b)), "gmres_rel_res": rel_res, "damp": damp,
           "T": new["T"], "sx": new["sx"], "H": H, "Q": Q, "step": step}
    return new, log


def adjoint_descent_iteration(params, m, v, t, pd, lr=1e-2):
    """One iteration of adjoint_descent.py:93-123 in mode Lawson_RK4: loss and
    gradient (:96), adam_update at step t+1 with betas (0.9, 0.999) and eps
    1e-6 (:108), then the fields dealiased (:111-114).  ``t`` counts from 0.
    Returns (loss, params, m, v)."""
    loss, grad = loss_RPO_value_and_grad(params, pd)
    params, m, v = adam_update(params, grad, m, v, t + 1, lr=lr, beta1=0.9, beta2=0.999, eps=1e-6)
    params = dict(params, fields=dealias_fields(params["fields"], pd))
    return loss, params, m, v
