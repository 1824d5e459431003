"""ADAM over dicts of tensors (jax_scripts/lib/adam.py:7-61)."""
from __future__ import annotations

import torch


def init_adam(params):
    """Zero first and second moments shaped like params (adam.py:7-21)."""
    if isinstance(params, dict):
        return ({k: torch.zeros_like(torch.as_tensor(v)) for k, v in params.items()},
                {k: torch.zeros_like(torch.as_tensor(v)) for k, v in params.items()})
    return torch.zeros_like(params), torch.zeros_like(params)


def adam_update(params, grads, m, v, t, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
    """One ADAM step with bias correction (adam.py:23-61)."""
    if isinstance(params, dict):
        m = {k: beta1 * m[k] + (1 - beta1) * grads[k] for k in params}
        v = {k: beta2 * v[k] + (1 - beta2) * (grads[k] ** 2) for k in params}
        m_hat = {k: m[k] / (1 - beta1 ** t) for k in params}
        v_hat = {k: v[k] / (1 - beta2 ** t) for k in params}
        params = {k: params[k] - lr * m_hat[k] / (torch.sqrt(v_hat[k]) + eps) for k in params}
        return params, m, v
    m = beta1 * m + (1 - beta1) * grads
    v = beta2 * v + (1 - beta2) * (grads ** 2)
    m_hat = m / (1 - beta1 ** t)
    v_hat = v / (1 - beta2 ** t)
    return params - lr * m_hat / (torch.sqrt(v_hat) + eps), m, v
