"""Plain Adam, as used by the reference's golden training trajectories.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

The reference drives ``optax.adam(lr)`` (tests/test_training.py:15-30);
optax is absent here, so its published update rule is restated with optax's
defaults (b1=0.9, b2=0.999, eps=1e-8, eps_root=0):

    m <- b1 m + (1-b1) g ;  v <- b2 v + (1-b2) g^2
    m_hat = m / (1-b1^t) ;  v_hat = v / (1-b2^t)
    p <- p - lr * m_hat / (sqrt(v_hat) + eps)

Pinned by the golden losses of tests/test_training.py:54-106 (rtol 1e-12).
"""

from __future__ import annotations

import numpy as np


def adam_trajectory(value_and_grad, params, lr, steps, b1=0.9, b2=0.999, eps=1e-8):
    """Losses after every step plus the final loss, and the final parameters."""
    p = np.array(params, dtype=np.float64)
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    losses = []
    for t in range(1, steps + 1):
        loss, g = value_and_grad(p)
        m = b1 * m + (1 - b1) * g
        v = b2 * v + (1 - b2) * g * g
        mh = m / (1 - b1**t)
        vh = v / (1 - b2**t)
        p = p - lr * mh / (np.sqrt(vh) + eps)
        losses.append(float(loss))
    losses.append(float(value_and_grad(p)[0]))
    return np.array(losses), p
