"""NumPy/SciPy float64 restatement of the pulse-propagator path.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Restates:

* the lifted Gaussian / DRAG envelopes       -- jaqsi/pulses.py:324-390
  (the cancellation-free form of pulses.py:337-352)
* the Magnus-2 midpoint step
  U <- expm(h * (-i H(t_n + h/2))) U         -- jaqsi/evolution.py:230-235
* the process infidelity 1 - |Tr(V^dag U)|^2 / d^2
                                             -- jaqsi/qoc.py:335-341

The two-qubit drive Hamiltonian (wire 0 = most significant bit):

    H(t) = w0/2 Z0 + w1/2 Z1 + J Z0 Z1
           + E(t) (cos phi X0 + sin phi Y0) + Q(t) (-sin phi X0 + cos phi Y0)

with E = A * lifted Gaussian(sigma) centred at T/2 and Q = -beta dE/dt (the
DRAG quadrature on axis phi + pi/2).

How it is pinned (tests/test_oracle_pulse.py): zero drive against the closed
form exp(-i T H0) of the diagonal Hamiltonian; the envelope against its plain
(cancelling) textbook form and against a finite-difference derivative for Q;
second-order convergence of Magnus-2 in the step count; unitarity; and the
per-block SU(2) closed form (H commutes with Z1) as an independent
restatement of every step.
"""

from __future__ import annotations

import math

import numpy as np
from scipy.linalg import expm

from . import sv

C = np.complex128
I2 = np.eye(2, dtype=C)


def lifted_gaussian(t, sigma, tc):
    """(g(t) - g(0)) / (1 - g(0)), g(t) = exp(-(t - tc)^2 / (2 sigma^2)),
    in the cancellation-free form of pulses.py:337-352."""
    a = -0.5 * ((t - tc) / sigma) ** 2
    b = -0.5 * (tc / sigma) ** 2
    d = t * (2 * tc - t) / (2 * sigma**2)
    rise = np.expm1(np.minimum(d, 0.0)) - np.expm1(np.minimum(-d, 0.0))
    return np.exp(np.maximum(a, b)) * rise / -np.expm1(b)


def envelopes(t, A, sigma, beta, tc):
    """(E, Q): in-phase lifted Gaussian and DRAG quadrature -beta dE/dt."""
    E = A * lifted_gaussian(t, sigma, tc)
    g = np.exp(-0.5 * ((t - tc) / sigma) ** 2)
    lift = -np.expm1(-0.5 * (tc / sigma) ** 2)
    Q = A * beta * (t - tc) / sigma**2 * g / lift
    return E, Q


def hamiltonian(t, params, omega, J, T):
    A, sigma, beta, phi = params
    E, Q = envelopes(t, A, sigma, beta, T / 2)
    a = E * math.cos(phi) - Q * math.sin(phi)
    b = E * math.sin(phi) + Q * math.cos(phi)
    Z0 = np.kron(sv.PZ, I2)
    Z1 = np.kron(I2, sv.PZ)
    H = 0.5 * omega[0] * Z0 + 0.5 * omega[1] * Z1 + J * (Z0 @ Z1)
    return H + np.kron(a * sv.PX + b * sv.PY, I2)


def propagate(params, omega, J, T, N):
    """Magnus-2: U <- expm(-i h H(t_n + h/2)) U for n = 0 .. N-1."""
    h = T / N
    U = np.eye(4, dtype=C)
    for n in range(N):
        U = expm(-1j * h * hamiltonian(n * h + 0.5 * h, params, omega, J, T)) @ U
    return U


def target_rx_half_pi() -> np.ndarray:
    return np.kron(sv.op_matrix(sv.Op("RX", (0,), (math.pi / 2,))), I2)


def infidelity(U, V) -> float:
    d = U.shape[0]
    tr = np.trace(V.conj().T @ U)
    return float(1.0 - abs(tr) ** 2 / d**2)


def workload(n_sets: int, seed: int = 1005):
    """The batch of parameter sets: rows (A, sigma, beta, phi) and (w0, w1)."""
    rng = np.random.default_rng(seed)
    A = rng.uniform(0.1, 1.0, n_sets)
    sigma = rng.uniform(5.0, 20.0, n_sets)
    beta = rng.uniform(-1.0, 1.0, n_sets)
    phi = rng.uniform(0.0, 2 * math.pi, n_sets)
    omega = rng.uniform(4.5, 5.5, (n_sets, 2)) * 2 * math.pi
    return np.stack([A, sigma, beta, phi], axis=1), omega


J_C5 = 2 * math.pi * 0.005
T_C5 = 50.0
N_C5 = 2000
