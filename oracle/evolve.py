"""NumPy/SciPy float64 restatement of the reference's pulse evolution.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Restates:

* envelopes (lifted Gaussian, square, cosine, DRAG and its quadrature,
  sech, legacy DRAG)                     -- jaqsi/pulses.py:324-420
* the interaction-picture coefficient functions under the RWA and in the
  drive / lab frames                    -- jaqsi/pulses.py:549-764
* the Magnus-2 and CFM4 fixed-step solvers
  (U <- expm(h Omega) U per step)        -- jaqsi/evolution.py:194-263
* the single-term closed form V exp(-i F lambda) V^dagger with the pulse
  area F integrated adaptively           -- jaqsi/evolution.py:353-404
* the pulse gates RX / RY / RZ / CZ      -- jaqsi/pulses.py:1370-1500, 1630-1660
* the exact gradient of a real loss of an ordered product of exponentials,
  through the Frechet derivative from the block-triangular exponential
  expm([[X, D], [0, X]]) = [[e^X, dexp_X(D)], [0, e^X]] (Higham, "Functions of
  Matrices", 2008, thm. 3.6).

How it is pinned (tests/test_oracle_evolve.py): Magnus-2 of a self-commuting
drive against exp(-i S H) with the midpoint-rule area S (exact); second-
and fourth-order convergence of Magnus-2 and CFM4; the closed form against
the erf formula of the reference's own test (tests/test_pulse_state.py:238-
259); the RWA factor 1/2 against the lab frame at w_c = w_q; the drive frame
against the lab frame (the reference states them equal); the DRAG
quadrature against a finite-difference derivative; and the product's
gradient against central differences.
"""

from __future__ import annotations

import math
from typing import Callable, List, Sequence

import numpy as np
from scipy.integrate import quad
from scipy.linalg import expm

C = np.complex128
X = np.array([[0, 1], [1, 0]], dtype=C)
Y = np.array([[0, -1j], [1j, 0]], dtype=C)
Z = np.array([[1, 0], [0, -1]], dtype=C)
I2 = np.eye(2, dtype=C)
H_CZ = (np.pi / 4) * (np.kron(I2, I2) - np.kron(Z, I2) - np.kron(I2, Z) + np.kron(Z, Z))
OMEGA = 10 * np.pi  # qubit and carrier frequency (pulses.py:1188-1189)


# ---------------------------------------------------------------------------
# envelopes (pulses.py:324-420), p = [envelope params..., T, w]
# ---------------------------------------------------------------------------
def lifted_gaussian(sigma, t, tc):
    a = -0.5 * ((t - tc) / sigma) ** 2
    b = -0.5 * (tc / sigma) ** 2
    d = t * (2 * tc - t) / (2 * sigma**2)
    rise = np.expm1(np.minimum(d, 0.0)) - np.expm1(np.minimum(-d, 0.0))
    return np.exp(np.maximum(a, b)) * rise / -np.expm1(b)


def envelope(name: str, p, t, tc):
    if name == "gaussian":
        return p[0] * lifted_gaussian(p[1], t, tc)
    if name == "square":
        return p[0] * (np.abs(t - tc) <= p[1] / 2)
    if name == "cosine":
        return p[0] * np.cos(np.pi * np.clip((t - tc) / p[1], -0.5, 0.5))
    if name == "drag":
        return p[0] * lifted_gaussian(p[2], t, tc)
    if name == "sech":
        return p[0] / np.cosh((t - tc) / p[1])
    if name == "drag_legacy":
        A, beta, sigma = p[0], p[1], p[2]
        c = t / 2
        g = A * np.exp(-0.5 * ((t - c) / sigma) ** 2)
        return g + beta * g * (-(t - c) / sigma**2)
    raise ValueError(name)


def quadrature(name: str, p, t, tc):
    if name != "drag":
        return np.zeros_like(np.asarray(t, dtype=float))
    A, beta, sigma = p[0], p[1], p[2]
    g = np.exp(-0.5 * ((t - tc) / sigma) ** 2)
    return A * beta * (t - tc) / sigma**2 * g / -np.expm1(-0.5 * (tc / sigma) ** 2)


def coeff_fns(name: str, rwa: bool = True, frame: str = "drive", wc: float = OMEGA, wq: float = OMEGA):
    """(RX_X, RX_Y, RY_X, RY_Y) of pulses.py:549-764."""
    E = lambda p, t: envelope(name, p, t, p[-2] / 2)  # noqa: E731
    Q = lambda p, t: quadrature(name, p, t, p[-2] / 2)  # noqa: E731
    if rwa:
        return (lambda p, t: 0.5 * E(p, t) * p[-1], lambda p, t: 0.5 * Q(p, t) * p[-1],
                lambda p, t: -0.5 * Q(p, t) * p[-1], lambda p, t: 0.5 * E(p, t) * p[-1])
    if frame == "lab":
        def drive(p, t, ph):
            return E(p, t) * np.cos(wc * t + ph) - Q(p, t) * np.sin(wc * t + ph)
        return (lambda p, t: drive(p, t, 0.0) * np.cos(wq * t) * p[-1],
                lambda p, t: -drive(p, t, 0.0) * np.sin(wq * t) * p[-1],
                lambda p, t: drive(p, t, np.pi / 2) * np.cos(wq * t) * p[-1],
                lambda p, t: -drive(p, t, np.pi / 2) * np.sin(wq * t) * p[-1])
    wd, ws = wc - wq, wc + wq
    return (
        lambda p, t: (E(p, t) * 0.5 * (np.cos(wd * t) + np.cos(ws * t)) - Q(p, t) * 0.5 * (np.sin(ws * t) + np.sin(wd * t))) * p[-1],
        lambda p, t: (-E(p, t) * 0.5 * (np.sin(ws * t) - np.sin(wd * t)) + Q(p, t) * 0.5 * (np.cos(wd * t) - np.cos(ws * t))) * p[-1],
        lambda p, t: (-E(p, t) * 0.5 * (np.sin(ws * t) + np.sin(wd * t)) - Q(p, t) * 0.5 * (np.cos(wd * t) + np.cos(ws * t))) * p[-1],
        lambda p, t: (-E(p, t) * 0.5 * (np.cos(ws * t) - np.cos(wd * t)) + Q(p, t) * 0.5 * (np.sin(ws * t) - np.sin(wd * t))) * p[-1],
    )


# ---------------------------------------------------------------------------
# solvers (evolution.py:194-263, 353-404)
# ---------------------------------------------------------------------------
def hamiltonian(fns, params, Hs, t):
    return sum(float(f(p, t)) * H for f, p, H in zip(fns, params, Hs))


def magnus2(fns, params, Hs, t0, t1, N):
    h = (t1 - t0) / N
    U = np.eye(Hs[0].shape[0], dtype=C)
    for n in range(N):
        U = expm(-1j * h * hamiltonian(fns, params, Hs, t0 + (n + 0.5) * h)) @ U
    return U


_S3 = math.sqrt(3.0)
_C1, _C2, _A1, _A2 = 0.5 - _S3 / 6, 0.5 + _S3 / 6, 0.25 + _S3 / 6, 0.25 - _S3 / 6


def cfm4(fns, params, Hs, t0, t1, N):
    """CFM4:2 (Blanes & Moan 2006): U <- exp(Omega_b) exp(Omega_a) U."""
    h = (t1 - t0) / N
    U = np.eye(Hs[0].shape[0], dtype=C)
    for n in range(N):
        H1 = hamiltonian(fns, params, Hs, t0 + (n + _C1) * h)
        H2 = hamiltonian(fns, params, Hs, t0 + (n + _C2) * h)
        U = expm(-1j * h * (_A2 * H1 + _A1 * H2)) @ expm(-1j * h * (_A1 * H1 + _A2 * H2)) @ U
    return U


def area(fn, p, t0, t1) -> float:
    val, _ = quad(lambda t: float(fn(p, t)), t0, t1, epsabs=1e-13, epsrel=1e-13, limit=500)
    return val


def closed_form(fn, p, H, t0, t1):
    """U = V exp(-i F lambda) V^dagger, F the integral of fn over [t0, t1]."""
    lam, V = np.linalg.eigh(H)
    F = area(fn, p, t0, t1)
    return (V * np.exp(-1j * F * lam)) @ V.conj().T


def frechet(Xm, D):
    """dexp_X(D) from the block-triangular exponential."""
    d = Xm.shape[0]
    big = np.zeros((2 * d, 2 * d), dtype=C)
    big[:d, :d] = Xm
    big[d:, d:] = Xm
    big[:d, d:] = D
    return expm(big)[:d, d:]


def product(coef, Hs):
    """prod_m exp(-i sum_k coef[m, k] H_k), m = 0 applied first."""
    U = np.eye(Hs[0].shape[0], dtype=C)
    for row in coef:
        U = expm(-1j * sum(c * H for c, H in zip(row, Hs))) @ U
    return U


def product_grad(coef, Hs, G):
    """dL/dcoef of a real loss with dL = Re Tr(G^dagger dU) (G = dL/dRe U + i dL/dIm U)."""
    M, K = coef.shape
    E = [expm(-1j * sum(c * H for c, H in zip(row, Hs))) for row in coef]
    d = Hs[0].shape[0]
    out = np.zeros((M, K))
    for m in range(M):
        L = np.eye(d, dtype=C)
        for q in range(m + 1, M):
            L = E[q] @ L
        R = np.eye(d, dtype=C)
        for q in range(m):
            R = E[q] @ R
        Xm = -1j * sum(c * H for c, H in zip(coef[m], Hs))
        for k in range(K):
            dU = L @ frechet(Xm, -1j * Hs[k]) @ R
            out[m, k] = np.real(np.trace(G.conj().T @ dU))
    return out


# ---------------------------------------------------------------------------
# pulse gates (pulses.py:1370-1500, 1630-1660) under the RWA
# ---------------------------------------------------------------------------
def rx_rwa(w, pp, name="gaussian"):
    """RX drive of one quadrature under the RWA: the single term
    (E / 2) w X over [0, T], solved in closed form."""
    p = list(pp) + [w]
    f = coeff_fns(name, rwa=True)[0]
    return closed_form(f, p, X, 0.0, pp[-1])


def ry_rwa(w, pp, name="gaussian"):
    p = list(pp) + [w]
    f = coeff_fns(name, rwa=True)[3]
    return closed_form(f, p, Y, 0.0, pp[-1])


def rz_pulse(w, p0=0.5):
    return expm(-1j * p0 * w * Z)


def cz_pulse(p=0.3183098783513154):
    return expm(-1j * p * np.pi * H_CZ)
