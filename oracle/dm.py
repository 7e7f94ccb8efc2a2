"""NumPy complex128 restatement of the reference density-matrix engine.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Restates:

* Kraus sets of the channels                 -- jaqsi/noise.py:123-465,
                                                jaqsi/unitary.py:21-79
* rho -> U rho U^dagger, rho -> sum K rho K^dagger
                                             -- jaqsi/simulation.py:348-381, 443-473
* Tr(O rho) and diag(rho)                    -- jaqsi/simulation.py:814-857

rho is a (2^n, 2^n) matrix; wire 0 is the most significant bit of both the
row (ket) and column (bra) index, as in the statevector engine.

How it is pinned (tests/test_oracle_dm.py): trace preservation of every
Kraus set; depolarizing and amplitude damping against their closed-form
actions on the Bloch vector / excited population; thermal relaxation against
the T1 / T2 decay of populations and coherences (the reference's
tests/test_jaqsi.py:689-711); the multi-qubit depolarizing channel against
rho -> (1 - q) rho + q I / 2^k with q = p (its Pauli twirl); noiseless circuits against
|psi><psi| from the statevector oracle; Hermiticity, unit trace and positive
semidefiniteness after noisy circuits; and gradients of Tr(Z rho) against the
exact parameter-shift rule (exact for exp(-i theta P / 2) gates inside any
fixed CPTP circuit, since the value is a degree-1 trigonometric polynomial in
theta) and central finite differences.
"""

from __future__ import annotations

import math
from typing import List, Sequence, Tuple

import numpy as np

from . import sv

C = np.complex128


# ---------------------------------------------------------------------------
# Kraus sets (noise.py)
# ---------------------------------------------------------------------------


def kraus(name: str, *params: float) -> List[np.ndarray]:
    I2, X, Y, Z = sv.I2, sv.PX, sv.PY, sv.PZ
    if name == "BitFlip":  # noise.py:123-158
        (p,) = params
        return [math.sqrt(1 - p) * I2, math.sqrt(p) * X]
    if name == "PhaseFlip":  # noise.py:161-196
        (p,) = params
        return [math.sqrt(1 - p) * I2, math.sqrt(p) * Z]
    if name == "DepolarizingChannel":  # noise.py:199-237
        (p,) = params
        return [math.sqrt(1 - p) * I2] + [math.sqrt(p / 3) * P for P in (X, Y, Z)]
    if name == "AmplitudeDamping":  # noise.py:240-277
        (g,) = params
        return [np.array([[1, 0], [0, math.sqrt(1 - g)]], dtype=C), np.array([[0, math.sqrt(g)], [0, 0]], dtype=C)]
    if name == "PhaseDamping":  # noise.py:280-316
        (g,) = params
        return [np.array([[1, 0], [0, math.sqrt(1 - g)]], dtype=C), np.array([[0, 0], [0, math.sqrt(g)]], dtype=C)]
    if name == "ThermalRelaxationError":  # noise.py:319-431, regime t2 <= t1
        # Markovian mixture: identity, a Z flip with the dephasing that
        # relaxation does not already cause, and resets to |0> / |1> with the
        # relaxed fraction split by the thermal population pe
        pe, t1, t2, tg = params
        if t2 > t1:
            raise ValueError("the oracle restates the t2 <= t1 regime only")
        relaxed = 1.0 - math.exp(-tg / t1)
        flip = 0.5 * (math.exp(-tg / t1) - math.exp(-tg / t2))
        r0, r1 = (1.0 - pe) * relaxed, pe * relaxed
        keep = 1.0 - flip - relaxed
        P0 = np.array([[1, 0], [0, 0]], dtype=C)
        P1 = np.array([[0, 0], [0, 1]], dtype=C)
        L0 = np.array([[0, 1], [0, 0]], dtype=C)  # |0><1|
        L1 = np.array([[0, 0], [1, 0]], dtype=C)  # |1><0|
        return [math.sqrt(keep) * I2, math.sqrt(flip) * Z, math.sqrt(r0) * P0, math.sqrt(r0) * L0,
                math.sqrt(r1) * L1, math.sqrt(r1) * P1]
    if name == "MultiQubitDepolarizing":  # unitary.py:21-79
        # every non-identity Pauli string on k wires with weight p / 4^k
        p, k = params
        ks = []
        for word in np.ndindex(*(4,) * int(k)):
            P = np.eye(1, dtype=C)
            for c in word:
                P = np.kron(P, (I2, X, Y, Z)[c])
            w = 1.0 - p * (4**k - 1) / 4**k if not any(word) else p / 4**k
            ks.append(math.sqrt(w) * P)
        return ks
    raise ValueError(name)


# ---------------------------------------------------------------------------
# evolution
# ---------------------------------------------------------------------------


def initial_rho(n: int) -> np.ndarray:
    rho = np.zeros((2**n, 2**n), dtype=C)
    rho[0, 0] = 1
    return rho


def apply_unitary(rho: np.ndarray, m: np.ndarray, wires: Sequence[int], n: int) -> np.ndarray:
    U = sv.lifted_kron(m, wires, n)
    return U @ rho @ U.conj().T


def apply_channel(rho: np.ndarray, ks: Sequence[np.ndarray], wires: Sequence[int], n: int) -> np.ndarray:
    out = np.zeros_like(rho)
    for k in ks:
        K = sv.lifted_kron(k, wires, n)
        out += K @ rho @ K.conj().T
    return out


class Chan(Tuple):
    pass


def simulate(items, n: int, rho0: np.ndarray = None) -> np.ndarray:
    """items: sequence of sv.Op (unitary gates) or ("kraus", [K...], wires)."""
    rho = initial_rho(n) if rho0 is None else np.array(rho0, dtype=C)
    for it in items:
        if isinstance(it, sv.Op):
            rho = apply_unitary(rho, sv.op_matrix(it), it.wires, n)
        else:
            _, ks, wires = it
            rho = apply_channel(rho, ks, wires, n)
    return rho


def expval(rho: np.ndarray, m: np.ndarray, wires: Sequence[int], n: int) -> float:
    return float(np.real(np.trace(sv.lifted_kron(m, wires, n) @ rho)))


def probs(rho: np.ndarray) -> np.ndarray:
    return np.real(np.diag(rho)).copy()


# ---------------------------------------------------------------------------
# gradients (exact parameter shift for Pauli-rotation parameters)
# ---------------------------------------------------------------------------

_SHIFTABLE = {"RX", "RY", "RZ", "RXX", "RYY", "RZZ", "RZX", "PauliRot"}


def value_and_shift_grad(items, n: int, obs: Sequence[Tuple[np.ndarray, Tuple[int, ...]]], ct: np.ndarray):
    """sum_o ct_o Tr(O_o rho) and its gradient w.r.t. the parameter of every
    shiftable gate in ``items`` (in order of appearance)."""

    def value(its):
        rho = simulate(its, n)
        return sum(c * expval(rho, m, w, n) for c, (m, w) in zip(ct, obs))

    v = value(items)
    grads = []
    for j, it in enumerate(items):
        if isinstance(it, sv.Op) and it.name in _SHIFTABLE:
            (t,) = it.params
            plus = list(items)
            minus = list(items)
            plus[j] = it._replace(params=(t + math.pi / 2,))
            minus[j] = it._replace(params=(t - math.pi / 2,))
            grads.append((value(plus) - value(minus)) / 2)
    return v, np.array(grads)


def noisy_hea(theta: np.ndarray, n: int, p_dep: float, gamma: float):
    """The noisy ansatz: after every gate, on each touched wire in order,
    DepolarizingChannel(p_dep) then AmplitudeDamping(gamma)."""
    items = []
    for op in sv.hea_ops(theta, n):
        items.append(op)
        for w in op.wires:
            items.append(("kraus", kraus("DepolarizingChannel", p_dep), (w,)))
            items.append(("kraus", kraus("AmplitudeDamping", gamma), (w,)))
    return items
