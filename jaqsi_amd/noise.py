"""Kraus noise channels (mirrors jaqsi/noise.py:32-465).

Recording a channel switches a Script to density-matrix simulation.  Each
channel lowers to its superoperator ``sum_k K_k (x) conj(K_k)`` on the
(ket, bra) wire pair, the same entry the reference builds in
``_compile_mixed`` (simulation.py:348-381).
"""

from __future__ import annotations

from typing import List, Union

import numpy as np
import torch

from jaqsi_amd.gateset import Id, PauliX, PauliY, PauliZ
from jaqsi_amd.operations import Operation, _as_tensor, _contract, cdtype


def _f(v) -> float:
    return float(v.detach()) if isinstance(v, torch.Tensor) else float(v)


def _check_unit_interval(value: float, name: str) -> None:
    if not 0.0 <= _f(value) <= 1.0:
        raise ValueError(f"{name} must be in [0, 1].")


def _needs_grad(*vals) -> bool:
    return any(isinstance(v, torch.Tensor) and v.requires_grad for v in vals)


def _real(v) -> torch.Tensor:
    """A channel parameter as a float64 0-d tensor (autograd kept)."""
    if isinstance(v, torch.Tensor):
        return v.to(torch.float64).reshape(())
    return torch.tensor(float(v), dtype=torch.float64)


def _unit(r: int, c: int) -> torch.Tensor:
    e = torch.zeros((4, 4), dtype=torch.complex128)
    e[r, c] = 1.0
    return e


def _entries(vals: dict) -> torch.Tensor:
    """A 4x4 superoperator from {(row, col): real 0-d tensor}."""
    dev = next(iter(vals.values())).device
    return sum(v.to(torch.complex128) * _unit(r, c).to(dev) for (r, c), v in vals.items())


def _pauli_mixture(weights) -> torch.Tensor:
    """sum_P w_P P (x) conj(P) for (w, P) pairs, in torch."""
    return sum(w.to(torch.complex128) * torch.as_tensor(np.kron(P, P.conj())).to(w.device) for w, P in weights)


class KrausChannel(Operation):
    """Base class of channels phi(rho) = sum_k K_k rho K_k^dagger.

    A channel whose parameters are tensors that require grad lowers to a
    superoperator built in torch (``superoperator_torch``), so the adjoint's
    gradient of that matrix reaches the parameters through autograd."""

    def kraus_matrices(self) -> List[np.ndarray]:
        raise NotImplementedError

    def superoperator_torch(self) -> torch.Tensor:
        raise NotImplementedError(f"gradients with respect to the parameters of {self.__class__.__name__} are not "
                                  "supported")

    @property
    def matrix(self):
        raise TypeError(
            f"{self.__class__.__name__} is a noise channel and has no single "
            "unitary matrix. Use apply_to_density() instead."
        )

    def apply_to_state(self, state, n_qubits):
        raise TypeError(
            f"{self.__class__.__name__} is a noise channel and cannot be "
            "applied to a pure statevector. Use execute(type='density') instead."
        )

    def superoperator(self):
        """sum_k K (x) conj(K) on (ket wires, bra wires) (simulation.py:373-375):
        numpy, or a torch tensor when a parameter requires grad."""
        if _needs_grad(*self.parameters):
            return self.superoperator_torch()
        ks = [np.asarray(_as_tensor(k).detach().cpu().numpy(), dtype=np.complex128) for k in self.kraus_matrices()]
        return sum(np.kron(k, np.conj(k)) for k in ks)

    def apply_to_density(self, rho, n_qubits):
        dim = 2**n_qubits
        bra = [w + n_qubits for w in self.wires]
        t0 = _as_tensor(rho).to(cdtype()).reshape((2,) * (2 * n_qubits))
        out = torch.zeros_like(t0)
        for k in self.kraus_matrices():
            kt = _as_tensor(k).to(cdtype())
            t = _contract(t0, kt, self.wires)
            out = out + _contract(t, kt.conj(), bra)
        return out.reshape(dim, dim)


class BitFlip(KrausChannel):
    """sqrt(1-p) I, sqrt(p) X (noise.py:123-158)."""

    _num_wires = 1
    _param_names = ("p",)

    def __init__(self, p: float, wires: Union[int, List[int]] = 0) -> None:
        _check_unit_interval(p, "p")
        self.p = p
        super().__init__(wires=wires)

    def kraus_matrices(self):
        p = _f(self.p)
        return [np.sqrt(1 - p) * Id._matrix, np.sqrt(p) * PauliX._matrix]

    def superoperator_torch(self):
        p = _real(self.p)
        return _pauli_mixture([(1 - p, Id._matrix), (p, PauliX._matrix)])


class PhaseFlip(KrausChannel):
    """sqrt(1-p) I, sqrt(p) Z (noise.py:161-196)."""

    _num_wires = 1
    _param_names = ("p",)

    def __init__(self, p: float, wires: Union[int, List[int]] = 0) -> None:
        _check_unit_interval(p, "p")
        self.p = p
        super().__init__(wires=wires)

    def kraus_matrices(self):
        p = _f(self.p)
        return [np.sqrt(1 - p) * Id._matrix, np.sqrt(p) * PauliZ._matrix]

    def superoperator_torch(self):
        p = _real(self.p)
        return _pauli_mixture([(1 - p, Id._matrix), (p, PauliZ._matrix)])


class DepolarizingChannel(KrausChannel):
    """sqrt(1-p) I, sqrt(p/3) X, Y, Z (noise.py:199-237)."""

    _num_wires = 1
    _param_names = ("p",)

    def __init__(self, p: float, wires: Union[int, List[int]] = 0) -> None:
        _check_unit_interval(p, "p")
        self.p = p
        super().__init__(wires=wires)

    def kraus_matrices(self):
        p = _f(self.p)
        return [
            np.sqrt(1 - p) * Id._matrix,
            np.sqrt(p / 3) * PauliX._matrix,
            np.sqrt(p / 3) * PauliY._matrix,
            np.sqrt(p / 3) * PauliZ._matrix,
        ]

    def superoperator_torch(self):
        p = _real(self.p)
        return _pauli_mixture([(1 - p, Id._matrix), (p / 3, PauliX._matrix), (p / 3, PauliY._matrix),
                               (p / 3, PauliZ._matrix)])


class AmplitudeDamping(KrausChannel):
    """[[1,0],[0,sqrt(1-g)]], [[0,sqrt(g)],[0,0]] (noise.py:240-277)."""

    _num_wires = 1
    _param_names = ("gamma",)

    def __init__(self, gamma: float, wires: Union[int, List[int]] = 0) -> None:
        _check_unit_interval(gamma, "gamma")
        self.gamma = gamma
        super().__init__(wires=wires)

    def kraus_matrices(self):
        g = _f(self.gamma)
        return [
            np.array([[1.0, 0.0], [0.0, np.sqrt(1 - g)]], dtype=np.complex128),
            np.array([[0.0, np.sqrt(g)], [0.0, 0.0]], dtype=np.complex128),
        ]

    def superoperator_torch(self):
        g = _real(self.gamma)
        c = torch.sqrt(1 - g)
        return _entries({(0, 0): torch.ones_like(g), (0, 3): g, (1, 1): c, (2, 2): c, (3, 3): 1 - g})


class PhaseDamping(KrausChannel):
    """[[1,0],[0,sqrt(1-g)]], [[0,0],[0,sqrt(g)]] (noise.py:280-316)."""

    _num_wires = 1
    _param_names = ("gamma",)

    def __init__(self, gamma: float, wires: Union[int, List[int]] = 0) -> None:
        _check_unit_interval(gamma, "gamma")
        self.gamma = gamma
        super().__init__(wires=wires)

    def kraus_matrices(self):
        g = _f(self.gamma)
        return [
            np.array([[1.0, 0.0], [0.0, np.sqrt(1 - g)]], dtype=np.complex128),
            np.array([[0.0, 0.0], [0.0, np.sqrt(g)]], dtype=np.complex128),
        ]

    def superoperator_torch(self):
        g = _real(self.gamma)
        c = torch.sqrt(1 - g)
        one = torch.ones_like(g)
        return _entries({(0, 0): one, (1, 1): c, (2, 2): c, (3, 3): one})


def kraus_from_superoperator(S: np.ndarray, tol: float = 1e-14) -> List[np.ndarray]:
    """Kraus operators of a completely positive map given by its
    superoperator S on row-major vec(rho) (S[(k,l),(i,j)] is the weight of
    rho_ij in rho'_kl).

    Reshuffling S to C[(k,i),(l,j)] = S[(k,l),(i,j)] gives the Choi matrix
    sum_m vec(K_m) vec(K_m)^dagger, vec(K)[(k,i)] = K_ki; its eigenvectors
    with positive eigenvalues, scaled by their square roots, are a minimal
    Kraus set (Choi 1975)."""
    d = int(round(np.sqrt(S.shape[0])))
    C = S.reshape(d, d, d, d).transpose(0, 2, 1, 3).reshape(d * d, d * d)
    lam, vec = np.linalg.eigh((C + C.conj().T) / 2)
    keep = lam > tol * max(1.0, float(lam.max()))
    return [np.sqrt(l) * vec[:, m].reshape(d, d) for m, l in zip(np.flatnonzero(keep), lam[keep])]


class ThermalRelaxationError(KrausChannel):
    """Relaxation toward the thermal state and dephasing over a gate of
    length ``tg`` (noise.py:319-431).

    With the relaxed fraction a = 1 - exp(-tg / t1), the excited-state
    population ``pe`` of the thermal state and the coherence decay
    c = exp(-tg / t2), the channel maps rho to

        rho'_11 = (1 - a) rho_11 + a pe,   rho'_00 = 1 - rho'_11,
        rho'_01 = c rho_01,                rho'_10 = c rho_10,

    which is completely positive for t2 <= 2 t1.  The Kraus operators are
    read off the map's Choi matrix (at most four; a Kraus set is fixed only up
    to a unitary mixing, the channel itself is unique)."""

    _num_wires = 1
    _param_names = ("pe", "t1", "t2", "tg")

    def __init__(self, pe: float, t1: float, t2: float, tg: float, wires: Union[int, List[int]] = 0) -> None:
        _check_unit_interval(pe, "pe")
        if not _f(t1) > 0 or not _f(t2) > 0:
            raise ValueError(f"relaxation times must be positive (t1={t1}, t2={t2}).")
        if _f(t2) > 2 * _f(t1):
            raise ValueError(f"t2={t2} exceeds 2*t1={2 * _f(t1)}: the channel would not be completely positive.")
        if _f(tg) < 0:
            raise ValueError(f"gate time tg={tg} must be non-negative.")
        self.pe, self.t1, self.t2, self.tg = pe, t1, t2, tg
        super().__init__(wires=wires)

    def relaxation_superoperator(self) -> np.ndarray:
        """The map above as a 4x4 matrix on row-major vec(rho)."""
        pe, t1, t2, tg = (_f(v) for v in (self.pe, self.t1, self.t2, self.tg))
        a = -np.expm1(-tg / t1)
        c = np.exp(-tg / t2)
        S = np.zeros((4, 4), dtype=np.complex128)
        S[3, 0] = a * pe            # rho_00 -> rho'_11
        S[3, 3] = 1.0 - a * (1.0 - pe)
        S[0, 0] = 1.0 - S[3, 0]     # trace preservation fixes the first row
        S[0, 3] = 1.0 - S[3, 3]
        S[1, 1] = S[2, 2] = c
        return S

    def kraus_matrices(self):
        return kraus_from_superoperator(self.relaxation_superoperator())

    def superoperator_torch(self):
        pe, t1, t2, tg = (_real(v) for v in (self.pe, self.t1, self.t2, self.tg))
        a = -torch.expm1(-tg / t1)
        c = torch.exp(-tg / t2)
        up, stay = a * pe, 1 - a * (1 - pe)
        return _entries({(3, 0): up, (3, 3): stay, (0, 0): 1 - up, (0, 3): 1 - stay, (1, 1): c, (2, 2): c})


class QubitChannel(KrausChannel):
    """A channel from user-supplied Kraus operators (noise.py:434-465)."""

    def __init__(self, kraus_ops, wires: Union[int, List[int]] = 0) -> None:
        self._kraus_ops = [np.asarray(_as_tensor(k).detach().cpu().numpy(), dtype=np.complex128) for k in kraus_ops]
        # tensors that require grad are kept: the superoperator is then built
        # from them in torch
        self._kraus_t = list(kraus_ops) if _needs_grad(*kraus_ops) else None
        super().__init__(wires=wires)

    def kraus_matrices(self):
        return self._kraus_ops

    def superoperator(self):
        if self._kraus_t is None:
            return super().superoperator()
        ks = [_as_tensor(k).to(torch.complex128) for k in self._kraus_t]
        return sum(torch.kron(k, k.conj()) for k in ks)
