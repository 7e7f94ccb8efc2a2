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


def _check_unit_interval(value: float, name: str) -> None:
    if not 0.0 <= float(value) <= 1.0:
        raise ValueError(f"{name} must be in [0, 1].")


class KrausChannel(Operation):
    """Base class of channels phi(rho) = sum_k K_k rho K_k^dagger."""

    def kraus_matrices(self) -> List[np.ndarray]:
        raise NotImplementedError

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

    def superoperator(self) -> np.ndarray:
        """sum_k K (x) conj(K) on (ket wires, bra wires) (simulation.py:373-375)."""
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
        p = float(self.p)
        return [np.sqrt(1 - p) * Id._matrix, np.sqrt(p) * PauliX._matrix]


class PhaseFlip(KrausChannel):
    """sqrt(1-p) I, sqrt(p) Z (noise.py:161-196)."""

    _num_wires = 1
    _param_names = ("p",)

    def __init__(self, p: float, wires: Union[int, List[int]] = 0) -> None:
        _check_unit_interval(p, "p")
        self.p = p
        super().__init__(wires=wires)

    def kraus_matrices(self):
        p = float(self.p)
        return [np.sqrt(1 - p) * Id._matrix, np.sqrt(p) * PauliZ._matrix]


class DepolarizingChannel(KrausChannel):
    """sqrt(1-p) I, sqrt(p/3) X, Y, Z (noise.py:199-237)."""

    _num_wires = 1
    _param_names = ("p",)

    def __init__(self, p: float, wires: Union[int, List[int]] = 0) -> None:
        _check_unit_interval(p, "p")
        self.p = p
        super().__init__(wires=wires)

    def kraus_matrices(self):
        p = float(self.p)
        return [
            np.sqrt(1 - p) * Id._matrix,
            np.sqrt(p / 3) * PauliX._matrix,
            np.sqrt(p / 3) * PauliY._matrix,
            np.sqrt(p / 3) * PauliZ._matrix,
        ]


class AmplitudeDamping(KrausChannel):
    """[[1,0],[0,sqrt(1-g)]], [[0,sqrt(g)],[0,0]] (noise.py:240-277)."""

    _num_wires = 1
    _param_names = ("gamma",)

    def __init__(self, gamma: float, wires: Union[int, List[int]] = 0) -> None:
        _check_unit_interval(gamma, "gamma")
        self.gamma = gamma
        super().__init__(wires=wires)

    def kraus_matrices(self):
        g = float(self.gamma)
        return [
            np.array([[1.0, 0.0], [0.0, np.sqrt(1 - g)]], dtype=np.complex128),
            np.array([[0.0, np.sqrt(g)], [0.0, 0.0]], dtype=np.complex128),
        ]


class PhaseDamping(KrausChannel):
    """[[1,0],[0,sqrt(1-g)]], [[0,0],[0,sqrt(g)]] (noise.py:280-316)."""

    _num_wires = 1
    _param_names = ("gamma",)

    def __init__(self, gamma: float, wires: Union[int, List[int]] = 0) -> None:
        _check_unit_interval(gamma, "gamma")
        self.gamma = gamma
        super().__init__(wires=wires)

    def kraus_matrices(self):
        g = float(self.gamma)
        return [
            np.array([[1.0, 0.0], [0.0, np.sqrt(1 - g)]], dtype=np.complex128),
            np.array([[0.0, 0.0], [0.0, np.sqrt(g)]], dtype=np.complex128),
        ]


class ThermalRelaxationError(KrausChannel):
    """T1/T2 relaxation (noise.py:319-431)."""

    _num_wires = 1
    _param_names = ("pe", "t1", "t2", "tg")

    def __init__(self, pe: float, t1: float, t2: float, tg: float, wires: Union[int, List[int]] = 0) -> None:
        _check_unit_interval(pe, "pe")
        if t1 <= 0:
            raise ValueError("t1 must be > 0.")
        if t2 <= 0:
            raise ValueError("t2 must be > 0.")
        if t2 > 2 * t1:
            raise ValueError("t2 must be <= 2·t1.")
        if tg < 0:
            raise ValueError("tg must be >= 0.")
        self.pe, self.t1, self.t2, self.tg = pe, t1, t2, tg
        super().__init__(wires=wires)

    def kraus_matrices(self):
        pe, t1, t2, tg = (float(v) for v in (self.pe, self.t1, self.t2, self.tg))
        eT1 = np.exp(-tg / t1)
        p_reset = 1.0 - eT1
        eT2 = np.exp(-tg / t2)
        if t2 <= t1:
            pz = (1.0 - p_reset) * (1.0 - eT2 / eT1) / 2.0
            pr0 = (1.0 - pe) * p_reset
            pr1 = pe * p_reset
            pid = 1.0 - pz - pr0 - pr1
            return [
                np.sqrt(pid) * np.eye(2, dtype=np.complex128),
                np.sqrt(pz) * np.array([[1, 0], [0, -1]], dtype=np.complex128),
                np.sqrt(pr0) * np.array([[1, 0], [0, 0]], dtype=np.complex128),
                np.sqrt(pr0) * np.array([[0, 1], [0, 0]], dtype=np.complex128),
                np.sqrt(pr1) * np.array([[0, 0], [1, 0]], dtype=np.complex128),
                np.sqrt(pr1) * np.array([[0, 0], [0, 1]], dtype=np.complex128),
            ]
        choi = np.array(
            [
                [1 - pe * p_reset, 0, 0, eT2],
                [0, pe * p_reset, 0, 0],
                [0, 0, (1 - pe) * p_reset, 0],
                [eT2, 0, 0, 1 - (1 - pe) * p_reset],
            ],
            dtype=np.complex128,
        )
        w, v = np.linalg.eigh(choi)
        return [np.sqrt(abs(w[i])) * v[:, i].reshape(2, 2, order="F") for i in range(4)]


class QubitChannel(KrausChannel):
    """A channel from user-supplied Kraus operators (noise.py:434-465)."""

    def __init__(self, kraus_ops, wires: Union[int, List[int]] = 0) -> None:
        self._kraus_ops = [np.asarray(_as_tensor(k).detach().cpu().numpy(), dtype=np.complex128) for k in kraus_ops]
        super().__init__(wires=wires)

    def kraus_matrices(self):
        return self._kraus_ops
