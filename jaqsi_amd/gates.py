"""``Gates.<Name>(...)``: the recording entry point (mirrors jaqsi/gates.py).

Routes a call to ``UnitaryGates`` (gates.py:88-176), filtering keyword
arguments the same way.  Pulse-level gates (``pulse=True``, the reference's
``PulseGates``, jaqsi/pulses.py:1370-1665) are not lowered by this backend:
such calls raise ``NotImplementedError``.  Batched pulse propagators of the
two-qubit drive Hamiltonian are available from ``jaqsi_amd.pulse``.
"""

from __future__ import annotations

import logging
from contextlib import contextmanager
from typing import Callable, List, Union

from jaqsi_amd.gateset import Barrier as BarrierOp
from jaqsi_amd.unitary import UnitaryGates

log = logging.getLogger(__name__)

_PULSE_GATES = ("pulse-level gates (pulse=True) are not lowered by this backend; batched two-qubit pulse "
                "propagators with gradients are available from jaqsi_amd.pulse")


class _GateHandler:
    """What ``Gates.<name>`` returns (gates.py:22-35)."""

    def __init__(self, gates, gate_name: str):
        self._gates = gates
        self.__name__ = gate_name

    def __call__(self, *args, **kwargs):
        return self._gates._inner_getattr(self.__name__, *args, **kwargs)


class GatesMeta(type):
    def __getattr__(cls, gate_name):
        if gate_name.startswith("__"):
            raise AttributeError(gate_name)
        return _GateHandler(cls, gate_name)


def _no_op(*args, **kwargs):
    return None


def Barrier(wires: Union[int, List[int]], *args, **kwargs):
    return BarrierOp(wires)


class Gates(metaclass=GatesMeta):
    """Entry point for applying gates inside a circuit function."""

    @classmethod
    def _inner_getattr(cls, gate_name, *args, **kwargs):
        if gate_name == "Barrier":
            return Barrier(*args, **kwargs)
        pulse = kwargs.pop("pulse", False)
        if not isinstance(pulse, bool):
            raise TypeError(f"'pulse' must be a bool, got {type(pulse).__name__}.")
        if pulse:
            raise NotImplementedError(_PULSE_GATES)
        allowed = ["w", "wires", "phi", "theta", "omega", "noise_params", "random_key"]
        backend = UnitaryGates
        extra = kwargs.keys() - allowed
        if extra:
            log.debug(f"Unsupported keyword arguments: {list(extra)}")
        kwargs = {k: v for k, v in kwargs.items() if k in allowed}
        gate = getattr(backend, gate_name, None)
        if gate is None:
            raise AttributeError(f"'{backend.__name__}' object has no attribute '{gate_name}'")
        return gate(*args, **kwargs)

    @classmethod
    @contextmanager
    def pulse_manager_context(cls, pulse_params):
        raise NotImplementedError(_PULSE_GATES)
        yield  # pragma: no cover  (keeps this a generator for contextmanager)

    @classmethod
    def parse_gates(cls, gates, set_of_gates=None):
        set_of_gates = set_of_gates or cls
        if isinstance(gates, str):
            return [getattr(set_of_gates, gates)]
        if isinstance(gates, list):
            out = []
            for enc in gates:
                if isinstance(enc, str):
                    out.append(getattr(set_of_gates, enc))
                elif callable(enc):
                    out.append(enc)
                else:
                    raise ValueError(f"Operation {enc} is not a valid gate or callable. Got {type(enc)}")
            return out
        if callable(gates):
            return [gates]
        if gates is None:
            return [_no_op]
        raise ValueError(f"Operation {gates} is not a valid gate or callable or list of both.")

    @classmethod
    def is_rotational(cls, gate: Callable) -> bool:
        return gate.__name__ in ["RX", "RY", "RZ", "Rot", "RXX", "RYY", "RZZ", "RZX", "CRX", "CRY", "CRZ", "CPhase"]
