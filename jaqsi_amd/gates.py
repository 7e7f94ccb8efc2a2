"""``Gates.<Name>(...)``: the recording entry point (mirrors jaqsi/gates.py).

Routes a call to ``UnitaryGates`` or, with ``pulse=True``, to
``PulseGates`` (gates.py:88-176), filtering keyword arguments the same way
and checking ``pulse_params`` as the reference does.  Under
``pulse_manager_context`` the model's pulse parameters scale each gate's
calibrated parameters element-wise (gates.py:142-176).
"""

from __future__ import annotations

import logging
import numbers
from contextlib import contextmanager
from typing import Callable, List, Union

import numpy as np
import torch

from jaqsi_amd.gateset import Barrier as BarrierOp
from jaqsi_amd.pulses import PulseEnvelope, PulseGates, PulseInformation, PulseParamManager, PulseParams  # noqa: F401
from jaqsi_amd.tape import require_per_sample_recording
from jaqsi_amd.unitary import UnitaryGates

log = logging.getLogger(__name__)


_PLAIN_KWARGS = frozenset(("wires",))


class _GateHandler:
    """What ``Gates.<name>`` returns (gates.py:22-35)."""

    def __init__(self, gates, gate_name: str):
        self._gates = gates
        self.__name__ = gate_name
        # the unitary gate of this name, called directly when only wires are
        # given (for rotations and fixed gates that is the gate class itself:
        # their noise-free wrapper only forwards to it)
        plain = None if gate_name == "Barrier" else getattr(UnitaryGates, gate_name, None)
        self._plain = getattr(plain, "gate_class", plain)

    def __call__(self, *args, **kwargs):
        if self._plain is not None and kwargs.keys() <= _PLAIN_KWARGS:
            return self._plain(*args, **kwargs)
        return self._gates._inner_getattr(self.__name__, *args, **kwargs)


class GatesMeta(type):
    def __getattr__(cls, gate_name):
        if gate_name.startswith("__"):
            raise AttributeError(gate_name)
        h = cls._handlers.get(gate_name)
        if h is None:
            h = cls._handlers[gate_name] = _GateHandler(cls, gate_name)
            # later lookups find the handler as a class attribute instead of
            # failing the normal lookup first (which costs more than the gate)
            type.__setattr__(cls, gate_name, h)
        return h


def _no_op(*args, **kwargs):
    return None


def Barrier(wires: Union[int, List[int]], *args, **kwargs):
    return BarrierOp(wires)


class Gates(metaclass=GatesMeta):
    """Entry point for applying gates inside a circuit function."""

    _handlers: dict = {}
    _pulse_mgr = None

    @classmethod
    def _inner_getattr(cls, gate_name, *args, **kwargs):
        if gate_name == "Barrier":
            return Barrier(*args, **kwargs)
        if kwargs.keys() <= _PLAIN_KWARGS:
            # the common case (a unitary gate given only wires) skips the
            # keyword filtering and pulse handling below
            gate = getattr(UnitaryGates, gate_name, None)
            if gate is None:
                raise AttributeError(f"'UnitaryGates' object has no attribute '{gate_name}'")
            return gate(*args, **kwargs)
        pulse = kwargs.pop("pulse", False)
        if not isinstance(pulse, bool):
            raise TypeError(f"'pulse' must be a bool, got {type(pulse).__name__}.")
        if pulse:
            require_per_sample_recording()
        allowed = ["w", "wires", "phi", "theta", "omega", "noise_params", "random_key"]
        if pulse:
            backend = PulseGates
            allowed += ["pulse_params"]
        else:
            backend = UnitaryGates
        extra = kwargs.keys() - allowed
        if extra:
            log.debug(f"Unsupported keyword arguments: {list(extra)}")
        kwargs = {k: v for k, v in kwargs.items() if k in allowed}
        pulse_params = kwargs.get("pulse_params")
        mgr = cls._pulse_mgr
        if pulse_params is not None:
            if isinstance(pulse_params, PulseParams):
                kwargs["pulse_params"] = pulse_params.params
                flat = list(np.ravel(np.asarray(pulse_params.params)))
            elif isinstance(pulse_params, torch.Tensor):
                flat = list(pulse_params.reshape(-1))
            elif isinstance(pulse_params, (list, tuple, np.ndarray)):
                flat = list(np.ravel(np.asarray(pulse_params, dtype=object)))
            else:
                raise TypeError(f"Unsupported pulse_params type: {type(pulse_params)}")
            if not all(isinstance(x, (numbers.Real, torch.Tensor)) for x in flat):
                raise TypeError(f"All elements in pulse_params must be int or float, got {pulse_params}, "
                                f"type {type(pulse_params)}. ")
            if not isinstance(mgr, PulseParamManager):
                n_params = PulseInformation.gate_by_name(gate_name).size
                if len(flat) != n_params:
                    raise ValueError(f"Gate '{gate_name}' expects {n_params} pulse parameters, got {len(flat)}")
            if isinstance(pulse_params, (list, tuple)) and any(isinstance(x, torch.Tensor) for x in flat):
                kwargs["pulse_params"] = torch.stack([torch.as_tensor(x, dtype=torch.float64) for x in flat])
        if pulse and isinstance(mgr, PulseParamManager):
            n_params = PulseInformation.gate_by_name(gate_name).size
            scalers = mgr.get(n_params)
            base = PulseInformation.gate_by_name(gate_name).params
            if isinstance(scalers, torch.Tensor):
                kwargs["pulse_params"] = torch.as_tensor(base, dtype=torch.float64, device=scalers.device) * scalers
            else:
                kwargs["pulse_params"] = np.asarray(base) * np.asarray(scalers)
        gate = getattr(backend, gate_name, None)
        if gate is None:
            raise AttributeError(f"'{backend.__name__}' object has no attribute '{gate_name}'")
        return gate(*args, **kwargs)

    @classmethod
    @contextmanager
    def pulse_manager_context(cls, pulse_params):
        """Scale every pulse gate recorded inside by consecutive slices of
        ``pulse_params`` (gates.py:178-186)."""
        cls._pulse_mgr = PulseParamManager(pulse_params)
        try:
            yield
        finally:
            cls._pulse_mgr = None

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
