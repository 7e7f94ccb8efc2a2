"""The seam to the reference's ``simulate_and_measure``
(jaqsi/simulation.py:658-732) for a ``jax.custom_vjp`` binding.

The reference hands its seam a tape of its own ``Operation`` objects
(jaqsi/operations.py:229-273, tape.py:24-55), whose parameters are tracers
under ``jax.grad`` / ``jax.jit``.  ``adapt_tape`` rebuilds that tape as this
package's operations by class name, wires and ``_param_names``
(jaqsi/gateset.py, noise.py), passing every parameter object through
untouched, and ``lower_for_jax`` lowers it to a plan structure plus the list
of those parameter objects in plan-slot order.  No parameter value is read,
so the caller can stack the (traced) parameters into the vector that is the
custom_vjp's primal input, and the cotangent its backward returns chains back
to the user's parameters through that stack.

What crosses as a constant (read once, concrete values required, no gradient
through this seam): explicit matrices (``DiagonalQubitUnitary``'s diagonal,
``Hermitian`` / ``RandomUnitary`` matrices, ``QubitChannel`` Kraus operators)
and channel strengths.
"""

from __future__ import annotations

from typing import List, Sequence

import numpy as np

from jaqsi_amd import gateset, lowering, noise
from jaqsi_amd.gateset import Barrier
from jaqsi_amd.operations import Operation
from jaqsi_amd.tape import recording

# gates rebuilt as cls(*params, wires=wires)
_PARAM_GATES = {"RX", "RY", "RZ", "Rot", "ControlledPhaseShift", "RXX", "RYY", "RZZ", "RZX", "CRX", "CRY", "CRZ"}
_FIXED_GATES = {"Id", "PauliX", "PauliY", "PauliZ", "H", "S", "SWAP", "CX", "CY", "CZ", "CCX", "CSWAP"}
_CHANNELS = {"BitFlip", "PhaseFlip", "DepolarizingChannel", "AmplitudeDamping", "PhaseDamping",
             "ThermalRelaxationError"}


def _params(op) -> list:
    return [getattr(op, n) for n in getattr(op, "_param_names", ())]


def _concrete(x) -> np.ndarray:
    if hasattr(x, "detach"):
        x = x.detach().cpu()
    return np.asarray(x)


def _adapt_one(op) -> Operation:
    name = type(op).__name__
    wires = list(op.wires)
    if name in _PARAM_GATES:
        return getattr(gateset, name)(*_params(op), wires=wires)
    if name in _FIXED_GATES:
        return getattr(gateset, name)(wires=wires)
    if name == "PauliRot":
        return gateset.PauliRot(op.theta, op.pauli_word, wires=wires)
    if name == "ControlledPauliRot":
        return gateset.ControlledPauliRot(op.theta, op.pauli_word, wires=wires, n_controls=op.n_controls)
    if name == "DiagonalQubitUnitary":
        return gateset.DiagonalQubitUnitary(_concrete(op.diag), wires=wires)
    if name in _CHANNELS:
        return getattr(noise, name)(*[float(_concrete(p)) for p in _params(op)], wires=wires)
    if name == "QubitChannel":
        ks = getattr(op, "_kraus_ops", None)
        ks = ks if ks is not None else op.kraus_matrices()
        return noise.QubitChannel([_concrete(k) for k in ks], wires=wires)
    m = getattr(op, "matrix", None)
    if m is None:
        raise NotImplementedError(f"no lowering for operation {name!r} on wires {wires}")
    # any other gate with a matrix (Hermitian, RandomUnitary, user classes)
    return Operation(wires=wires, matrix=_concrete(m), name=name)


def adapt_tape(tape: Sequence) -> List[Operation]:
    """This package's operations for a recorded tape: operations of this
    package pass through, others are rebuilt by class name (Barriers are
    dropped)."""
    out = []
    with recording():  # constructors record; keep them off any caller's tape
        for op in tape:
            if isinstance(op, Barrier) or type(op).__name__ == "Barrier":
                continue
            out.append(op if isinstance(op, Operation) else _adapt_one(op))
    return out


def lower_for_jax(tape: Sequence, n_qubits: int, use_density: bool):
    """(lowered, params): the plan structure of *tape* and its angle
    parameters in plan-slot order, as the objects the tape holds (floats,
    arrays or tracers; none is converted)."""
    ops = adapt_tape(tape)
    low = lowering.lower_mixed(ops, n_qubits) if use_density else lowering.lower_pure(ops, n_qubits)
    return (low, n_qubits, use_density), list(low.angle_values)
