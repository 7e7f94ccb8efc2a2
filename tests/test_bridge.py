"""The custom_vjp seam (jaqsi_amd.bridge): a tape of foreign operations
laid out like the reference's (class names, ``wires``, ``_param_names``,
``pauli_word`` / ``n_controls`` / ``diag`` / ``_kraus_ops``) is rebuilt
without reading any parameter value, and its angle parameters come back as
the same objects in the plan's slot order (CPU only: JAX is not part of this
environment, so objects that refuse conversion stand in for tracers)."""

import numpy as np
import pytest

from jaqsi_amd import lowering
from jaqsi_amd.bridge import adapt_tape, lower_for_jax


class Traced:
    """Stands in for a JAX tracer: any attempt to read its value fails."""

    def __init__(self, tag):
        self.tag = tag

    def __float__(self):
        raise TypeError("concretization of a traced value")

    __int__ = __index__ = __bool__ = __float__

    def __array__(self, *a, **k):
        raise TypeError("concretization of a traced value")


def _foreign(name, wires, params=(), **attrs):
    """An object shaped like a reference Operation of class *name*."""
    names = attrs.pop("_param_names", ("theta", "phi", "omega")[: len(params)])
    cls = type(name, (), {"_param_names": tuple(names)})
    op = cls()
    op.wires = list(wires)
    for n, v in zip(names, params):
        setattr(op, n, v)
    for k, v in attrs.items():
        setattr(op, k, v)
    return op


def _tape(p):
    F = _foreign
    return [
        F("H", [0]), F("RX", [3], [p[0]]), F("RY", [1], [p[1]]), F("CX", [0, 1]),
        F("CRX", [1, 2], [p[2]]), F("Rot", [4], [p[3], p[4], p[5]], _param_names=("phi", "theta", "omega")),
        F("PauliRot", [0, 2], [p[6]], pauli_word="XY"), F("ControlledPhaseShift", [2, 1], [p[7]], _param_names=("phi",)),
        F("CCX", [3, 0, 2]), F("Barrier", [0, 1]), F("RZZ", [3, 0], [p[8]]),
        F("ControlledPauliRot", [2, 0, 3], [p[9]], pauli_word="Y", n_controls=2),
        F("DiagonalQubitUnitary", [1, 4], diag=np.exp(1j * np.arange(4.0))),
        F("Hermitian", [2], matrix=np.array([[0.0, 1.0], [1.0, 0.0]])),
    ]


def test_parameters_cross_untouched_in_slot_order():
    traced = [Traced(i) for i in range(10)]
    (low, n, dens), params = lower_for_jax(_tape(traced), 5, False)
    assert n == 5 and not dens
    # the same circuit with float parameters gives the same plan structure and
    # the same slot order
    floats = [float(i) for i in range(10)]
    (low_f, _, _), params_f = lower_for_jax(_tape(floats), 5, False)
    assert low.key == low_f.key
    assert all(isinstance(p, Traced) for p in params)
    assert [p.tag for p in params] == [int(v) for v in params_f]
    # and it is the structure of this package's own recording of the circuit
    ops = adapt_tape(_tape(floats))
    assert lowering.lower_pure(ops, 5).key == low.key
    assert [type(o).__name__ for o in ops][:3] == ["H", "RX", "RY"]


def test_channels_and_density_lowering():
    F = _foreign
    tape = [F("RX", [0], [Traced(0)]), F("BitFlip", [0], [0.1], _param_names=("p",)),
            F("ThermalRelaxationError", [1], [0.1, 1.0, 0.8, 0.2], _param_names=("pe", "t1", "t2", "tg")),
            F("QubitChannel", [0, 1], _kraus_ops=[np.eye(4) * np.sqrt(0.9), np.kron(np.eye(2), [[0, 1], [1, 0]]) * np.sqrt(0.1)]),
            F("CZ", [0, 1])]
    (low, n, dens), params = lower_for_jax(tape, 2, True)
    assert dens and len(params) == 1 and params[0].tag == 0
    assert low.n_wires == 4


def test_unknown_operation_without_a_matrix_is_rejected():
    with pytest.raises(NotImplementedError, match="Mystery"):
        adapt_tape([_foreign("Mystery", [0])])
