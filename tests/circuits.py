"""Shared circuits for the tests: the same gates recorded through jaqsi_amd
and described for the oracle."""

import functools

import numpy as np

from oracle import sv


def hea_theta(n, layers, seed):
    return np.random.default_rng(seed).uniform(0, 2 * np.pi, (layers, n, 3))


def record_hea(theta, n, features=None):
    """Record the hardware-efficient ansatz through Gates (jaqsi_amd)."""
    from jaqsi_amd import Gates

    if features is not None:
        for q in range(n):
            Gates.RX(features[q], wires=q)
    for layer in range(theta.shape[0]):
        for q in range(n):
            Gates.RX(theta[layer, q, 0], wires=q)
            Gates.RY(theta[layer, q, 1], wires=q)
            Gates.RZ(theta[layer, q, 2], wires=q)
        for q in range(n):
            Gates.CX(wires=[q, (q + 1) % n])


def zoo_params(seed=3):
    return np.random.default_rng(seed).uniform(0.1, 2.0, 12)


def record_zoo(p):
    """Every gate kind of the library on 5 wires (shuffled wire orders)."""
    import jaqsi_amd as J

    J.H(0)
    J.RX(p[0], 3)
    J.RY(p[1], 1)
    J.CX([0, 1])
    J.CRX(p[2], [1, 2])
    J.Rot(p[3], p[4], p[5], 4)
    J.PauliRot(p[6], "XY", [0, 2])
    J.ControlledPhaseShift(p[7], [2, 1])
    J.CCX([3, 0, 2])
    J.CSWAP([2, 3, 0])
    J.SWAP([4, 1])
    J.CY([0, 3])
    J.CZ([4, 2])
    J.S(3)
    J.RZZ(p[8], [3, 0])
    J.CRZ(p[9], [0, 4])
    J.ControlledPauliRot(p[10], "XY", [3, 0, 2])
    J.ControlledPauliRot(p[11], "Y", [2, 0, 3], n_controls=2)
    J.PauliX(1)
    J.PauliY(2)
    J.RZ(p[0] * 0.5, 4)


def zoo_ops(p):
    O = sv.Op
    return [
        O("H", (0,)), O("RX", (3,), (p[0],)), O("RY", (1,), (p[1],)), O("CX", (0, 1)),
        O("CRX", (1, 2), (p[2],)), O("Rot", (4,), (p[3], p[4], p[5])), O("PauliRot", (0, 2), (p[6],), "XY"),
        O("CPhase", (2, 1), (p[7],)), O("CCX", (3, 0, 2)), O("CSWAP", (2, 3, 0)), O("SWAP", (4, 1)),
        O("CY", (0, 3)), O("CZ", (4, 2)), O("S", (3,)), O("RZZ", (3, 0), (p[8],)), O("CRZ", (0, 4), (p[9],)),
        O("ControlledPauliRot", (3, 0, 2), (p[10],), ("XY", 1)),
        O("ControlledPauliRot", (2, 0, 3), (p[11],), ("Y", 2)),
        O("PauliX", (1,)), O("PauliY", (2,)), O("RZ", (4,), (p[0] * 0.5,)),
    ]


# parameter index of every oracle parameter in the flat zoo parameter vector
ZOO_GRAD_MAP = [0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 0]


@functools.lru_cache(maxsize=None)
def noisy_hea_reference(n, layers, seed, p_dep=0.01, gamma=0.005):
    """Oracle value and parameter-shift gradient of <Z_0> for the noisy ansatz
    (cached: the density-matrix shift rule is the slow part of a GPU test)."""
    from oracle import dm

    theta = hea_theta(n, layers, seed)
    v, g = dm.value_and_shift_grad(dm.noisy_hea(theta, n, p_dep, gamma), n, [(sv.PZ, (0,))], np.array([1.0]))
    return v, g
