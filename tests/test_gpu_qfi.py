"""GPU quantum Fisher information (jaqsi_amd.math, math.py:331-419) against
the oracle: state derivatives by the exact shift rule of Pauli rotations
(d psi / d t = [psi(t + pi) - psi(t - pi)] / 4, d rho / d t =
[rho(t + pi/2) - rho(t - pi/2)] / 2), the pure QFI from its Fubini-Study
form and the mixed QFI from the symmetric logarithmic derivative solved as a
Lyapunov equation (rho L + L rho = 2 d rho), an independent route to the
eigenbasis formula the library uses."""

import numpy as np
import pytest
import torch
from scipy.linalg import solve_continuous_lyapunov

from oracle import dm, sv
from parity import assert_parity

pytestmark = pytest.mark.gpu


def _shift_jac(fn, x, shift, scale):
    cols = []
    for e in np.eye(len(x)):
        cols.append((fn(x + shift * e) - fn(x - shift * e)) / scale)
    return np.stack(cols, axis=-1)


def _pure_qfi(psi, J):
    A = J.conj().T @ J
    v = J.conj().T @ psi
    return 4 * np.real(A - np.outer(v, v.conj()))


def _sld_qfi(rho, dr):
    P = dr.shape[-1]
    Ls = [solve_continuous_lyapunov(rho, 2 * dr[:, :, i]) for i in range(P)]
    return np.array([[np.real(np.trace(rho @ (Ls[i] @ Ls[j] + Ls[j] @ Ls[i]) / 2)) for j in range(P)]
                     for i in range(P)])


def _hea(t, n):
    import jaqsi_amd as J

    for layer in range(t.shape[0]):
        for q in range(n):
            J.Gates.RY(t[layer, q, 0], wires=q)
            J.Gates.RZ(t[layer, q, 1], wires=q)
        for q in range(n - 1):
            J.Gates.CX(wires=[q, q + 1])


def _hea_ops(t, n):
    ops = []
    for layer in range(t.shape[0]):
        for q in range(n):
            ops += [sv.Op("RY", (q,), (t[layer, q, 0],)), sv.Op("RZ", (q,), (t[layer, q, 1],))]
        ops += [sv.Op("CX", (q, q + 1)) for q in range(n - 1)]
    return ops


@pytest.mark.parametrize("x64", [False, True])
def test_pure_state_qfi_and_metric(x64):
    import jaqsi_amd as J
    from jaqsi_amd.math import fubini_study_metric, quantum_fisher_information

    print(f"qfi pure x64={x64}", flush=True)
    J.enable_x64(x64)
    try:
        n, shape = 3, (2, 3, 2)
        x = np.random.default_rng(5).uniform(0, 2 * np.pi, shape)
        script = J.Script(lambda t: _hea(t, n), n_qubits=n)
        fn = lambda t: script.execute(type="state", args=(t,))  # noqa: E731
        F = quantum_fisher_information(fn, torch.tensor(x, device="cuda")).cpu().numpy()
        G = fubini_study_metric(fn, torch.tensor(x, device="cuda")).cpu().numpy()
        flat = lambda v: sv.simulate(_hea_ops(v.reshape(shape), n), n)  # noqa: E731
        Jx = _shift_jac(flat, x.reshape(-1), np.pi, 4.0)
        ref = _pure_qfi(flat(x.reshape(-1)), Jx)
        assert F.shape == (x.size, x.size)
        assert_parity(F, ref, x64, scale=np.abs(ref).max(), err_msg="QFI")
        assert_parity(G, ref / 4, x64, scale=np.abs(ref).max() / 4, err_msg="Fubini-Study metric")
    finally:
        J.enable_x64(False)


def test_chain_rule_through_shared_and_scaled_parameters():
    """One parameter feeds two gates, another enters scaled and shifted: the
    QFI of the user parameters is A^T F A with A the angle Jacobian."""
    import jaqsi_amd as J
    from jaqsi_amd.math import quantum_fisher_information

    print("qfi chain rule", flush=True)
    J.enable_x64(True)
    try:
        def circ(p):
            J.Gates.RX(2.0 * p[0], wires=0)
            J.Gates.RY(p[0] + 0.3, wires=1)
            J.Gates.CX(wires=[0, 1])
            J.Gates.RZ(-1.5 * p[1], wires=1)
            J.Gates.RY(p[1], wires=0)
            J.Gates.CX(wires=[1, 0])

        def ops(p):
            return [sv.Op("RX", (0,), (2.0 * p[0],)), sv.Op("RY", (1,), (p[0] + 0.3,)), sv.Op("CX", (0, 1)),
                    sv.Op("RZ", (1,), (-1.5 * p[1],)), sv.Op("RY", (0,), (p[1],)), sv.Op("CX", (1, 0))]

        x = np.array([0.4, 1.2])
        script = J.Script(circ, n_qubits=2)
        F = quantum_fisher_information(lambda p: script.execute(type="state", args=(p,)),
                                       torch.tensor(x, device="cuda")).cpu().numpy()
        angles = lambda v: np.array([2 * v[0], v[0] + 0.3, -1.5 * v[1], v[1]])  # noqa: E731
        A = np.array([[2.0, 0.0], [1.0, 0.0], [0.0, -1.5], [0.0, 1.0]])

        def by_angles(a):
            return sv.simulate([sv.Op("RX", (0,), (a[0],)), sv.Op("RY", (1,), (a[1],)), sv.Op("CX", (0, 1)),
                                sv.Op("RZ", (1,), (a[2],)), sv.Op("RY", (0,), (a[3],)), sv.Op("CX", (1, 0))], 2)

        a0 = angles(x)
        ref = A.T @ _pure_qfi(by_angles(a0), _shift_jac(by_angles, a0, np.pi, 4.0)) @ A
        assert_parity(F, ref, True, scale=np.abs(ref).max())
        assert np.allclose(sv.simulate(ops(x), 2), by_angles(a0))
    finally:
        J.enable_x64(False)


@pytest.mark.parametrize("x64", [False, True])
def test_mixed_state_qfi(x64):
    import jaqsi_amd as J
    from jaqsi_amd import noise
    from jaqsi_amd.math import quantum_fisher_information

    print(f"qfi mixed x64={x64}", flush=True)
    J.enable_x64(x64)
    try:
        def circ(p):
            J.Gates.RY(p[0], wires=0)
            noise.DepolarizingChannel(0.1, wires=0)
            J.Gates.RX(p[1], wires=1)
            J.Gates.CX(wires=[0, 1])
            noise.AmplitudeDamping(0.2, wires=1)
            J.Gates.RZ(p[2], wires=0)
            J.Gates.RY(p[3], wires=1)
            noise.BitFlip(0.05, wires=1)

        def items(p):
            return [sv.Op("RY", (0,), (p[0],)), ("kraus", dm.kraus("DepolarizingChannel", 0.1), (0,)),
                    sv.Op("RX", (1,), (p[1],)), sv.Op("CX", (0, 1)), ("kraus", dm.kraus("AmplitudeDamping", 0.2), (1,)),
                    sv.Op("RZ", (0,), (p[2],)), sv.Op("RY", (1,), (p[3],)), ("kraus", dm.kraus("BitFlip", 0.05), (1,))]

        x = np.array([0.9, -0.6, 1.4, 0.3])
        script = J.Script(circ, n_qubits=2)
        F = quantum_fisher_information(lambda p: script.execute(type="density", args=(p,)),
                                       torch.tensor(x, device="cuda")).cpu().numpy()
        rho_fn = lambda v: dm.simulate(items(v), 2)  # noqa: E731
        ref = _sld_qfi(rho_fn(x), _shift_jac(rho_fn, x, np.pi / 2, 2.0))
        assert_parity(F, ref, x64, scale=np.abs(ref).max())
    finally:
        J.enable_x64(False)
