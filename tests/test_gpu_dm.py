"""GPU parity of the density-matrix path (superoperators on 2n wires, Kraus
channels, checkpointed adjoint) against the complex128 density oracle."""

import numpy as np
import pytest
import torch

from circuits import hea_theta, noisy_hea_reference
from oracle import dm, sv
from parity import assert_parity

pytestmark = pytest.mark.gpu



def _record_noisy_hea(theta, n, p_dep, gamma):
    """The noisy ansatz: after every gate, on each touched wire in order,
    DepolarizingChannel(p_dep) then AmplitudeDamping(gamma)."""
    import jaqsi_amd as J
    from jaqsi_amd import noise

    def after(wires):
        for w in wires:
            noise.DepolarizingChannel(p_dep, wires=w)
            noise.AmplitudeDamping(gamma, wires=w)

    for layer in range(theta.shape[0]):
        for q in range(n):
            J.Gates.RX(theta[layer, q, 0], wires=q)
            after([q])
            J.Gates.RY(theta[layer, q, 1], wires=q)
            after([q])
            J.Gates.RZ(theta[layer, q, 2], wires=q)
            after([q])
        for q in range(n):
            J.Gates.CX(wires=[q, (q + 1) % n])
            after([q, (q + 1) % n])


@pytest.mark.parametrize("x64", [False, True])
@pytest.mark.parametrize("n,layers", [(2, 1), (3, 2), (5, 1), (7, 1)])
def test_noisy_hea_expval_and_gradient(n, layers, x64):
    import jaqsi_amd as J

    print(f"dm noisy hea n={n} layers={layers} x64={x64}", flush=True)
    J.enable_x64(x64)
    theta = hea_theta(n, layers, 2000 + n)
    th = torch.tensor(theta, dtype=torch.float64, device="cuda", requires_grad=True)
    script = J.Script(lambda t: _record_noisy_hea(t, n, 0.01, 0.005), n_qubits=n)
    ev = script.execute(type="expval", obs=[J.PauliZ(wires=0, record=False)], args=(th,))
    ev.sum().backward()
    v, g = noisy_hea_reference(n, layers, 2000 + n)
    assert_parity(ev.detach().cpu().numpy()[0], v, x64)
    assert_parity(th.grad.cpu().numpy().reshape(-1), g, x64)


@pytest.mark.parametrize("x64", [False, True])
def test_density_and_probs_outputs(x64):
    import jaqsi_amd as J
    from jaqsi_amd import noise

    print(f"dm outputs x64={x64}", flush=True)
    J.enable_x64(x64)
    n = 3

    def circ():
        J.Gates.H(wires=0)
        J.Gates.CX(wires=[0, 1])
        noise.BitFlip(0.1, wires=1)
        J.Gates.RY(0.4, wires=2)
        noise.PhaseDamping(0.2, wires=2)
        noise.DepolarizingChannel(0.05, wires=0)
        J.Gates.CRY(0.7, wires=[1, 2])
        noise.AmplitudeDamping(0.3, wires=1)

    items = [
        sv.Op("H", (0,)),
        sv.Op("CX", (0, 1)),
        ("kraus", dm.kraus("BitFlip", 0.1), (1,)),
        sv.Op("RY", (2,), (0.4,)),
        ("kraus", dm.kraus("PhaseDamping", 0.2), (2,)),
        ("kraus", dm.kraus("DepolarizingChannel", 0.05), (0,)),
        sv.Op("CRY", (1, 2), (0.7,)),
        ("kraus", dm.kraus("AmplitudeDamping", 0.3), (1,)),
    ]
    ref = dm.simulate(items, n)
    rho = J.Script(circ, n_qubits=n).execute(type="density").cpu().numpy()
    assert_parity(rho, ref, x64)
    probs = J.Script(circ, n_qubits=n).execute(type="probs").cpu().numpy()
    assert_parity(probs, dm.probs(ref), x64)
    ev = J.Script(circ, n_qubits=n).execute(type="expval", obs=[J.PauliX(wires=2, record=False), J.PauliZ(wires=1, record=False)])
    assert_parity(ev.cpu().numpy(), [dm.expval(ref, sv.PX, (2,), n), dm.expval(ref, sv.PZ, (1,), n)], x64)


def test_noise_params_attach_channels_on_gpu():
    import jaqsi_amd as J

    print("dm noise_params", flush=True)
    J.enable_x64(True)
    n = 2

    def circ():
        J.Gates.RX(0.3, wires=0, noise_params={"BitFlip": 0.1, "Depolarizing": 0.02})
        J.Gates.CX(wires=[0, 1], noise_params={"PhaseFlip": 0.05})

    items = [
        sv.Op("RX", (0,), (0.3,)),
        ("kraus", dm.kraus("BitFlip", 0.1), (0,)),
        ("kraus", dm.kraus("DepolarizingChannel", 0.02), (0,)),
        sv.Op("CX", (0, 1)),
        ("kraus", dm.kraus("PhaseFlip", 0.05), (0,)),
        ("kraus", dm.kraus("PhaseFlip", 0.05), (1,)),
    ]
    ref = dm.simulate(items, n)
    rho = J.Script(circ, n_qubits=n).execute(type="density").cpu().numpy()
    assert_parity(rho, ref, True)
