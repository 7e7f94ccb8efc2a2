"""Edge cases of the Script path on the GPU: an empty circuit, one qubit
with the closed form of RX (the reference's tests/test_jaqsi.py:818-837),
a circuit without parameters, and a batch of one."""

import numpy as np
import pytest
import torch

from oracle import sv
from parity import assert_parity

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("x64", [False, True])
def test_empty_circuit_is_the_zero_state(x64):
    import jaqsi_amd as J

    print(f"edge empty x64={x64}", flush=True)
    J.enable_x64(x64)
    n = 3
    script = J.Script(lambda: None, n_qubits=n)
    ev = script.execute(type="expval", obs=[J.PauliZ(wires=q, record=False) for q in range(n)])
    probs = script.execute(type="probs")
    assert_parity(ev.cpu().numpy(), np.ones(n), x64)
    expect = np.zeros(1 << n)
    expect[0] = 1
    assert_parity(probs.cpu().numpy(), expect, x64)


@pytest.mark.parametrize("x64", [False, True])
def test_single_qubit_rx_closed_form(x64):
    import jaqsi_amd as J

    print(f"edge single qubit x64={x64}", flush=True)
    J.enable_x64(x64)
    thetas = np.linspace(-3.0, 3.0, 7)
    for t in thetas:
        th = torch.tensor(t, dtype=torch.float64, device="cuda", requires_grad=True)
        ev = J.Script(lambda a: J.RX(a, wires=0), n_qubits=1).execute(
            type="expval", obs=[J.PauliZ(wires=0, record=False)], args=(th,))
        ev.sum().backward()
        assert_parity(ev.item(), np.cos(t), x64, scale=1.0)
        assert_parity(th.grad.item(), -np.sin(t), x64, scale=1.0)


def test_parameter_free_circuit():
    import jaqsi_amd as J

    print("edge parameter-free", flush=True)
    J.enable_x64(True)
    n = 4

    def circ():
        J.H(0)
        for q in range(n - 1):
            J.CX(wires=[q, q + 1])

    st = J.Script(circ, n_qubits=n).execute(type="state").cpu().numpy()
    ghz = np.zeros(1 << n, dtype=complex)
    ghz[0] = ghz[-1] = 1 / np.sqrt(2)
    assert_parity(st, ghz, True)
    ops = [sv.Op("H", (0,))] + [sv.Op("CX", (q, q + 1)) for q in range(n - 1)]
    assert_parity(st, sv.simulate(ops, n), True)


def test_batch_of_one_keeps_its_axis():
    import jaqsi_amd as J

    print("edge batch of one", flush=True)
    J.enable_x64(True)
    th = torch.tensor([[0.3, 0.7]], dtype=torch.float64, device="cuda")

    def circ(t):
        J.RX(t[0], wires=0)
        J.RY(t[1], wires=1)

    ev = J.Script(circ, n_qubits=2).execute(type="expval", obs=[J.PauliZ(wires=0, record=False)],
                                            args=(th,), in_axes=(0,))
    assert tuple(ev.shape)[0] == 1
    assert_parity(ev.cpu().numpy().reshape(-1), [np.cos(0.3)], True)
