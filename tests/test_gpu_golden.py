"""The GPU path against the committed golden fixtures (tests/golden)."""

import os

import numpy as np
import pytest
import torch

from circuits import record_hea
from parity import assert_parity

pytestmark = pytest.mark.gpu

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    return dict(np.load(os.path.join(HERE, name + ".npz")))


@pytest.mark.parametrize("x64", [False, True])
def test_c1_fixture(x64):
    import jaqsi_amd as J

    print(f"golden c1 x64={x64}", flush=True)
    J.enable_x64(x64)
    g = _load("c1_hea")
    n = 10
    th = torch.tensor(g["theta"], device="cuda", requires_grad=True)
    ev = J.Script(lambda t: record_hea(t, n), n_qubits=n).execute(
        type="expval", obs=[J.PauliZ(wires=q, record=False) for q in range(n)], args=(th,))
    ev.sum().backward()
    assert_parity(ev.detach().cpu().numpy(), g["expvals"], x64)
    assert_parity(th.grad.cpu().numpy().reshape(-1), g["grads"], x64)


@pytest.mark.parametrize("x64", [False, True])
def test_c3_head_fixture(x64):
    import jaqsi_amd as J
    from jaqsi_amd.ragged import RaggedBatch, record

    print(f"golden c3 x64={x64}", flush=True)
    J.enable_x64(x64)
    g = _load("c3_head")
    n = 14
    depths = [int(d) for d in g["depths"]]
    uniq = sorted(set(depths))
    members = {d: [b for b in range(len(depths)) if depths[b] == d] for d in uniq}
    tapes = [record(record_hea, np.zeros((d, n, 3)), n, np.zeros(n)) for d in uniq]
    rb = RaggedBatch(tapes, [len(members[d]) for d in uniq], n, [(q, 1.0, -1.0) for q in range(n)])
    angles = [torch.tensor(np.stack([np.concatenate([g["feats"][b], g[f"theta_{b}"].reshape(-1)]) for b in members[d]]),
                           device="cuda") for d in uniq]
    cts = [torch.ones((len(members[d]), n), dtype=torch.float64, device="cuda") for d in uniq]
    evs, grads = rb.adjoint(angles, cts)
    for gi, d in enumerate(uniq):
        for r, b in enumerate(members[d]):
            assert_parity(evs[gi][r].cpu().numpy(), g[f"expvals_{b}"], x64)
            assert_parity(grads[gi][r].cpu().numpy(), g[f"grads_{b}"], x64)


@pytest.mark.parametrize("x64", [False, True])
def test_dm_fixture(x64):
    import jaqsi_amd as J
    from test_gpu_dm import _record_noisy_hea

    print(f"golden dm x64={x64}", flush=True)
    J.enable_x64(x64)
    g = _load("dm_small")
    n = 3
    th = torch.tensor(g["theta"], device="cuda", requires_grad=True)
    ev = J.Script(lambda t: _record_noisy_hea(t, n, 0.01, 0.005), n_qubits=n).execute(
        type="expval", obs=[J.PauliZ(wires=0, record=False)], args=(th,))
    ev.sum().backward()
    assert_parity(ev.detach().cpu().numpy()[0], g["value"], x64)
    assert_parity(th.grad.cpu().numpy().reshape(-1), g["grads"], x64)


def test_c5_fixture():
    from test_gpu_pulse import _run

    print("golden c5", flush=True)
    g = _load("c5_small")
    U, inf, grads = _run(g["params"], g["omega"], int(g["n_steps"]))
    assert_parity(U, g["U"], True)
    assert_parity(inf, g["infidelity"], True, scale=1.0)
    np.testing.assert_allclose(grads, g["grads"], atol=2e-7, rtol=1e-6)
