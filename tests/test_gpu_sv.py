"""GPU parity of the statevector path against the complex128 oracle."""

import numpy as np
import pytest
import torch

from circuits import hea_theta, record_hea, record_zoo, zoo_ops, zoo_params
from oracle import sv
from parity import assert_parity

pytestmark = pytest.mark.gpu



def _script_expval(n, theta, x64, obs_wires=None):
    import jaqsi_amd as J

    J.enable_x64(x64)
    th = torch.tensor(theta, dtype=torch.float64, device="cuda", requires_grad=True)
    obs = [J.PauliZ(wires=q, record=False) for q in (obs_wires if obs_wires is not None else range(n))]
    script = J.Script(lambda t: record_hea(t, n), n_qubits=n)
    ev = script.execute(type="expval", obs=obs, args=(th,))
    ev.sum().backward()
    return ev.detach().cpu().numpy(), th.grad.detach().cpu().numpy().reshape(-1)


@pytest.mark.parametrize("x64", [False, True])
@pytest.mark.parametrize("n,layers", [(2, 3), (5, 2), (10, 5), (14, 2), (16, 2)])
def test_hea_expval_and_adjoint(n, layers, x64):
    print(f"hea n={n} layers={layers} x64={x64}", flush=True)
    theta = hea_theta(n, layers, 1000 + n)
    ev, grad = _script_expval(n, theta, x64)
    ops = sv.hea_ops(theta, n)
    vals, grads, _, _ = sv.adjoint(ops, n, sv.z_obs(n))
    assert_parity(ev, vals, x64)
    assert_parity(grad, sv.flat_grads(ops, grads), x64)


@pytest.mark.parametrize("x64", [False, True])
def test_deep_circuit_split_by_slot_budget(x64):
    """13 wires, 40 layers: one tile holds every wire, so only the LDS budget
    of the cotangent slots splits the passes."""
    n, layers = 13, 40
    print(f"deep hea n={n} layers={layers} x64={x64}", flush=True)
    theta = hea_theta(n, layers, 4000 + n)
    ev, grad = _script_expval(n, theta, x64)
    ops = sv.hea_ops(theta, n)
    vals, grads, _, _ = sv.adjoint(ops, n, sv.z_obs(n))
    assert_parity(ev, vals, x64)
    assert_parity(grad, sv.flat_grads(ops, grads), x64)


@pytest.mark.parametrize("x64", [False, True])
def test_gate_zoo_state_and_gradients(x64):
    import jaqsi_amd as J

    print(f"zoo x64={x64}", flush=True)
    J.enable_x64(x64)
    p = zoo_params()
    ops = zoo_ops(p)
    ref = sv.simulate(ops, 5)
    st = J.Script(lambda: record_zoo(p), n_qubits=5).execute(type="state").cpu().numpy()
    assert_parity(st, ref, x64)
    pt = torch.tensor(p, dtype=torch.float64, device="cuda", requires_grad=True)
    obs = [J.PauliZ(wires=q, record=False) for q in range(5)]
    ev = J.Script(record_zoo, n_qubits=5).execute(type="expval", obs=obs, args=(pt,))
    ev.sum().backward()
    vals, grads, _, _ = sv.adjoint(ops, 5, sv.z_obs(5))
    assert_parity(ev.detach().cpu().numpy(), vals, x64)
    flat = sv.flat_grads(ops, grads)
    # oracle parameters in tape order: p[0..11], then the last RZ's angle 0.5 * p[0]
    expect = flat[:12].copy()
    expect[0] += 0.5 * flat[12]
    assert_parity(pt.grad.cpu().numpy(), expect, x64)


def test_batched_in_axes_matches_loop():
    import jaqsi_amd as J

    n = 6
    rng = np.random.default_rng(7)
    x = rng.uniform(-np.pi, np.pi, (9, n))
    theta = rng.uniform(0, 2 * np.pi, (2, n, 3))
    obs = [J.PauliZ(wires=q, record=False) for q in range(n)]

    def circ(xi, th):
        record_hea(th, n, features=xi)

    s = J.Script(circ, n_qubits=n)
    xt = torch.tensor(x, device="cuda")
    tht = torch.tensor(theta, device="cuda", requires_grad=True)
    ev = s.execute(type="expval", obs=obs, args=(xt, tht), in_axes=(0, None))
    assert ev.shape == (9, n)
    ev.sum().backward()
    g = np.zeros(theta.size)
    for b in range(9):
        ops = sv.hea_ops(theta, n, features=x[b])
        vals, grads, _, _ = sv.adjoint(ops, n, sv.z_obs(n))
        assert_parity(ev[b].detach().cpu().numpy(), vals, False)
        g += sv.flat_grads(ops, grads)[n:]
    assert_parity(tht.grad.cpu().numpy().reshape(-1), g, False)


def test_probs_state_density_outputs():
    import jaqsi_amd as J

    n = 4
    theta = hea_theta(n, 2, 5)
    ref = sv.simulate(sv.hea_ops(theta, n), n)
    s = J.Script(lambda: record_hea(theta, n), n_qubits=n)
    assert_parity(s.execute(type="state").cpu().numpy(), ref, False)
    assert_parity(s.execute(type="probs").cpu().numpy(), np.abs(ref) ** 2, False)
    assert_parity(s.execute(type="density").cpu().numpy(), np.outer(ref, ref.conj()), False)


def test_initial_state_and_its_cotangent():
    import jaqsi_amd as J

    J.enable_x64(True)
    n = 3
    rng = np.random.default_rng(11)
    psi0 = rng.normal(size=2**n) + 1j * rng.normal(size=2**n)
    psi0 /= np.linalg.norm(psi0)
    theta = hea_theta(n, 1, 2)
    p0 = torch.tensor(psi0, device="cuda", requires_grad=True)
    obs = [J.PauliZ(wires=q, record=False) for q in range(n)]
    ev = J.Script(lambda: record_hea(theta, n), n_qubits=n).execute(type="expval", obs=obs, initial_state=p0)
    ev.sum().backward()
    ops = sv.hea_ops(theta, n)
    vals, _, _, lam0 = sv.adjoint(ops, n, sv.z_obs(n), psi0=psi0)
    assert_parity(ev.detach().cpu().numpy(), vals, True)
    # PyTorch's cotangent is the conjugate of JAX's
    assert_parity(p0.grad.cpu().numpy(), np.conj(lam0), True)


def test_lowered_entry_points_for_a_custom_vjp_bridge():
    import jaqsi_amd as J
    from jaqsi_amd import simulation as S
    from jaqsi_amd.tape import recording

    print("lowered entry points", flush=True)
    J.enable_x64(True)
    n = 4
    theta = hea_theta(n, 2, 77)
    with recording() as tape:
        record_hea(theta, n)
    obs = [J.PauliZ(wires=q, record=False) for q in range(n)]
    lowered, params = S.lower_for_jax(tape, n, False)
    a = torch.tensor([float(p) for p in params], dtype=torch.float64, device="cuda")
    ev = S.run_lowered(lowered, a, obs).cpu().numpy()
    ct = torch.arange(1, n + 1, dtype=torch.float64, device="cuda")
    g = S.adjoint_lowered(lowered, a, obs, ct).cpu().numpy()
    ops = sv.hea_ops(theta, n)
    vals, grads, _, _ = sv.adjoint(ops, n, sv.z_obs(n), ct=np.arange(1, n + 1, dtype=float))
    assert_parity(ev, vals, True)
    assert_parity(g, sv.flat_grads(ops, grads), True)


@pytest.mark.parametrize("x64", [False, True])
def test_general_observables_expval_and_gradient(x64):
    """Pauli words, products across wires and Hermitian matrices: measured
    through basis-rotated parity groups, differentiated by the adjoint."""
    import jaqsi_amd as J

    print(f"general observables x64={x64}", flush=True)
    J.enable_x64(x64)
    n = 4
    theta = hea_theta(n, 2, 4321)
    rng = np.random.default_rng(5)
    h2 = rng.normal(size=(2, 2)) + 1j * rng.normal(size=(2, 2))
    h2 = h2 + h2.conj().T
    h4 = rng.normal(size=(4, 4)) + 1j * rng.normal(size=(4, 4))
    h4 = h4 + h4.conj().T
    obs = [
        J.PauliX(0, record=False),
        J.PauliZ(0, record=False) @ J.PauliY(2, record=False),
        J.Hermitian(h2, wires=1, record=False),
        J.PauliY(1, record=False) @ J.PauliX(3, record=False),
        J.Hermitian(h4, wires=[2, 3], record=False),
    ]
    ref_obs = [(sv.PX, (0,)), (np.kron(sv.PZ, sv.PY), (0, 2)), (h2, (1,)), (np.kron(sv.PY, sv.PX), (1, 3)), (h4, (2, 3))]
    ct = np.array([1.0, -0.5, 0.7, 1.3, 0.25])
    th = torch.tensor(theta, device="cuda", requires_grad=True)
    ev = J.Script(lambda t: record_hea(t, n), n_qubits=n).execute(type="expval", obs=obs, args=(th,))
    (ev * torch.tensor(ct, device="cuda", dtype=ev.dtype)).sum().backward()
    ops = sv.hea_ops(theta, n)
    vals, grads, _, _ = sv.adjoint(ops, n, ref_obs, ct=ct)
    assert_parity(ev.detach().cpu().numpy(), vals, x64)
    assert_parity(th.grad.cpu().numpy().reshape(-1), sv.flat_grads(ops, grads), x64)
