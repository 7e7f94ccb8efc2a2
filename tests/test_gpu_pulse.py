"""GPU parity of the batched pulse propagators (FP64): propagators and
infidelities against the SciPy-expm Magnus-2 oracle, gradients against
central finite differences of the oracle."""

import math

import numpy as np
import pytest
import torch

from oracle import pulse
from parity import assert_parity

pytestmark = pytest.mark.gpu


def _run(params, omega, n_steps, T=50.0, J=pulse.J_C5, want_grad=True):
    from jaqsi_amd import _native as nat

    n = params.shape[0]
    P = torch.tensor(np.concatenate([params, omega], axis=1), dtype=torch.float64, device="cuda").contiguous()
    V = pulse.target_rx_half_pi()
    Vh = np.ascontiguousarray(np.stack([V.real, V.imag], axis=-1).reshape(-1))
    U = torch.empty((n, 4, 4), dtype=torch.complex128, device="cuda")
    inf = torch.empty(n, dtype=torch.float64, device="cuda")
    g = torch.empty((n, 4), dtype=torch.float64, device="cuda") if want_grad else None
    d = nat.PulseDesc(n, n_steps, T, J, nat.JQ_ENV_DRAG_GAUSSIAN, 0)
    nat.check(nat.lib().jq_pulse_propagate(d, nat.ptr(P), nat.ptr(Vh), nat.ptr(U), nat.ptr(inf), nat.ptr(g),
                                           torch.cuda.current_stream().cuda_stream))
    torch.cuda.synchronize()
    return U.cpu().numpy(), inf.cpu().numpy(), (g.cpu().numpy() if g is not None else None)


@pytest.mark.parametrize("n_steps", [7, 200, 333])
def test_propagators_and_infidelity(n_steps):
    print(f"pulse propagate n_steps={n_steps}", flush=True)
    params, omega = pulse.workload(5, seed=11)
    U, inf, _ = _run(params, omega, n_steps, want_grad=False)
    V = pulse.target_rx_half_pi()
    for s in range(5):
        ref = pulse.propagate(params[s], omega[s], pulse.J_C5, 50.0, n_steps)
        assert_parity(U[s], ref, True)
        assert_parity(inf[s], pulse.infidelity(ref, V), True, scale=1.0)


def test_gradient_against_finite_differences():
    print("pulse gradient", flush=True)
    params, omega = pulse.workload(3, seed=12)
    n_steps = 150
    _, _, g = _run(params, omega, n_steps)
    V = pulse.target_rx_half_pi()
    for s in range(3):
        fd = []
        for k in range(4):
            h = 1e-5 * max(1.0, abs(params[s, k]))
            pp, pm = params[s].copy(), params[s].copy()
            pp[k] += h
            pm[k] -= h
            fp = pulse.infidelity(pulse.propagate(pp, omega[s], pulse.J_C5, 50.0, n_steps), V)
            fm = pulse.infidelity(pulse.propagate(pm, omega[s], pulse.J_C5, 50.0, n_steps), V)
            fd.append((fp - fm) / (2 * h))
        np.testing.assert_allclose(g[s], fd, atol=2e-7, rtol=1e-6)


def test_zero_drive_is_diagonal_phase():
    print("pulse zero drive", flush=True)
    params = np.array([[0.0, 10.0, 0.3, 0.2]])
    omega = np.array([[30.0, 31.0]])
    U, _, _ = _run(params, omega, 64, want_grad=False)
    z = np.array([1, -1])
    e = np.array([15.0 * z0 + 15.5 * z1 + pulse.J_C5 * z0 * z1 for z0 in z for z1 in z])
    assert_parity(U[0], np.diag(np.exp(-1j * 50.0 * e)), True)


def test_autograd_api_matches_oracle_and_kernel_gradient():
    """jaqsi_amd.pulse: propagators / infidelities against the oracle, and the
    autograd gradient of sum(w * infidelity) equal to w times the kernel's
    own per-set gradient."""
    from jaqsi_amd import pulse as P

    print("pulse autograd api", flush=True)
    params, omega = pulse.workload(4, seed=13)
    n_steps = 120
    V = pulse.target_rx_half_pi()
    pt = torch.tensor(params, dtype=torch.float64, device="cuda", requires_grad=True)
    wt = torch.tensor(omega, dtype=torch.float64, device="cuda")
    U, inf, g = P.propagate(pt, wt, V, n_steps=n_steps, coupling=pulse.J_C5)
    for s in range(4):
        ref = pulse.propagate(params[s], omega[s], pulse.J_C5, 50.0, n_steps)
        assert_parity(U[s].cpu().numpy(), ref, True)
        assert_parity(inf[s].item(), pulse.infidelity(ref, V), True, scale=1.0)
    w = torch.tensor([1.0, -2.0, 0.5, 3.0], dtype=torch.float64, device="cuda")
    f = P.infidelity(pt, wt, torch.tensor(V), n_steps=n_steps, coupling=pulse.J_C5)
    np.testing.assert_allclose(f.detach().cpu().numpy(), inf.cpu().numpy(), atol=1e-15, rtol=0)
    (w * f).sum().backward()
    np.testing.assert_allclose(pt.grad.cpu().numpy(), (w[:, None] * g).cpu().numpy(), atol=1e-15, rtol=1e-13)
