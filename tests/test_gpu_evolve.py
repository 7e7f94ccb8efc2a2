"""GPU parity of the general propagator kernel (jq_evolve / jq_evolve_vjp)
and of pulse-level circuits (Gates.<G>(pulse=True), ParametrizedHamiltonian,
resolve_pending) against the float64 evolution oracle."""

import math

import numpy as np
import pytest
import torch

from oracle import evolve as ev
from parity import assert_parity

pytestmark = pytest.mark.gpu


def _random_terms(d, K, rng):
    Hs = []
    for _ in range(K):
        A = rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d))
        Hs.append((A + A.conj().T) / 2)
    return Hs


@pytest.mark.parametrize("d,K,M", [(2, 2, 1), (2, 3, 77), (4, 3, 130), (4, 1, 5)])
def test_product_of_exponentials_and_its_vjp(d, K, M):
    from jaqsi_amd.evolution import product_of_exponentials

    print(f"evolve d={d} K={K} M={M}", flush=True)
    rng = np.random.default_rng(100 * d + K)
    G = 3
    Hs = [_random_terms(d, K, rng) for _ in range(G)]
    scale = 2.0 if M == 1 else 0.3  # one factor: ||X|| > 1 exercises the squarings
    coef = rng.normal(size=(G, M, K)) * scale
    ct = torch.tensor(coef, device="cuda", requires_grad=True)
    Ht = torch.tensor(np.stack([np.stack(h) for h in Hs]), device="cuda")
    U = product_of_exponentials(ct, Ht)
    W = rng.normal(size=(G, d, d)) + 1j * rng.normal(size=(G, d, d))
    L = (U * torch.tensor(W, device="cuda").conj()).sum().real  # dL/dRe U + i dL/dIm U = W
    L.backward()
    for g in range(G):
        assert_parity(U[g].detach().cpu().numpy(), ev.product(coef[g], Hs[g]), True, scale=1.0)
        ref = ev.product_grad(coef[g], Hs[g], W[g])
        assert_parity(ct.grad[g].cpu().numpy(), ref, True, scale=max(1.0, np.abs(ref).max()))


def test_pulse_rx_state_matches_the_closed_form():
    """Gates.RX(w, pulse=True) under the default RWA gaussian envelope is the
    closed form exp(-i F X), F = (w / 2) * pulse area."""
    import jaqsi_amd as J

    J.enable_x64(True)
    pp = [0.44506904345408577, 1.9072729604580203, 3.5213966860043175]
    for w in (math.pi / 2, 2.1, -0.7):
        st = J.Script(lambda a: J.Gates.RX(a, wires=0, pulse=True), n_qubits=1).execute(type="state", args=(w,))
        ref = ev.rx_rwa(w, pp) @ np.array([1, 0], dtype=complex)
        assert_parity(st.cpu().numpy(), ref, True, scale=1.0)


def test_single_term_closed_form_erf():
    """The reference's tests/test_pulse_state.py:238-259 through the
    evolution API: the RWA gaussian RY at w = 60 (far beyond pi)."""
    from scipy.special import erf

    import jaqsi_amd as J
    from jaqsi_amd.pulses import PulseGates, PulseInformation

    J.enable_x64(True)
    PulseInformation.set_envelope("gaussian", rwa=True)
    A, sigma, T = [float(x) for x in PulseInformation.RY.params]
    w = 60.0
    H = PulseGates._coeff_RY_Y * J.Hermitian(PulseGates.Y, wires=0, record=False)
    U = H.evolve()([np.array([A, sigma, T, w])], T).matrix.cpu().numpy()
    g0 = math.exp(-(T**2) / (8 * sigma**2))
    area = sigma * math.sqrt(2 * math.pi) * erf(T / (2 * math.sqrt(2) * sigma)) - T * g0
    F = w * A * area / (2 * (1 - g0))
    assert_parity(U, math.cos(F) * np.eye(2) - 1j * math.sin(F) * ev.Y, True, scale=1.0)


@pytest.mark.parametrize("solver", ["magnus2", "magnus4", "dopri8"])
def test_two_term_drive_solvers(solver):
    """A DRAG RX (two non-commuting terms) with each solver against the
    oracle's CFM4 at a fine grid: Magnus-2 / CFM4 on their fixed grids
    against the oracle on the same grid, the adaptive solver to its
    tolerance."""
    import jaqsi_amd as J
    from jaqsi_amd.pulses import PulseGates, PulseInformation

    print(f"two-term drive {solver}", flush=True)
    J.enable_x64(True)
    with PulseInformation.preserve_state():
        PulseInformation.set_envelope("drag", rwa=True)
        p = np.array([0.9, 0.6, 1.8, 4.0, 1.7])
        H = (PulseGates._coeff_RX_X * J.Hermitian(ev.X, wires=0, record=False)
             + PulseGates._coeff_RX_Y * J.Hermitian(ev.Y, wires=0, record=False))
        U = H.evolve(solver=solver, magnus_steps=64)([p, p], 4.0).matrix.cpu().numpy()
    fns = list(ev.coeff_fns("drag")[:2])
    if solver == "magnus2":
        ref = ev.magnus2(fns, [p, p], [ev.X, ev.Y], 0.0, 4.0, 64)
    elif solver == "magnus4":
        ref = ev.cfm4(fns, [p, p], [ev.X, ev.Y], 0.0, 4.0, 64)
    else:
        ref = ev.cfm4(fns, [p, p], [ev.X, ev.Y], 0.0, 4.0, 2048)
    assert_parity(U, ref, True, scale=1.0) if solver != "dopri8" else np.testing.assert_allclose(U, ref, atol=1e-9)


def test_pulse_gate_set_against_the_oracle():
    """CZ, RZ, RY and the composite CX (H-CZ-H) at pulse level against the
    oracle's gates (RWA gaussian defaults), and identical gates solved once
    per group (resolve_pending's deduplication)."""
    import jaqsi_amd as J
    from jaqsi_amd import evolution
    from jaqsi_amd.tape import recording

    J.enable_x64(True)

    def circ():
        J.Gates.RY(0.7, wires=0, pulse=True)
        J.Gates.RZ(0.4, wires=1, pulse=True)
        J.Gates.CZ(wires=[0, 1], pulse=True)
        J.Gates.RY(0.7, wires=1, pulse=True)

    st = J.Script(circ, n_qubits=2).execute(type="state").cpu().numpy()
    ry = ev.ry_rwa(0.7, [0.449245182920846, 1.8896592598281816, 3.4886409030956145])
    psi = np.zeros(4, dtype=complex)
    psi[0] = 1
    psi = np.kron(ry, np.eye(2)) @ psi
    psi = np.kron(np.eye(2), ev.rz_pulse(0.4)) @ psi
    psi = ev.cz_pulse() @ psi
    psi = np.kron(np.eye(2), ry) @ psi
    assert_parity(st, psi, True, scale=1.0)
    with recording() as tape:
        circ()
    sizes = evolution.resolve_pending(tape)
    assert sorted(sizes) == [1, 1, 1]  # the two RY(0.7) pulses are one solve


def _adam(loss_fn, x0, lr, steps):
    """Plain Adam with optax defaults (b1 0.9, b2 0.999, eps 1e-8)."""
    x = x0.clone().detach()
    m = torch.zeros_like(x)
    v = torch.zeros_like(x)
    losses = []
    for k in range(1, steps + 1):
        xr = x.clone().requires_grad_(True)
        L = loss_fn(xr)
        (g,) = torch.autograd.grad(L, xr)
        m = 0.9 * m + 0.1 * g
        v = 0.999 * v + 0.001 * g * g
        mh = m / (1 - 0.9**k)
        vh = v / (1 - 0.999**k)
        x = x - lr * mh / (torch.sqrt(vh) + 1e-8)
        losses.append(float(L.detach()))
    losses.append(float(loss_fn(x)))
    return np.array(losses), x.cpu().numpy()


def _pulse_infidelity():
    import jaqsi_amd as J
    from jaqsi_amd.math import fidelity

    target = J.Script(lambda w: J.Gates.RX(w, wires=0), n_qubits=1).execute(type="state", args=(math.pi / 2,))
    script = J.Script(lambda w, pp: J.Gates.RX(w, wires=0, pulse=True, pulse_params=pp), n_qubits=1)

    def f(pp):
        return 1.0 - fidelity(target, script.execute(type="state", args=(math.pi / 2, pp)))

    return f


@pytest.mark.parametrize("envelope", ["gaussian", "sech", "drag_legacy"])
def test_pulse_training_matches_analytic_trajectory(envelope):
    """The reference's tests/test_training.py:131-157: Adam on the pulse-level
    infidelity (GPU circuit, state adjoint, propagator VJP) equals Adam on
    the analytic infidelity sin^2(theta (a - 1) / 2) of the RWA drive, with
    the pulse area a by 200-point Gauss-Legendre."""
    import jaqsi_amd as J
    from jaqsi_amd.pulses import PulseEnvelope, PulseInformation

    print(f"pulse training {envelope}", flush=True)
    J.enable_x64(True)
    nodes, weights = np.polynomial.legendre.leggauss(200)
    with PulseInformation.preserve_state():
        PulseInformation.set_envelope(envelope, rwa=True)
        fn = PulseEnvelope.get(envelope)["fn"]
        nd = torch.tensor(nodes, device="cuda")
        wt = torch.tensor(weights, device="cuda")

        def analytic(pp):
            T = pp[-1]
            t = 0.5 * T * (nd + 1.0)
            a = 0.5 * T * torch.sum(wt * fn(pp, t, T / 2))
            return torch.sin(math.pi / 2 * (a - 1.0) / 2) ** 2

        start = torch.tensor(np.asarray(PulseInformation.gate_by_name("RX").params) * 1.15, device="cuda")
        losses, params = _adam(_pulse_infidelity(), start, 0.01, 30)
        ref_losses, ref_params = _adam(analytic, start, 0.01, 30)
    np.testing.assert_allclose(losses, ref_losses, rtol=1e-10, atol=1e-14)
    np.testing.assert_allclose(params, ref_params, rtol=1e-10)


def test_legacy_pulse_training_golden_trajectory():
    """The reference's tests/test_training.py:160-185: the legacy DRAG doc
    example (Adam, lr 0.01, 30 steps) reproduces the published losses and
    parameters."""
    import jaqsi_amd as J
    from jaqsi_amd.pulses import PulseInformation

    print("legacy pulse training", flush=True)
    J.enable_x64(True)
    with PulseInformation.preserve_state():
        PulseInformation.set_envelope("drag_legacy", rwa=True)
        start = torch.tensor(np.asarray(PulseInformation.gate_by_name("RX").params) * 1.15, device="cuda")
        losses, params = _adam(_pulse_infidelity(), start, 0.01, 30)
    np.testing.assert_allclose(
        losses[[0, 9, 17, 28, 30]],
        [0.06279595635035018, 2.455984355853591e-05, 0.005852024299794256, 4.172439736160882e-06,
         0.0002575426158604177],
        rtol=1e-10,
    )
    np.testing.assert_allclose(params, [0.297624817828573, 0.5495358398394896, 6.03496082067464, 3.5238765222566895],
                               rtol=1e-10)
