"""The density-matrix oracle is pinned by channel closed forms, the
statevector oracle on noiseless circuits, rho invariants and exact
parameter-shift / finite-difference gradients (CPU only)."""

import math

import numpy as np
import pytest

from oracle import dm, sv

RNG = np.random.default_rng(7)


@pytest.mark.parametrize(
    "name,p", [("BitFlip", 0.1), ("PhaseFlip", 0.3), ("DepolarizingChannel", 0.2), ("AmplitudeDamping", 0.4), ("PhaseDamping", 0.25)]
)
def test_kraus_sets_are_trace_preserving(name, p):
    ks = dm.kraus(name, p)
    assert np.allclose(sum(k.conj().T @ k for k in ks), np.eye(2))


def _bloch(rho):
    return np.array([np.real(np.trace(rho @ P)) for P in (sv.PX, sv.PY, sv.PZ)])


def test_channel_closed_forms():
    psi = sv.simulate([sv.Op("RY", (0,), (0.7,)), sv.Op("RZ", (0,), (0.4,))], 1)
    rho = np.outer(psi, psi.conj())
    p = 0.3
    out = dm.apply_channel(rho, dm.kraus("DepolarizingChannel", p), (0,), 1)
    np.testing.assert_allclose(_bloch(out), (1 - 4 * p / 3) * _bloch(rho), atol=1e-14)
    g = 0.35
    out = dm.apply_channel(rho, dm.kraus("AmplitudeDamping", g), (0,), 1)
    np.testing.assert_allclose(out[1, 1], (1 - g) * rho[1, 1], atol=1e-14)
    np.testing.assert_allclose(out[0, 1], math.sqrt(1 - g) * rho[0, 1], atol=1e-14)
    out = dm.apply_channel(rho, dm.kraus("BitFlip", 0.2), (0,), 1)
    b, b0 = _bloch(out), _bloch(rho)
    np.testing.assert_allclose(b, [b0[0], 0.6 * b0[1], 0.6 * b0[2]], atol=1e-14)


@pytest.mark.parametrize("pe,t1,t2,tg", [(0.0, 1e-4, 5e-5, 1e-6), (0.3, 1.0, 0.6, 0.7), (0.1, 2.0, 2.0, 1.5)])
def test_thermal_relaxation_decay(pe, t1, t2, tg):
    """Populations relax to (1 - pe, pe) with exp(-tg/t1) and coherences
    decay with exp(-tg/t2) (the reference's tests/test_jaqsi.py:689-711)."""
    ks = dm.kraus("ThermalRelaxationError", pe, t1, t2, tg)
    assert np.allclose(sum(k.conj().T @ k for k in ks), np.eye(2), atol=1e-14)
    psi = sv.simulate([sv.Op("RX", (0,), (1.0,)), sv.Op("RZ", (0,), (0.3,))], 1)
    rho = np.outer(psi, psi.conj())
    out = dm.apply_channel(rho, ks, (0,), 1)
    e1, e2 = math.exp(-tg / t1), math.exp(-tg / t2)
    p1 = e1 * rho[1, 1] + (1 - e1) * pe
    np.testing.assert_allclose(out, [[1 - p1, e2 * rho[0, 1]], [e2 * rho[1, 0], p1]], atol=1e-14)


@pytest.mark.parametrize("k", [1, 2])
def test_multi_qubit_depolarizing_is_a_mix_with_the_maximally_mixed_state(k):
    """Averaging rho over all 4^k Pauli strings gives I / 2^k, so the channel
    is rho -> (1 - p) rho + p I / 2^k."""
    p = 0.37
    n = k + 1
    ops = [sv.Op("RY", (q,), (0.3 + q,)) for q in range(n)] + [sv.Op("CX", (0, 1))]
    psi = sv.simulate(ops, n)
    rho = np.outer(psi, psi.conj())
    wires = tuple(range(1, k + 1))
    out = dm.apply_channel(rho, dm.kraus("MultiQubitDepolarizing", p, k), wires, n)
    # partial trace over the channel's wires, then the mixed state on them
    r = rho.reshape(2, 2**k, 2, 2**k)
    red = np.einsum("aibi->ab", r)
    want = (1 - p) * rho + p * np.kron(red, np.eye(2**k) / 2**k)
    np.testing.assert_allclose(out, want, atol=1e-14)


def test_noiseless_matches_statevector():
    n = 4
    theta = RNG.uniform(0, 2 * np.pi, (2, n, 3))
    ops = sv.hea_ops(theta, n)
    rho = dm.simulate(ops, n)
    psi = sv.simulate(ops, n)
    np.testing.assert_allclose(rho, np.outer(psi, psi.conj()), atol=1e-13)


def test_noisy_rho_invariants():
    n = 3
    rho = dm.simulate(dm.noisy_hea(RNG.uniform(0, 2 * np.pi, (2, n, 3)), n, 0.05, 0.03), n)
    assert np.isclose(np.trace(rho), 1)
    np.testing.assert_allclose(rho, rho.conj().T, atol=1e-14)
    assert np.linalg.eigvalsh(rho).min() > -1e-12
    assert np.real(np.trace(rho @ rho)) < 1


def test_shift_gradient_matches_finite_differences():
    n = 3
    theta = RNG.uniform(0, 2 * np.pi, (1, n, 3))
    obs = [(sv.PZ, (0,)), (sv.PX, (1,))]
    ct = np.array([1.0, 0.5])
    _, g = dm.value_and_shift_grad(dm.noisy_hea(theta, n, 0.05, 0.02), n, obs, ct)

    def f(t):
        rho = dm.simulate(dm.noisy_hea(t.reshape(theta.shape), n, 0.05, 0.02), n)
        return sum(c * dm.expval(rho, m, w, n) for c, (m, w) in zip(ct, obs))

    h = 1e-6
    x = theta.reshape(-1)
    fd = np.array([(f(x + h * e) - f(x - h * e)) / (2 * h) for e in np.eye(x.size)])
    np.testing.assert_allclose(g, fd, atol=1e-8)
