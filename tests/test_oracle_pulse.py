"""The pulse oracle is pinned by closed forms, convergence order and an
independent per-block SU(2) restatement (CPU only)."""

import math

import numpy as np

from oracle import pulse, sv


def test_envelope_stable_form_matches_plain_form():
    t = np.linspace(0.0, 50.0, 101)
    for sigma in (5.0, 12.0, 20.0):
        g = np.exp(-((t - 25.0) ** 2) / (2 * sigma**2))
        g0 = math.exp(-(25.0**2) / (2 * sigma**2))
        np.testing.assert_allclose(pulse.lifted_gaussian(t, sigma, 25.0), (g - g0) / (1 - g0), atol=1e-13)
        E, Q = pulse.envelopes(t, 0.7, sigma, 0.3, 25.0)
        h = 1e-5
        Ep, _ = pulse.envelopes(t + h, 0.7, sigma, 0.3, 25.0)
        Em, _ = pulse.envelopes(t - h, 0.7, sigma, 0.3, 25.0)
        np.testing.assert_allclose(Q, -0.3 * (Ep - Em) / (2 * h), atol=1e-8)
    assert abs(pulse.lifted_gaussian(0.0, 10.0, 25.0)) < 1e-15
    assert abs(pulse.lifted_gaussian(25.0, 10.0, 25.0) - 1.0) < 1e-15


def test_zero_drive_closed_form():
    omega = (30.0, 31.5)
    U = pulse.propagate((0.0, 10.0, 0.5, 0.3), omega, pulse.J_C5, 50.0, 50)
    z = np.array([1, -1])
    e = np.array([omega[0] / 2 * z0 + omega[1] / 2 * z1 + pulse.J_C5 * z0 * z1 for z0 in z for z1 in z])
    np.testing.assert_allclose(U, np.diag(np.exp(-1j * 50.0 * e)), atol=1e-12)


def test_unitary_and_second_order_convergence():
    p, om = (0.5, 8.0, 0.4, 1.1), (29.0, 32.0)
    ref = pulse.propagate(p, om, pulse.J_C5, 50.0, 6400)
    errs = []
    for N in (800, 1600):
        U = pulse.propagate(p, om, pulse.J_C5, 50.0, N)
        np.testing.assert_allclose(U @ U.conj().T, np.eye(4), atol=1e-12)
        errs.append(np.abs(U - ref).max())
    assert 3.5 < errs[0] / errs[1] < 5.0  # Magnus-2 is second order


def test_block_su2_restatement():
    """H commutes with Z1, so each step is exp(-i h e_s) times an SU(2)
    rotation per value of qubit 1: an independent form of the same product."""
    p, om, N = (0.9, 15.0, -0.6, 4.0), (28.5, 33.0), 300
    U = pulse.propagate(p, om, pulse.J_C5, 50.0, N)
    h = 50.0 / N
    blocks = [np.eye(2, dtype=complex), np.eye(2, dtype=complex)]
    for n in range(N):
        t = (n + 0.5) * h
        E, Q = pulse.envelopes(t, p[0], p[1], p[2], 25.0)
        a = E * math.cos(p[3]) - Q * math.sin(p[3])
        b = E * math.sin(p[3]) + Q * math.cos(p[3])
        for s, z1 in enumerate((1, -1)):
            c = om[0] / 2 + pulse.J_C5 * z1
            r = math.sqrt(a * a + b * b + c * c)
            K = (a * sv.PX + b * sv.PY + c * sv.PZ) / r
            step = np.exp(-1j * h * om[1] / 2 * z1) * (math.cos(h * r) * np.eye(2) - 1j * math.sin(h * r) * K)
            blocks[s] = step @ blocks[s]
    full = np.zeros((4, 4), dtype=complex)
    for s in range(2):
        full[np.ix_([s, 2 + s], [s, 2 + s])] = blocks[s]
    np.testing.assert_allclose(U, full, atol=1e-12)


def test_infidelity_of_target_is_zero():
    V = pulse.target_rx_half_pi()
    assert abs(pulse.infidelity(V, V)) < 1e-15
    assert abs(pulse.infidelity(np.exp(0.3j) * V, V)) < 1e-15
